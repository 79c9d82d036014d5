// jst_fp64.hip — FP64 plain path for gfx950.
//
// Built with -ffp-contract=off: the reference (Rust) never fuses a*b+c, and this path is
// the parity path (per-residual relative error <= 1e-10 against the reference algorithm).
//
// Kernels
//   geom_*          FP64 metric precompute (jst-rs geometry.rs:76-216), once per case.
//   stage_kernel    one fused RK stage on a TI x TJ tile: ghost synthesis (boundary.rs),
//                   central flux (convection.rs), Green-Gauss gradient (gradient.rs),
//                   viscous/SA diffusion (viscous.rs), JST dissipation (dissipation.rs),
//                   SA source (source.rs), update + non-physical check (solver.rs:123-185).
//   post_end_kernel end of step: L2 density residual (solver.rs:190-200) from stage 5's
//                   per-tile partials in a fixed order, next step's global dt
//                   (timestep.rs:24-57), step control; dt_tile_kernel the dt-only refresh.
//   forces_kernel   wall force / CL / CD (no reference counterpart).
#include <hip/hip_runtime.h>

#include "jst_dev.h"
#include "jst_launch.h"

namespace {

__device__ __forceinline__ int wrapj(int j, int nj)
{
    j %= nj;
    return j < 0 ? j + nj : j;
}

__device__ __forceinline__ size_t cidx(int i, int j, int pitch) { return (size_t)i * pitch + j; }

struct Prim {
    double rho, vx, vy, p, t, e, nut;
};

// jst-rs solver.rs:151-185 (Solver::unpack), without the error branch.
__device__ __forceinline__ void unpack(const PhysConst &k, const double u[5], Prim &q)
{
    double rho = u[0];
    double inv = 1.0 / rho;
    q.rho = rho;
    q.vx = u[1] * inv;
    q.vy = u[2] * inv;
    q.e = u[3] * inv;
    q.p = (k.gamma - 1.0) * (u[3] - rho * (q.vx * q.vx + q.vy * q.vy) * 0.5);
    q.t = q.p / (k.r_gas * rho);
    q.nut = u[4] * inv;
}

// jst-rs state.rs:142-156 (Cells::pack)
__device__ __forceinline__ void pack(const Prim &q, double u[5])
{
    u[0] = q.rho;
    u[1] = q.rho * q.vx;
    u[2] = q.rho * q.vy;
    u[3] = q.rho * q.e;
    u[4] = q.rho * q.nut;
}

__device__ __forceinline__ void load_u(const double *U, size_t plane, size_t c, double u[5])
{
#pragma unroll
    for (int m = 0; m < 5; ++m) u[m] = U[m * plane + c];
}

// jst-rs boundary.rs:112-185 (riemann_face + extrapolate_nut) for far-field column jw,
// from the current interior state U.  Out: {rho, p, t, vx, vy, e, nut}.
__device__ void far_state(const CaseDev &C, const double *U, int jw, Prim &s)
{
    const PhysConst &k = C.k;
    const int ni = C.ni, pitch = C.pitch;
    const size_t plane = (size_t)ni * pitch;
    Prim q[3];
    for (int m = 0; m < 3; ++m) {
        double u[5];
        load_u(U, plane, cidx(ni - 1 - m, jw, pitch), u);
        unpack(k, u, q[m]);
    }
    const size_t f = cidx(ni, jw, pitch);
    const double nx = C.tnx[f], ny = C.tny[f], fmx = C.tmx[f], fmy = C.tmy[f];
    const double len = hypot(nx, ny);
    const double gamma = k.gamma;
    const double vn_inf = (k.u_inf * nx + k.v_inf * ny) / len;
    const double vt_inf = (k.u_inf * ny - k.v_inf * nx) / len;
    const double vn_in = (q[0].vx * nx + q[0].vy * ny) / len;
    const double vt_in = (q[0].vx * ny - q[0].vy * nx) / len;
    const double c_in = sqrt(k.r_gas * gamma * q[0].t);
    double c_face, vn_face, vt_face;
    bool inflow;
    if (!k.supersonic) {
        double r_in = vn_in + 2.0 * c_in / (gamma - 1.0);
        double r_inf = vn_inf - 2.0 * k.c_inf / (gamma - 1.0);
        double vn = 0.5 * (r_in + r_inf);
        double c = (gamma - 1.0) / 4.0 * (r_in - r_inf);
        inflow = vn <= 0.0;
        c_face = c;
        vn_face = vn;
        vt_face = inflow ? vt_inf : vt_in;
    } else if (vn_inf <= 0.0) {
        c_face = k.c_inf;
        vn_face = vn_inf;
        vt_face = vt_inf;
        inflow = true;
    } else {
        c_face = c_in;
        vn_face = vn_in;
        vt_face = vt_in;
        inflow = false;
    }
    const double t = c_face * c_face / (k.r_gas * gamma);
    const double rho = k.rho_inf * pow(t / k.t_inf, 1.0 / (gamma - 1.0));
    const double p = rho * k.r_gas * t;
    const double vx = (vt_face * ny + vn_face * nx) / len;
    const double vy = (-vt_face * nx + vn_face * ny) / len;
    const double e = p / (rho * (gamma - 1.0)) + 0.5 * (vx * vx + vy * vy);
    double nut;
    if (inflow) {
        nut = k.nut_inf;
    } else {
        const double FLOOR = 1e-10;
        double dist[3];
        for (int m = 0; m < 3; ++m) {
            size_t c = cidx(ni - 1 - m, jw, pitch);
            dist[m] = hypot(C.cx[c] - fmx, C.cy[c] - fmy);
        }
        double l1 = dist[0], l2 = dist[1], l3 = dist[2];
        double d12 = l1 - l2, d13 = l1 - l3, d23 = l2 - l3;
        if (fabs(d12) < 1e-14 || fabs(d13) < 1e-14 || fabs(d23) < 1e-14) {
            nut = fmax(q[0].nut, FLOOR);
        } else {
            double v = l2 * l3 / (d12 * d13) * q[0].nut + l1 * l3 / (-d12 * d23) * q[1].nut +
                       l1 * l2 / (d23 * -d13) * q[2].nut;
            nut = v > FLOOR ? v : FLOOR;
        }
    }
    s.rho = rho;
    s.p = p;
    s.t = t;
    s.vx = vx;
    s.vy = vy;
    s.e = e;
    s.nut = nut;
}

// jst-rs convection.rs:25-40 (euler_flux) on the face state fu = (U_L + U_R)/2.
__device__ __forceinline__ void euler_flux(const double ul[5], const double ur[5], double nx,
                                           double ny, double gamma, double out[5])
{
    double fu[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) fu[m] = 0.5 * (ul[m] + ur[m]);
    double rho = fu[0];
    double inv_rho = 1.0 / rho;
    double u = fu[1] * inv_rho;
    double v = fu[2] * inv_rho;
    double rho_e = fu[3];
    double p = (gamma - 1.0) * (rho_e - rho * (u * u + v * v) * 0.5);
    double vn = nx * u + ny * v;
    out[0] = rho * vn;
    out[1] = fu[1] * vn + p * nx;
    out[2] = fu[2] * vn + p * ny;
    out[3] = (rho_e + p) * vn;
    out[4] = fu[4] * vn;
}

// jst-rs dissipation.rs:29-31
__device__ __forceinline__ double pressure_sensor(double pm, double p0, double pp)
{
    return fabs((pm - 2.0 * p0 + pp) / (pm + 2.0 * p0 + pp));
}

// Viscous spectral radius of a cell over both directions (extension simulation.viscous_time_step;
// Blazek, Computational Fluid Dynamics, 3rd ed., eqs. 6.21-6.22): max(4/3, gamma)/rho
// (mu/Pr + mu_t/Prt) (|A|^2 + |C|^2)/V with (a, b), (c, d) the mean face normals of the
// convective radius and mu_t >= 0.  The local dt becomes CFL V / (conv + acou + 4 visc); the
// JST spectral radius keeps the convective part only.
__device__ __forceinline__ double visc_radius(const PhysConst &k, double rho, double t, double rho_nu,
                                              double a, double b, double c, double d, double vol)
{
    const double r = t * k.inv_t_ref;
    const double mu = r * sqrt(r) * k.suth_num / (t + k.t_suth); // config.rs:210-213
    const double chi = rho_nu / mu, c3 = chi * chi * chi;
    double mut = rho_nu * (c3 / (c3 + k.cv1_cubed));
    if (!(mut > 0.0) || (k.sa_neg && rho_nu < 0.0)) mut = 0.0;
    const double g = k.gamma > 4.0 / 3.0 ? k.gamma : 4.0 / 3.0;
    return g / rho * (mu * k.inv_pr + mut * k.inv_prt) * (a * a + b * b + c * c + d * d) / vol;
}

// jst-rs viscous.rs:37-71 (diffusion_tensor).  d[0..3] = D_x[1..4], d[4..7] = D_y[1..4]
// (component 0 of both columns is identically zero).
__device__ __forceinline__ void diffusion_tensor(const PhysConst &k, double t, double rho_nu,
                                                 double vx, double vy, const double g[8],
                                                 double d[8], double &chi_out, double &fv1_out)
{
    double r = t * k.inv_t_ref;
    double mu = r * sqrt(r) * k.suth_num / (t + k.t_suth); // config.rs:210-213
    double chi = rho_nu / mu;
    double c3 = chi * chi * chi;
    double f1 = c3 / (c3 + k.cv1_cubed);
    double mu_t = rho_nu * f1;
    double nu_coef = rho_nu;
    if (k.sa_neg && rho_nu < 0.0) {
        // extension spalart_allmaras.negative (Allmaras, Johnson & Spalart, ICCFD7-1902,
        // eqs. 20-21): no eddy viscosity, nu~ diffusion mu + rho nu~ fn with
        // fn = (cn1 + chi^3) / (cn1 - chi^3), cn1 = 16; also for the wall ghosts (nu~ =
        // -nu~ of the first cell), where the reference's fv1 has its pole at chi = +cv1
        mu_t = 0.0;
        nu_coef = rho_nu * ((16.0 + c3) / (16.0 - c3));
    }
    double mu_eff = mu + mu_t;
    double lam_eff = mu * k.inv_pr + mu_t * k.inv_prt;
    double txx = mu_eff * (4.0 / 3.0 * g[0] - 2.0 / 3.0 * g[3]);
    double tyy = mu_eff * (4.0 / 3.0 * g[3] - 2.0 / 3.0 * g[0]);
    double txy = mu_eff * (g[1] + g[2]);
    double qx = lam_eff * k.cp * g[4];
    double qy = lam_eff * k.cp * g[5];
    double nu_diff = k.sigma * (mu + nu_coef);
    d[0] = txx;
    d[1] = txy;
    d[2] = vx * txx + vy * txy + qx;
    d[3] = nu_diff * g[6];
    d[4] = txy;
    d[5] = tyy;
    d[6] = vx * txy + vy * tyy + qy;
    d[7] = nu_diff * g[7];
    chi_out = chi;
    fv1_out = f1;
}

// LLVM's constant powi expansion order (x^2, x^4, x^2*x^4), as Rust's f64::powi(6).
__device__ __forceinline__ double pow6(double x)
{
    double x2 = x * x;
    double x4 = x2 * x2;
    return x2 * x4;
}

// jst-rs source.rs:45-88 for one cell (already multiplied by the cell volume).
__device__ __forceinline__ double sa_source(const PhysConst &k, double rho, double rho_nu,
                                            double chi, double fv1, const double g[8],
                                            double inv_dist2, double vol)
{
    double cb1_over_kappa2 = k.cb1 * k.inv_kappa2;
    double cb2_sigma = k.cb2 * k.sigma;
    double nu_tilde = rho_nu / rho;
    double inv_kd2 = inv_dist2 * k.inv_kappa2;
    double ft2 = k.ct3 * exp(-k.ct4 * chi * chi);
    double fv2 = 1.0 - chi / (1.0 + chi * fv1);
    double s = k.fv3 * fabs(g[2] - g[1]);
    double s_tilde = fmax(s + nu_tilde * inv_kd2 * fv2, 1e-10);
    double prod = k.cb1 * (1.0 - ft2) * s_tilde * rho_nu;
    double r = fmin(nu_tilde / s_tilde * inv_kd2, k.rmax);
    double gg = r + k.cw2 * (pow6(r) - r);
    double fw = gg * cbrt(sqrt(k.one_plus_cw3_6 / (pow6(gg) + k.cw3_6)));
    double destr = (k.cw1 * fw - cb1_over_kappa2 * ft2) * rho * (nu_tilde * nu_tilde * inv_dist2);
    double diff = cb2_sigma * rho * (g[6] * g[6] + g[7] * g[7]);
    if (k.sa_neg && rho_nu < 0.0) {
        // extension spalart_allmaras.negative (ICCFD7-1902 eq. 19): production on the
        // vorticity cb1 (1 - ct3) S rho nu~, destruction -cw1 rho (nu~ / d)^2
        prod = k.cb1 * (1.0 - k.ct3) * s * rho_nu;
        destr = -k.cw1 * rho * (nu_tilde * nu_tilde * inv_dist2);
    }
    return (prod - destr + diff) * vol;
}

} // namespace

// ---------------------------------------------------------------------------------------
// geometry (jst-rs geometry.rs:76-216): nodes are (ring, theta) ring-major, theta wraps.
// ---------------------------------------------------------------------------------------

__global__ void geom_cells_kernel(const double *xy, int n_theta, int ni, int nj, int pitch,
                                  double *vol, double *cx, double *cy)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y;
    if (j >= nj || i >= ni) return;
    auto node = [&](int r, int t, double &x, double &y) {
        size_t q = (size_t)r * n_theta + (t % n_theta);
        x = xy[2 * q];
        y = xy[2 * q + 1];
    };
    double ax, ay, bx, by, qx, qy, dx, dy;
    node(i, j, ax, ay);
    node(i + 1, j, bx, by);
    node(i + 1, j + 1, qx, qy);
    node(i, j + 1, dx, dy);
    double d1x = qx - ax, d1y = qy - ay;
    double d2x = bx - dx, d2y = by - dy;
    double v = 0.5 * fabs(d1x * d2y - d1y * d2x);
    double cr0 = ax * by - bx * ay;
    double cr1 = bx * qy - qx * by;
    double cr2 = qx * dy - dx * qy;
    double cr3 = dx * ay - ax * dy;
    double sgn = 0.5 * (cr0 + cr1 + cr2 + cr3);
    double sx = (ax + bx) * cr0 + (bx + qx) * cr1 + (qx + dx) * cr2 + (dx + ax) * cr3;
    double sy = (ay + by) * cr0 + (by + qy) * cr1 + (qy + dy) * cr2 + (dy + ay) * cr3;
    size_t c = cidx(i, j, pitch);
    vol[c] = v;
    cx[c] = fabs(sgn) > 1e-30 ? sx / (6.0 * sgn) : 0.0;
    cy[c] = fabs(sgn) > 1e-30 ? sy / (6.0 * sgn) : 0.0;
}

__global__ void geom_faces_kernel(const double *xy, int n_theta, int ni, int nj, int pitch,
                                  double *tnx, double *tny, double *tmx, double *tmy,
                                  double *nnx, double *nny, double *nmx, double *nmy)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y;
    if (j >= nj || i > ni) return;
    auto node = [&](int r, int t, double &x, double &y) {
        size_t q = (size_t)r * n_theta + (t % n_theta);
        x = xy[2 * q];
        y = xy[2 * q + 1];
    };
    size_t c = cidx(i, j, pitch);
    double ax, ay, bx, by;
    node(i, j, ax, ay);
    node(i, j + 1, bx, by);
    double dx = bx - ax, dy = by - ay;
    tnx[c] = dy; // tau normal (dy, -dx) points radially outward
    tny[c] = -dx;
    tmx[c] = 0.5 * (ax + bx);
    tmy[c] = 0.5 * (ay + by);
    if (i < ni) {
        node(i + 1, j, bx, by);
        dx = bx - ax;
        dy = by - ay;
        nnx[c] = -dy; // n normal (-dy, dx) points toward +j
        nny[c] = dx;
        nmx[c] = 0.5 * (ax + bx);
        nmy[c] = 0.5 * (ay + by);
    }
}

// geometry.rs:192-216: window = min(max(15, nj/5), nj/2), then 1/V and 1/d^2 (95-101).
__global__ void geom_walldist_kernel(int ni, int nj, int pitch, const double *cx, const double *cy,
                                     const double *tmx, const double *tmy, const double *vol,
                                     double *wd, double *inv_vol, double *inv_wd2)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y;
    if (j >= nj || i >= ni) return;
    int w = nj / 5 > 15 ? nj / 5 : 15;
    if (w > nj / 2) w = nj / 2;
    size_t c = cidx(i, j, pitch);
    double px = cx[c], py = cy[c];
    double best = __builtin_huge_val();
    for (int dk = -w; dk <= w; ++dk) {
        int k = wrapj(j + dk, nj);
        double ex = px - tmx[k], ey = py - tmy[k];
        double d = sqrt(ex * ex + ey * ey);
        if (d < best) best = d;
    }
    wd[c] = best;
    inv_vol[c] = 1.0 / vol[c];
    inv_wd2[c] = 1.0 / (best * best);
}

// ---------------------------------------------------------------------------------------
// Fused RK stage.
//
// A block owns a TI x TJ tile of interior cells.  LDS holds the tile plus a 3-cell ring
// ("R3": U, p and the gradient primitives), the tile plus a 1-cell ring ("R1": viscous
// tensor D and the frozen V/dt), and the pressure sensors.  Ghost cells of the ring are
// synthesised on load: rows < 0 mirror rows -1-k (boundary.rs:30-51), rows >= NI take the
// far-field Riemann state of their column (boundary.rs:87-155), columns wrap modulo NJ
// (boundary.rs:54-72).  Every face flux is evaluated from identical inputs by both cells
// that share it, so the discrete scheme stays exactly conservative.
// ---------------------------------------------------------------------------------------

template <int TI, int TJ>
struct StageSmem {
    static constexpr int H3 = TI + 6, W3 = TJ + 6; // R3 rectangle
    static constexpr int H1 = TI + 2, W1 = TJ + 2; // R1 rectangle
    double u[5][H3][W3];
    // The primitives on R3 are read in phase 2 only, and the viscous tensor is written at the
    // end of phase 2 (after a barrier) and read in phase 3: they share storage, which brings
    // an 8 x 32 tile to 53 184 B, three workgroups per CU.
    union {
        struct {
            double p[H3][W3];
            double t[H3][W3];
            double vx[H3][W3];
            double vy[H3][W3];
            double nu[H3][W3];
        } q;
        double d[8][H1][W1]; // viscous tensor on R1
    };
    double si[TI + 4][TJ]; // sensor_i rows i0-2 .. i0+TI+1
    double sj[TI][TJ + 4]; // sensor_j cols j0-2 .. j0+TJ+1
    double ratio[H1][W1];  // V/dt rows clamp(i0-1 .. i0+TI), cols j0-1 .. j0+TJ (wrapped)
    double src[TI][TJ];
};

template <int TI, int TJ, int NT, bool TERMS>
__global__ __launch_bounds__(NT) void stage_kernel(const CaseDev *__restrict__ cases,
                                                   const TileDesc *__restrict__ tiles, int stage,
                                                   double alpha)
{
    using S = StageSmem<TI, TJ>;
    constexpr int H3 = S::H3, W3 = S::W3, H1 = S::H1, W1 = S::W1;
    static_assert((TI + 4) * TJ >= NT && TI * (TJ + 4) >= NT, "step-end tree reuses si and sj");
    __shared__ S sm;

    const TileDesc tl = tiles[blockIdx.x];
    const CaseDev &C = cases[tl.c];
    const PhysConst &k = C.k;
    CaseCtl *ctl = C.ctl;
    const int ni = C.ni, nj = C.nj, pitch = C.pitch;
    const size_t plane = (size_t)ni * pitch;
    const int i0 = tl.i0, j0 = tl.j0;
    const int tid = threadIdx.x;

    // control words as they stand when the launch starts (the step end that writes them is an
    // earlier launch on the stream; within this launch only err_key changes, atomically through
    // the global pointer, and no tile needs its new value): scalar loads through the constant
    // address space, issued together instead of as dependent vector loads
    const auto *cc = (const __attribute__((address_space(4))) CaseCtl *)ctl;
    const int status = cc->status;
    const long long step = cc->step;
    const unsigned long long ek = cc->err_key;
    const int rot = cc->rot;
    const long long seq = step * 5 + stage;
    if (!TERMS) {
        if (status != JST_RUN_ACTIVE) return;
        if (ek != JST_ERR_NONE && (long long)(ek >> 32) < seq) return;
    }
    // buffer rotation: stage 0 reads U^n = ubuf[rot]; stages alternate between the two others.
    const double *src;
    double *dst;
    const double *former = C.ubuf[rot];
    {
        const int b1 = rot == 2 ? 0 : rot + 1, b2 = b1 == 2 ? 0 : b1 + 1;
        if (TERMS) {
            src = C.ubuf[rot];
            dst = nullptr;
        } else if (stage == 0) {
            src = C.ubuf[rot];
            dst = C.ubuf[b1];
        } else if (stage & 1) {
            src = C.ubuf[b1];
            dst = C.ubuf[b2];
        } else {
            src = C.ubuf[b2];
            dst = C.ubuf[b1];
        }
    }
    const double dt = cc->dt;
    // frozen V/dt and local dt of this step: plane (step & 1); stage 5 writes the next step's
    // pair into the other plane while neighbouring tiles still read this one
    const size_t cur = (size_t)(step & 1) * plane, nxt = (size_t)((step + 1) & 1) * plane;

    // ---- phase 1: R3 load with ghost synthesis ----
    // A thread's interior and wall-mirror cells are loaded first, all together, so that their
    // round trips to memory overlap instead of following one another's unpack.
    constexpr int NL3 = (H3 * W3 + NT - 1) / NT;
    double ul[NL3][5];
#pragma unroll
    for (int s = 0; s < NL3; ++s) {
        const int idx = tid + s * NT;
        if (idx >= H3 * W3) continue;
        const int r = idx / W3, c = idx - r * W3;
        const int i = i0 - JST_H + r;
        if (i >= -JST_H && i < ni)
            load_u(src, plane, cidx(i >= 0 ? i : -1 - i, wrapj(j0 - JST_H + c, nj), pitch), ul[s]);
    }
    // frozen V/dt on R1: rows clamped to [0, NI-1] (dissipation.rs:59), columns wrapped; loaded
    // with the state, stored to LDS after it
    constexpr int NL1 = (H1 * W1 + NT - 1) / NT;
    double rl[NL1];
#pragma unroll
    for (int s = 0; s < NL1; ++s) {
        const int idx = tid + s * NT;
        if (idx >= H1 * W1) continue;
        const int r = idx / W1, c = idx - r * W1;
        int i = i0 - 1 + r;
        i = i < 0 ? 0 : (i > ni - 1 ? ni - 1 : i);
        rl[s] = C.ratio[cur + cidx(i, wrapj(j0 - 1 + c, nj), pitch)];
    }
#pragma unroll
    for (int s = 0; s < NL3; ++s) {
        const int idx = tid + s * NT;
        if (idx >= H3 * W3) continue;
        const int r = idx / W3, c = idx - r * W3;
        const int i = i0 - JST_H + r, jj = j0 - JST_H + c;
        const int jw = wrapj(jj, nj);
        double u[5];
        Prim q;
        if (i >= 0 && i < ni) {
#pragma unroll
            for (int m = 0; m < 5; ++m) u[m] = ul[s][m];
            unpack(k, u, q);
        } else if (i < 0 && i >= -JST_H) {
#pragma unroll
            for (int m = 0; m < 5; ++m) u[m] = ul[s][m];
            unpack(k, u, q);
            q.vx = -q.vx;
            q.vy = -q.vy;
            q.nut = -q.nut;
            pack(q, u);
        } else if (i >= ni && i < ni + JST_H) {
            far_state(C, src, jw, q);
            pack(q, u);
        } else { // outside the ghost layers: never read by a valid cell
            q.rho = 1.0; q.vx = 0.0; q.vy = 0.0; q.p = 1.0; q.t = 1.0; q.e = 1.0; q.nut = 0.0;
            pack(q, u);
        }
#pragma unroll
        for (int m = 0; m < 5; ++m) sm.u[m][r][c] = u[m];
        sm.q.p[r][c] = q.p;
        sm.q.t[r][c] = q.t;
        sm.q.vx[r][c] = q.vx;
        sm.q.vy[r][c] = q.vy;
        sm.q.nu[r][c] = q.nut;
    }
#pragma unroll
    for (int s = 0; s < NL1; ++s) {
        const int idx = tid + s * NT;
        if (idx >= H1 * W1) continue;
        const int r = idx / W1, c = idx - r * W1;
        sm.ratio[r][c] = rl[s];
    }
    __syncthreads();

    // ---- phase 2: sensors, gradients + viscous tensor on R1, SA source on the tile ----
    for (int idx = tid; idx < (TI + 4) * TJ; idx += NT) {
        const int r = idx / TJ, c = idx - r * TJ; // cell (i0-2+r, j0+c) -> R3 (r+1, c+3)
        sm.si[r][c] = pressure_sensor(sm.q.p[r][c + 3], sm.q.p[r + 1][c + 3], sm.q.p[r + 2][c + 3]);
    }
    for (int idx = tid; idx < TI * (TJ + 4); idx += NT) {
        const int r = idx / (TJ + 4), c = idx - r * (TJ + 4); // cell (i0+r, j0-2+c)
        sm.sj[r][c] = pressure_sensor(sm.q.p[r + 3][c], sm.q.p[r + 3][c + 1], sm.q.p[r + 3][c + 2]);
    }
    // the viscous tensor of each R1 cell stays in registers until every thread has read the
    // primitives it overwrites
    constexpr int NR1 = (H1 * W1 + NT - 1) / NT;
    double dreg[NR1][8];
    bool dset[NR1];
#pragma unroll
    for (int slot = 0; slot < NR1; ++slot) {
        const int idx = tid + slot * NT;
        dset[slot] = false;
        if (idx >= H1 * W1) continue;
        const int r = idx / W1, c = idx - r * W1;
        const int i = i0 - 1 + r, jj = j0 - 1 + c;
        const int R = r + 2, Q = c + 2; // R3 coordinates of the cell
        double g[8];
        if (i < -1 || i > ni) continue;
        if (i == ni) {
#pragma unroll
            for (int m = 0; m < 8; ++m) g[m] = 0.0; // far field (gradient.rs:97)
        } else {
            // gradient.rs:22-80 at (ie, jj); the wall ghost row copies row 0 (gradient.rs:93-96)
            const int ie = i < 0 ? 0 : i;
            const int Re = i < 0 ? R + 1 : R;
            const int jw = wrapj(jj, nj);
            const int jp = jw + 1 < nj ? jw + 1 : 0;
            const double upx = C.tnx[cidx(ie + 1, jw, pitch)], upy = C.tny[cidx(ie + 1, jw, pitch)];
            const double dnx = C.tnx[cidx(ie, jw, pitch)], dny = C.tny[cidx(ie, jw, pitch)];
            const double lfx = C.nnx[cidx(ie, jw, pitch)], lfy = C.nny[cidx(ie, jw, pitch)];
            const double rtx = C.nnx[cidx(ie, jp, pitch)], rty = C.nny[cidx(ie, jp, pitch)];
            const double inv_v = C.inv_vol[cidx(ie, jw, pitch)];
            const double(*fld[4])[W3] = {sm.q.vx, sm.q.vy, sm.q.t, sm.q.nu};
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                const double cc = fld[f][Re][Q];
                const double f_up = 0.5 * (cc + fld[f][Re + 1][Q]);
                const double f_dn = 0.5 * (cc + fld[f][Re - 1][Q]);
                const double f_lf = 0.5 * (cc + fld[f][Re][Q - 1]);
                const double f_rt = 0.5 * (cc + fld[f][Re][Q + 1]);
                g[2 * f] = (f_up * upx - f_dn * dnx + f_rt * rtx - f_lf * lfx) * inv_v;
                g[2 * f + 1] = (f_up * upy - f_dn * dny + f_rt * rty - f_lf * lfy) * inv_v;
            }
            if (i < 0) {
                g[4] = 0.0; // adiabatic wall: no temperature gradient in the ghost
                g[5] = 0.0;
            }
        }
        double chi, fv1;
        diffusion_tensor(k, sm.q.t[R][Q], sm.u[4][R][Q], sm.q.vx[R][Q], sm.q.vy[R][Q], g,
                         dreg[slot], chi, fv1);
        dset[slot] = true;
        const bool own = r >= 1 && r <= TI && c >= 1 && c <= TJ && i < ni && jj < nj;
        if (own) {
            const size_t cc = cidx(i, jj, pitch);
            sm.src[r - 1][c - 1] = sa_source(k, sm.u[0][R][Q], sm.u[4][R][Q], chi, fv1, g,
                                             C.inv_wd2[cc], C.vol[cc]);
            if (TERMS) {
#pragma unroll
                for (int m = 0; m < 8; ++m) C.dbg[(DBG_GRAD + m) * plane + cc] = g[m];
                C.dbg[DBG_CHI * plane + cc] = chi;
                C.dbg[DBG_FV1 * plane + cc] = fv1;
                C.dbg[DBG_S * plane + cc] = sm.src[r - 1][c - 1];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int slot = 0; slot < NR1; ++slot) {
        if (!dset[slot]) continue;
        const int idx = tid + slot * NT;
        const int r = idx / W1, c = idx - r * W1;
#pragma unroll
        for (int m = 0; m < 8; ++m) sm.d[m][r][c] = dreg[slot][m];
    }
    __syncthreads();

    // ---- phase 3: face fluxes, residual and update of the tile's own cells ----
    // Per cell, the per-face scalars come first (face-average normal velocity and pressure
    // of the convective flux, the JST coefficients); the five components then follow one at
    // a time with the reference expressions, which keeps few values live at once.
    const double gamma = k.gamma, half_cfl = 0.5 * k.cfl;
    double racc = 0.0, dmin = __builtin_huge_val(); // stage 5: this thread's step-end terms
    for (int o = tid; o < TI * TJ; o += NT) {
        const int ti = o / TJ, tj = o - ti * TJ;
        const int i = i0 + ti, j = j0 + tj;
        if (i >= ni || j >= nj) continue;
        const int R = ti + 3, Q = tj + 3;    // R3
        const int R1 = ti + 1, Q1 = tj + 1;  // R1
        const int jp = j + 1 < nj ? j + 1 : 0;
        // faces: 0 = tau face i (dn), 1 = tau face i+1 (up), 2 = n face j (lf), 3 = n face j+1 (rt)
        const size_t c_dn = cidx(i, j, pitch), c_up = cidx(i + 1, j, pitch), c_rt = cidx(i, jp, pitch);
        const double nx[4] = {C.tnx[c_dn], C.tnx[c_up], C.nnx[c_dn], C.nnx[c_rt]};
        const double ny[4] = {C.tny[c_dn], C.tny[c_up], C.nny[c_dn], C.nny[c_rt]};
        // the two cells of each face in R3: first (R - 1 | R | R | R), second (R | R + 1 | R | R)
        const int ar[4] = {R - 1, R, R, R}, ac[4] = {Q, Q, Q - 1, Q};
        const int br[4] = {R, R + 1, R, R}, bc[4] = {Q, Q, Q, Q + 1};

        // convection.rs:25-40 per face: fu = (U_a + U_b)/2, u, v, p, v_n
        double vn[4], pf[4];
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            const double rho = 0.5 * (sm.u[0][ar[f]][ac[f]] + sm.u[0][br[f]][bc[f]]);
            const double inv_rho = 1.0 / rho;
            const double u = 0.5 * (sm.u[1][ar[f]][ac[f]] + sm.u[1][br[f]][bc[f]]) * inv_rho;
            const double v = 0.5 * (sm.u[2][ar[f]][ac[f]] + sm.u[2][br[f]][bc[f]]) * inv_rho;
            const double rho_e = 0.5 * (sm.u[3][ar[f]][ac[f]] + sm.u[3][br[f]][bc[f]]);
            pf[f] = (gamma - 1.0) * (rho_e - rho * (u * u + v * v) * 0.5);
            vn[f] = nx[f] * u + ny[f] * v;
        }

        // dissipation.rs:44-185 per face: eps2 from the 4-cell sensor maximum, eps4, lambda
        double e2[4], e4[4], lam[4];
        {
            const int ri = ti + 2; // sensor_i row of cell i (rows start at i0-2)
            const int cj = tj + 2; // sensor_j column of cell j (columns start at j0-2)
            double m4[4];
            m4[0] = fmax(fmax(fmax(sm.si[ri - 2][tj], sm.si[ri - 1][tj]), sm.si[ri][tj]), sm.si[ri + 1][tj]);
            m4[1] = fmax(fmax(fmax(sm.si[ri - 1][tj], sm.si[ri][tj]), sm.si[ri + 1][tj]), sm.si[ri + 2][tj]);
            m4[2] = fmax(fmax(fmax(sm.sj[ti][cj - 2], sm.sj[ti][cj - 1]), sm.sj[ti][cj]), sm.sj[ti][cj + 1]);
            m4[3] = fmax(fmax(fmax(sm.sj[ti][cj - 1], sm.sj[ti][cj]), sm.sj[ti][cj + 1]), sm.sj[ti][cj + 2]);
            lam[0] = half_cfl * (sm.ratio[R1][Q1] + sm.ratio[R1 - 1][Q1]);
            lam[1] = half_cfl * (sm.ratio[R1 + 1][Q1] + sm.ratio[R1][Q1]);
            lam[2] = half_cfl * (sm.ratio[R1][Q1] + sm.ratio[R1][Q1 - 1]);
            lam[3] = half_cfl * (sm.ratio[R1][Q1 + 1] + sm.ratio[R1][Q1]);
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                e2[f] = k.k2 * m4[f];
                e4[f] = fmax(k.k4 - e2[f], 0.0);
            }
        }
        // 4-cell stencils um2 | um1 || u0 | up1 across each face (R3 coordinates)
        const int sr[4][4] = {{R - 2, R - 1, R, R + 1}, {R - 1, R, R + 1, R + 2}, {R, R, R, R}, {R, R, R, R}};
        const int sc[4][4] = {{Q, Q, Q, Q}, {Q, Q, Q, Q}, {Q - 2, Q - 1, Q, Q + 1}, {Q - 1, Q, Q + 1, Q + 2}};
        // viscous faces in R1: the neighbour across each face
        const int vr[4] = {R1 - 1, R1 + 1, R1, R1}, vc[4] = {Q1, Q1, Q1 - 1, Q1 + 1};

        const size_t cc = cidx(i, j, pitch);
        const double a_dt = k.lts ? alpha * C.localdt[cur + cc] : alpha * dt;
        const double scl = a_dt * C.inv_vol[cc];
        double un[5], rho_n = 0.0;
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            double fl[4], vf[4], df[4];
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                // convection.rs:34-39, component m of F(fu) . n
                const double fu = 0.5 * (sm.u[m][ar[f]][ac[f]] + sm.u[m][br[f]][bc[f]]);
                if (m == 0) fl[f] = fu * vn[f];
                else if (m == 1) fl[f] = fu * vn[f] + pf[f] * nx[f];
                else if (m == 2) fl[f] = fu * vn[f] + pf[f] * ny[f];
                else if (m == 3) fl[f] = (fu + pf[f]) * vn[f];
                else fl[f] = fu * vn[f];
                // viscous.rs:128-164: 0.5 (D_a + D_b) . n; component 0 of D is identically
                // zero, so its (0 + 0) * (0.5 n) terms sum to +0
                if (m == 0) {
                    vf[f] = 0.0 * (0.5 * nx[f]) + 0.0 * (0.5 * ny[f]);
                } else {
                    const double cx0 = sm.d[m - 1][R1][Q1], cy0 = sm.d[3 + m][R1][Q1];
                    const double cx1 = sm.d[m - 1][vr[f]][vc[f]], cy1 = sm.d[3 + m][vr[f]][vc[f]];
                    // the lower/left neighbour is the first operand of each sum
                    vf[f] = (f == 0 || f == 2) ? (cx1 + cx0) * (0.5 * nx[f]) + (cy1 + cy0) * (0.5 * ny[f])
                                               : (cx0 + cx1) * (0.5 * nx[f]) + (cy0 + cy1) * (0.5 * ny[f]);
                }
                // dissipation.rs:140-170: D = (eps2 d1 - eps4 d3) lambda
                const double um2 = sm.u[m][sr[f][0]][sc[f][0]], um1 = sm.u[m][sr[f][1]][sc[f][1]];
                const double u0 = sm.u[m][sr[f][2]][sc[f][2]], up1 = sm.u[m][sr[f][3]][sc[f][3]];
                const double d1 = u0 - um1;
                const double d3 = up1 - u0 * 3.0 + um1 * 3.0 - um2;
                df[f] = (d1 * e2[f] - d3 * e4[f]) * lam[f];
            }
            const double fc = fl[1] - fl[0] + fl[3] - fl[2];   // convection.rs:77
            const double fv = vf[1] - vf[0] + vf[3] - vf[2];   // viscous.rs:160
            const double fd = df[1] + df[3] - df[0] - df[2];   // dissipation.rs:182
            if (TERMS) {
                C.dbg[(DBG_FC + m) * plane + cc] = fc;
                C.dbg[(DBG_FV + m) * plane + cc] = fv;
                C.dbg[(DBG_FD + m) * plane + cc] = fd;
            } else {
                // solver.rs:123-145: U = U_former - (Fc - Fv - Fd - S) * (alpha dt / V)
                double r = fc - fv - fd;
                if (m == 4) r -= sm.src[ti][tj];
                const double fo = former[m * plane + cc];
                const double v = fo - r * scl;
                dst[m * plane + cc] = v;
                un[m] = v;
                if (m == 0) rho_n = fo;
            }
        }
        if (!TERMS) {
            // solver.rs:151-171: first row-major cell with rho or p NaN or <= 1e-15
            bool bad = isnan(un[0]) || un[0] <= 1e-15;
            if (!bad) {
                const double inv = 1.0 / un[0];
                const double vx = un[1] * inv, vy = un[2] * inv;
                const double p = (gamma - 1.0) * (un[3] - un[0] * (vx * vx + vy * vy) * 0.5);
                bad = isnan(p) || p <= 1e-15;
            }
            if (bad) {
                const unsigned long long key =
                    ((unsigned long long)seq << 32) | (unsigned long long)((long long)i * nj + j);
                atomicMin(&ctl->err_key, key);
            }
            if (stage == 4) {
                // step end, per cell: (rho^{n+1} - rho^n)^2 (solver.rs:190-200) and the next
                // step's local dt and V/dt from U^{n+1} (timestep.rs:24-47), with the face
                // normals this cell already holds
                const double d = un[0] - rho_n;
                racc += d * d;
                const double inv = 1.0 / un[0];
                const double vx = un[1] * inv, vy = un[2] * inv;
                const double p = (gamma - 1.0) * (un[3] - un[0] * (vx * vx + vy * vy) * 0.5);
                const double t = p / (k.r_gas * un[0]);
                const double cs = sqrt(k.r_gas * k.gamma * t);
                const double aa = 0.5 * (nx[0] + nx[1]), bb = 0.5 * (ny[0] + ny[1]);
                const double ccv = 0.5 * (nx[2] + nx[3]), ddv = 0.5 * (ny[2] + ny[3]);
                const double conv = fabs(vx * aa + vy * bb) + fabs(vx * ccv + vy * ddv);
                const double acou = cs * (sqrt(aa * aa + bb * bb) + sqrt(ccv * ccv + ddv * ddv));
                const double vol = C.vol[cc];
                const double dtc = k.cfl * vol / (conv + acou);
                const double ldt = k.visc_dt ? k.cfl * vol / (conv + acou + 4.0 * visc_radius(k, un[0], t,
                                                                un[0] * (un[4] * inv), aa, bb, ccv, ddv, vol))
                                             : dtc;
                C.localdt[nxt + cc] = ldt;
                C.ratio[nxt + cc] = vol / dtc;
                dmin = fmin(dmin, ldt);
            }
        }
    }
    if (!TERMS && stage == 4) {
        // the tile's residual sum and dt minimum, in the fixed tree of the step-end kernels
        __syncthreads(); // the sensor arrays are free now: reuse them for the tree
        double *red_s = &sm.si[0][0], *red_m = &sm.sj[0][0];
        red_s[tid] = racc;
        red_m[tid] = dmin;
        __syncthreads();
        for (int h = NT / 2; h > 0; h >>= 1) {
            if (tid < h) {
                red_s[tid] = red_s[tid] + red_s[tid + h];
                red_m[tid] = fmin(red_m[tid], red_m[tid + h]);
            }
            __syncthreads();
        }
        if (tid == 0) {
            C.partial[tl.t] = red_s[0];
            C.partial[C.ntiles + tl.t] = red_m[0];
        }
    }
}

// ---------------------------------------------------------------------------------------
// End of step.  Stage 5 leaves per tile the sum of (rho^{n+1}-rho^n)^2 (solver.rs:190-200) in
// partial[t] and the minimum of the next local dt in partial[ntiles + t], and writes the next
// step's local dt and V/dt (timestep.rs:24-47).  post_end_kernel (one block per case) then
// adds the tile sums in a fixed order (so the value does not depend on scheduling):
//   mode 1: residual sqrt(sum / NI / NJ), totaltime += dt, step += 1, rotation advance,
//           convergence / stop control;
//   both:   dt of the next step = minimum over the tiles (exact, order-free).
// dt_tile_kernel is the dt-only refresh of the current state (mode 0: at creation, after a
// state injection, before a residual-terms export): local dt and V/dt into the current plane
// and the tile minima.
// ---------------------------------------------------------------------------------------
template <int TI, int TJ, int NT>
__global__ __launch_bounds__(NT) void dt_tile_kernel(const CaseDev *__restrict__ cases,
                                                     const TileDesc *__restrict__ tiles)
{
    __shared__ double red_m[NT];
    const TileDesc tl = tiles[blockIdx.x];
    const CaseDev &C = cases[tl.c];
    const PhysConst &k = C.k;
    const CaseCtl *ctl = C.ctl;
    const int ni = C.ni, nj = C.nj, pitch = C.pitch;
    const size_t plane = (size_t)ni * pitch;
    const int tid = threadIdx.x;
    const int rot = *(volatile const int *)&ctl->rot;
    const long long step = *(volatile const long long *)&ctl->step;
    const size_t cur = (size_t)(step & 1) * plane;
    const double *u_now = C.ubuf[rot];
    double dmin = __builtin_huge_val();
    for (int o = tid; o < TI * TJ; o += NT) {
        const int ti = o / TJ, tj = o - ti * TJ;
        const int i = tl.i0 + ti, j = tl.j0 + tj;
        if (i >= ni || j >= nj) continue;
        const size_t c = cidx(i, j, pitch);
        double u[5];
        load_u(u_now, plane, c, u);
        Prim q;
        unpack(k, u, q);
        const double cs = sqrt(k.r_gas * k.gamma * q.t);
        const int jp = j + 1 < nj ? j + 1 : 0;
        const size_t cu = cidx(i + 1, j, pitch), cr = cidx(i, jp, pitch);
        const double aa = 0.5 * (C.tnx[c] + C.tnx[cu]);
        const double bb = 0.5 * (C.tny[c] + C.tny[cu]);
        const double ccv = 0.5 * (C.nnx[c] + C.nnx[cr]);
        const double ddv = 0.5 * (C.nny[c] + C.nny[cr]);
        const double conv = fabs(q.vx * aa + q.vy * bb) + fabs(q.vx * ccv + q.vy * ddv);
        const double acou = cs * (sqrt(aa * aa + bb * bb) + sqrt(ccv * ccv + ddv * ddv));
        const double dtc = k.cfl * C.vol[c] / (conv + acou);
        const double ldt = k.visc_dt ? k.cfl * C.vol[c] / (conv + acou + 4.0 * visc_radius(k, q.rho, q.t,
                                                            q.rho * q.nut, aa, bb, ccv, ddv, C.vol[c]))
                                     : dtc;
        C.localdt[cur + c] = ldt;
        C.ratio[cur + c] = C.vol[c] / dtc;
        dmin = fmin(dmin, ldt);
    }
    red_m[tid] = dmin;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (tid < s) red_m[tid] = fmin(red_m[tid], red_m[tid + s]);
        __syncthreads();
    }
    if (tid == 0) {
        C.partial[tl.t] = 0.0;
        C.partial[C.ntiles + tl.t] = red_m[0];
    }
}

template <int NT>
__global__ __launch_bounds__(NT) void post_end_kernel(const CaseDev *__restrict__ cases, int mode)
{
    __shared__ double red_s[NT];
    __shared__ double red_m[NT];
    const CaseDev &C = cases[blockIdx.x];
    const PhysConst &k = C.k;
    CaseCtl *ctl = C.ctl;
    const int tid = threadIdx.x;
    if (mode == 1) {
        // the stage launches before this one wrote them: scalar loads through the constant
        // address space, issued together
        const auto *cc = (const __attribute__((address_space(4))) CaseCtl *)ctl;
        const int status = cc->status;
        const unsigned long long ek = cc->err_key;
        if (status != JST_RUN_ACTIVE) return;
        if (ek != JST_ERR_NONE) {
            if (tid == 0) ctl->status = JST_RUN_NONPHYSICAL;
            return;
        }
    }
    // fixed-order reduction of the per-tile partials
    double s = 0.0, m = __builtin_huge_val();
    for (int t = tid; t < C.ntiles; t += NT) {
        s += C.partial[t];
        m = fmin(m, C.partial[C.ntiles + t]);
    }
    red_s[tid] = s;
    red_m[tid] = m;
    __syncthreads();
    for (int h = NT / 2; h > 0; h >>= 1) {
        if (tid < h) {
            red_s[tid] = red_s[tid] + red_s[tid + h];
            red_m[tid] = fmin(red_m[tid], red_m[tid + h]);
        }
        __syncthreads();
    }
    if (tid == 0) {
        // this thread is the only one to touch the control block in this launch: plain
        // loads (issued together) and stores instead of one round trip per volatile access
        CaseCtl &v = *ctl;
        if (mode == 1) {
            const int rot = v.rot;
            const double res = sqrt(red_s[0] / (double)C.ni / (double)C.nj);
            v.totaltime = v.totaltime + v.dt;
            const long long st = v.step + 1;
            v.step = st;
            v.rot = rot == 2 ? 0 : rot + 1;
            v.residual = res;
            C.res_hist[(st - 1) % C.res_cap] = res;
            if (st == 1) v.res_first = res;
            bool conv = res < k.targetres;
            if (k.residual_drop > 0.0 && v.res_first > 0.0)
                conv = conv || res < v.res_first * pow(10.0, -k.residual_drop);
            if (!v.timing) {
                if (conv) v.status = JST_RUN_CONVERGED;
                else if (st >= v.stop_step) v.status = JST_RUN_STOPPED;
            }
        }
        v.dt = red_m[0];
    }
}

// ---------------------------------------------------------------------------------------
// Wall forces: per wall face j, the momentum components of Fc_tau(0,j) and Fv_tau(0,j) from
// the current state into scratch[4][nj] (Fc_x | Fc_y | Fv_x | Fv_y; the wall distribution
// reads them back), then a serial FP64 sum of -(Fc - Fv) in j order by one thread.
// out = {Fx, Fy}.
// ---------------------------------------------------------------------------------------
__global__ void forces_kernel(const CaseDev *__restrict__ cases, int ci, double *scratch,
                              double *out)
{
    const CaseDev &C = cases[ci];
    const PhysConst &k = C.k;
    const int ni = C.ni, nj = C.nj, pitch = C.pitch;
    const size_t plane = (size_t)ni * pitch;
    const double *U = C.ubuf[C.ctl->rot];
    for (int j = threadIdx.x; j < nj; j += blockDim.x) {
        // prims at (0, j-1), (0, j), (0, j+1), (1, j) and the wall ghost (-1, j)
        Prim q0, ql, qr, q1, qg;
        double u0[5], ug[5], tmp[5];
        const int jm = j > 0 ? j - 1 : nj - 1, jp = j + 1 < nj ? j + 1 : 0;
        load_u(U, plane, cidx(0, j, pitch), u0);
        unpack(k, u0, q0);
        load_u(U, plane, cidx(0, jm, pitch), tmp);
        unpack(k, tmp, ql);
        load_u(U, plane, cidx(0, jp, pitch), tmp);
        unpack(k, tmp, qr);
        load_u(U, plane, cidx(1, j, pitch), tmp);
        unpack(k, tmp, q1);
        qg = q0;
        qg.vx = -q0.vx;
        qg.vy = -q0.vy;
        qg.nut = -q0.nut;
        pack(qg, ug);
        const size_t c0 = cidx(0, j, pitch);
        const double nx = C.tnx[c0], ny = C.tny[c0];
        double fc[5];
        euler_flux(ug, u0, nx, ny, k.gamma, fc);
        // gradient at (0, j)
        const double upx = C.tnx[cidx(1, j, pitch)], upy = C.tny[cidx(1, j, pitch)];
        const double lfx = C.nnx[c0], lfy = C.nny[c0];
        const double rtx = C.nnx[cidx(0, jp, pitch)], rty = C.nny[cidx(0, jp, pitch)];
        const double inv_v = C.inv_vol[c0];
        const double cen[4] = {q0.vx, q0.vy, q0.t, q0.nut};
        const double upv[4] = {q1.vx, q1.vy, q1.t, q1.nut};
        const double dnv[4] = {qg.vx, qg.vy, qg.t, qg.nut};
        const double lfv[4] = {ql.vx, ql.vy, ql.t, ql.nut};
        const double rtv[4] = {qr.vx, qr.vy, qr.t, qr.nut};
        double g[8];
        for (int f = 0; f < 4; ++f) {
            const double f_up = 0.5 * (cen[f] + upv[f]);
            const double f_dn = 0.5 * (cen[f] + dnv[f]);
            const double f_lf = 0.5 * (cen[f] + lfv[f]);
            const double f_rt = 0.5 * (cen[f] + rtv[f]);
            g[2 * f] = (f_up * upx - f_dn * nx + f_rt * rtx - f_lf * lfx) * inv_v;
            g[2 * f + 1] = (f_up * upy - f_dn * ny + f_rt * rty - f_lf * lfy) * inv_v;
        }
        double d0[8], dg[8], chi, fv1;
        diffusion_tensor(k, q0.t, u0[4], q0.vx, q0.vy, g, d0, chi, fv1);
        g[4] = 0.0;
        g[5] = 0.0;
        diffusion_tensor(k, qg.t, ug[4], qg.vx, qg.vy, g, dg, chi, fv1);
        const double fv1x = (dg[0] + d0[0]) * (0.5 * nx) + (dg[4] + d0[4]) * (0.5 * ny);
        const double fv2y = (dg[1] + d0[1]) * (0.5 * nx) + (dg[5] + d0[5]) * (0.5 * ny);
        scratch[j] = fc[1];
        scratch[nj + j] = fc[2];
        scratch[2 * nj + j] = fv1x;
        scratch[3 * nj + j] = fv2y;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double fx = 0.0, fy = 0.0;
        for (int j = 0; j < nj; ++j) {
            fx -= scratch[j] - scratch[2 * nj + j];
            fy -= scratch[nj + j] - scratch[3 * nj + j];
        }
        out[0] = fx;
        out[1] = fy;
    }
}

// ---------------------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------------------
namespace jst {

hipError_t launch_geometry(const double *xy, int n_theta, int ni, int nj, int pitch,
                           const GeomPtrs &g, hipStream_t s)
{
    dim3 blk(128), grd_c((nj + 127) / 128, ni), grd_f((nj + 127) / 128, ni + 1);
    hipLaunchKernelGGL(geom_cells_kernel, grd_c, blk, 0, s, xy, n_theta, ni, nj, pitch, g.vol,
                       g.cx, g.cy);
    hipLaunchKernelGGL(geom_faces_kernel, grd_f, blk, 0, s, xy, n_theta, ni, nj, pitch, g.tnx,
                       g.tny, g.tmx, g.tmy, g.nnx, g.nny, g.nmx, g.nmy);
    hipLaunchKernelGGL(geom_walldist_kernel, grd_c, blk, 0, s, ni, nj, pitch, g.cx, g.cy, g.tmx,
                       g.tmy, g.vol, g.wd, g.inv_vol, g.inv_wd2);
    return hipGetLastError();
}

hipError_t launch_stage_fp64(const CaseDev *cases, const TileDesc *tiles, int ntiles, int stage,
                             double alpha, bool terms, hipStream_t s)
{
    if (terms)
        hipLaunchKernelGGL((stage_kernel<FP64_TI, FP64_TJ, FP64_NT, true>), dim3(ntiles),
                           dim3(FP64_NT), 0, s, cases, tiles, stage, alpha);
    else
        hipLaunchKernelGGL((stage_kernel<FP64_TI, FP64_TJ, FP64_NT, false>), dim3(ntiles),
                           dim3(FP64_NT), 0, s, cases, tiles, stage, alpha);
    return hipGetLastError();
}

// mode 1: end of step (stage 5 has left the tile partials); mode 0: dt refresh of the current
// state.
hipError_t launch_post_fp64(const CaseDev *cases, const TileDesc *tiles, int ntiles, int mode,
                            hipStream_t s)
{
    if (mode == 0)
        hipLaunchKernelGGL((dt_tile_kernel<FP64_TI, FP64_TJ, FP64_NT>), dim3(ntiles), dim3(FP64_NT),
                           0, s, cases, tiles);
    hipLaunchKernelGGL((post_end_kernel<256>), dim3(1), dim3(256), 0, s, cases, mode);
    return hipGetLastError();
}

hipError_t launch_forces_fp64(const CaseDev *cases, int ci, double *scratch, double *out,
                              hipStream_t s)
{
    hipLaunchKernelGGL(forces_kernel, dim3(1), dim3(256), 0, s, cases, ci, scratch, out);
    return hipGetLastError();
}

} // namespace jst
