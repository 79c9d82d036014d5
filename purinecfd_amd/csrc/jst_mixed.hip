// jst_mixed.hip — FP32-mixed path for gfx950 (batched: one launch covers many cases).
//
// The stage stencils read an FP32 shadow of dU = U - U_inf and compute in FP32; the RK
// update U^(k) = U^n - alpha dt R / V is accumulated in FP64 from the FP64 U^n.  Face fluxes
// are evaluated in perturbation form (F(U) - F(U_inf)) . n: the freestream part cancels
// exactly over a closed cell (metric closure, jst-rs geometry.rs:17-18) and so never enters
// a single-precision sum.  Every face flux is computed once per tile; results depend only
// on the case's own data, not on the tile shape or on the other cases of the launch.
//
// Scheme (same as the FP64 plain path): jst-rs boundary.rs (ghosts), convection.rs,
// gradient.rs, viscous.rs, dissipation.rs, source.rs, solver.rs:123-200 (update, residual).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "jst_dev.h"
#include "jst_launch.h"

namespace {

__device__ __forceinline__ int wrapj(int j, int nj)
{
    j %= nj;
    return j < 0 ? j + nj : j;
}
// 32-bit cell offsets: the host rejects cases whose planes exceed 2^31 / 8 elements
__device__ __forceinline__ int cidx(int i, int j, int pitch) { return i * pitch + j; }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Primitive perturbations of a cell from its conserved perturbation d = U - U_inf:
// dp = p - p_inf, du = u - u_inf, dv = v - v_inf, dT = T - T_inf, and nu~ itself.
__device__ __forceinline__ void prim_from_du(const MixConst &k, const float d[5], float &dp,
                                             float &du, float &dv, float &dT, float &nut)
{
    const float rho = k.rho_inf + d[0];
    const float inv = frcp(rho);
    du = (d[1] - k.u_inf * d[0]) * inv;
    dv = (d[2] - k.v_inf * d[0]) * inv;
    const float u = k.u_inf + du, v = k.v_inf + dv;
    // rho |u|^2 - rho_inf |u_inf|^2, without forming either product
    const float dkin = d[0] * (u * u + v * v) + k.rho_inf * (du * (u + k.u_inf) + dv * (v + k.v_inf));
    dp = k.gm1 * (d[3] - 0.5f * dkin);
    dT = (dp - k.r_gas * k.t_inf * d[0]) * inv * (1.0f / k.r_gas);
    nut = (k.rnu_inf + d[4]) * inv;
}

// jst-rs boundary.rs:112-185 in FP64 from the FP64 reconstruction U_inf + dU of the three
// outermost cells.  Out: rho p t vx vy e nut.
__device__ __noinline__ void far_state64(const CaseMix &C, const float *S, int jw, double s[7])
{
    const PhysConst &k = C.k.k;
    const int ni = C.ni, pitch = C.pitch;
    const int plane = ni * pitch;
    double vx[3], vy[3], tt[3], nu[3];
    for (int m = 0; m < 3; ++m) {
        const int c = cidx(ni - 1 - m, jw, pitch);
        double u[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) u[q] = C.k.u64[q] + (double)S[q * plane + c];
        const double inv = 1.0 / u[0];
        vx[m] = u[1] * inv;
        vy[m] = u[2] * inv;
        const double p = (k.gamma - 1.0) * (u[3] - u[0] * (vx[m] * vx[m] + vy[m] * vy[m]) * 0.5);
        tt[m] = p / (k.r_gas * u[0]);
        nu[m] = u[4] * inv;
    }
    const int f = cidx(ni, jw, pitch);
    const double nx = C.tnx64[f], ny = C.tny64[f], fmx = C.tmx64[f], fmy = C.tmy64[f];
    const double len = hypot(nx, ny);
    const double gamma = k.gamma;
    const double vn_inf = (k.u_inf * nx + k.v_inf * ny) / len;
    const double vt_inf = (k.u_inf * ny - k.v_inf * nx) / len;
    const double vn_in = (vx[0] * nx + vy[0] * ny) / len;
    const double vt_in = (vx[0] * ny - vy[0] * nx) / len;
    const double c_in = sqrt(k.r_gas * gamma * tt[0]);
    double c_face, vn_face, vt_face;
    bool inflow;
    if (!k.supersonic) {
        const double r_in = vn_in + 2.0 * c_in / (gamma - 1.0);
        const double r_inf = vn_inf - 2.0 * k.c_inf / (gamma - 1.0);
        const double vn = 0.5 * (r_in + r_inf);
        c_face = (gamma - 1.0) / 4.0 * (r_in - r_inf);
        inflow = vn <= 0.0;
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
    const double ux = (vt_face * ny + vn_face * nx) / len;
    const double uy = (-vt_face * nx + vn_face * ny) / len;
    const double e = p / (rho * (gamma - 1.0)) + 0.5 * (ux * ux + uy * uy);
    double nut;
    if (inflow) {
        nut = k.nut_inf;
    } else {
        double dist[3];
        for (int m = 0; m < 3; ++m) {
            const int c = cidx(ni - 1 - m, jw, pitch);
            dist[m] = hypot(C.cx64[c] - fmx, C.cy64[c] - fmy);
        }
        const double l1 = dist[0], l2 = dist[1], l3 = dist[2];
        const double d12 = l1 - l2, d13 = l1 - l3, d23 = l2 - l3;
        if (fabs(d12) < 1e-14 || fabs(d13) < 1e-14 || fabs(d23) < 1e-14) {
            nut = fmax(nu[0], 1e-10);
        } else {
            const double v = l2 * l3 / (d12 * d13) * nu[0] + l1 * l3 / (-d12 * d23) * nu[1] +
                             l1 * l2 / (d23 * -d13) * nu[2];
            nut = v > 1e-10 ? v : 1e-10;
        }
    }
    s[0] = rho;
    s[1] = p;
    s[2] = t;
    s[3] = ux;
    s[4] = uy;
    s[5] = e;
    s[6] = nut;
}

// Perturbation Euler flux (F(U_f) - F(U_inf)) . n of the face state dU_f = (dU_L + dU_R)/2
// (jst-rs convection.rs:25-40 rewritten around the freestream).
__device__ __forceinline__ void euler_flux_pert(const MixConst &k, const float a[5],
                                                const float b[5], float nx, float ny, float g[5])
{
    float d[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) d[m] = 0.5f * (a[m] + b[m]);
    const float rho = k.rho_inf + d[0];
    const float inv = frcp(rho);
    const float du = (d[1] - k.u_inf * d[0]) * inv;
    const float dv = (d[2] - k.v_inf * d[0]) * inv;
    const float u = k.u_inf + du, v = k.v_inf + dv;
    const float dkin = d[0] * (u * u + v * v) + k.rho_inf * (du * (u + k.u_inf) + dv * (v + k.v_inf));
    const float dp = k.gm1 * (d[3] - 0.5f * dkin);
    const float dvn = du * nx + dv * ny;
    const float vn = (k.u_inf * nx + k.v_inf * ny) + dvn;
    g[0] = d[1] * nx + d[2] * ny;
    g[1] = d[1] * vn + k.m_inf * dvn + dp * nx;
    g[2] = d[2] * vn + k.n_inf * dvn + dp * ny;
    g[3] = (d[3] + dp) * vn + k.h_inf * dvn;
    g[4] = d[4] * vn + k.rnu_inf * dvn;
}

// jst-rs viscous.rs:37-71 in FP32; g = gradients of (du, dv, dT, nu~) == of (u, v, T, nu~).
__device__ __forceinline__ void diffusion_tensor32(const MixConst &k, float dT, float drn, float du,
                                                   float dv, const float g[8], float d[8],
                                                   float &chi_out, float &fv1_out)
{
    const float t = k.t_inf + dT;
    const float rnu = k.rnu_inf + drn;
    const float u = k.u_inf + du, v = k.v_inf + dv;
    const float r = t * k.inv_t_ref;
    const float mu = r * fsqrt(r) * k.suth_num * frcp(t + k.t_suth);
    const float chi = rnu * frcp(mu);
    const float c3 = chi * chi * chi;
    const float f1 = c3 * frcp(c3 + k.cv1_cubed);
    const float mu_t = rnu * f1;
    const float mu_eff = mu + mu_t;
    const float lam = mu * k.inv_pr + mu_t * k.inv_prt;
    const float txx = mu_eff * (4.0f / 3.0f * g[0] - 2.0f / 3.0f * g[3]);
    const float tyy = mu_eff * (4.0f / 3.0f * g[3] - 2.0f / 3.0f * g[0]);
    const float txy = mu_eff * (g[1] + g[2]);
    const float nu_diff = k.sigma * (mu + rnu);
    d[0] = txx;
    d[1] = txy;
    d[2] = u * txx + v * txy + lam * k.cp * g[4];
    d[3] = nu_diff * g[6];
    d[4] = txy;
    d[5] = tyy;
    d[6] = u * txy + v * tyy + lam * k.cp * g[5];
    d[7] = nu_diff * g[7];
    chi_out = chi;
    fv1_out = f1;
}

// jst-rs source.rs:45-88 in FP32 (x^(1/6) via exp2/log2).
__device__ __forceinline__ float sa_source32(const MixConst &k, float drho, float drn, float chi,
                                             float fv1, const float g[8], float inv_d2, float vol)
{
    const float rho = k.rho_inf + drho;
    const float rnu = k.rnu_inf + drn;
    const float nu = rnu * frcp(rho);
    const float inv_kd2 = inv_d2 * k.inv_kappa2;
    const float ft2 = k.ct3 * __expf(-k.ct4 * chi * chi);
    const float fv2 = 1.0f - chi * frcp(1.0f + chi * fv1);
    const float s = k.fv3 * fabsf(g[2] - g[1]);
    const float st = fmaxf(s + nu * inv_kd2 * fv2, 1e-10f);
    const float prod = k.cb1 * (1.0f - ft2) * st * rnu;
    const float r = fminf(nu * frcp(st) * inv_kd2, k.rmax);
    const float r2 = r * r, r6 = r2 * r2 * r2;
    const float gg = r + k.cw2 * (r6 - r);
    const float g2 = gg * gg, g6 = g2 * g2 * g2;
    const float q = k.one_plus_cw3_6 * frcp(g6 + k.cw3_6);
    const float fw = gg * __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(q) * (1.0f / 6.0f));
    const float destr = (k.cw1 * fw - k.cb1 * k.inv_kappa2 * ft2) * rho * (nu * nu * inv_d2);
    const float diff = k.cb2 * k.sigma * rho * (g[6] * g[6] + g[7] * g[7]);
    return (prod - destr + diff) * vol;
}

__device__ __forceinline__ float sensor32(float pinf4, float dm, float d0, float dp)
{
    // |p- - 2 p0 + p+| / (p- + 2 p0 + p+) with p = p_inf + dp (jst-rs dissipation.rs:29-31)
    return fabsf((dm - 2.0f * d0 + dp) * frcp(pinf4 + dm + 2.0f * d0 + dp));
}

template <int TI, int TJ>
struct MixSmem {
    static constexpr int H3 = TI + 6, W3 = TJ + 6, H1 = TI + 2, W1 = TJ + 2;
    union {
        struct {
            float u[5][H3][W3];
            float p[H3][W3]; // dp
            float du[H3][W3];
            float dv[H3][W3];
            float dT[H3][W3];
            float nu[H3][W3];
        } a;
        struct {
            float gt[5][TI + 1][TJ]; // tau faces i0 .. i0+TI
            float gn[5][TI][TJ + 1]; // n faces j0 .. j0+TJ
        } f;
    };
    float si[TI + 4][TJ];
    float sj[TI][TJ + 4];
    float ratio[H1][W1];
    float d[8][H1][W1];
    float src[TI][TJ];
    int col[W3]; // wrapped column index of R3 column c (j0 - 3 + c mod NJ)
};

} // namespace

// ---------------------------------------------------------------------------------------
// Fused FP32 stage, grid-strided over a tile list spanning any number of cases.  Stage 5
// also computes the step-end quantities of its own cells: the next local dt and V/dt (into
// the other half of their double buffer) and, per row and 32-column chunk, the residual
// partial sum and the dt minimum.  Partials are keyed by cell position, not by tile, so the
// step end does not depend on the tile shape or on the other cases of the batch.
// ---------------------------------------------------------------------------------------
template <int TI, int TJ, int NT, bool TERMS>
__global__ __launch_bounds__(NT) void mix_stage_kernel(const CaseMix *__restrict__ cases,
                                                       const TileDesc *__restrict__ tiles,
                                                       int ntiles, int stage, double alpha)
{
    using S = MixSmem<TI, TJ>;
    constexpr int H3 = S::H3, W3 = S::W3, H1 = S::H1, W1 = S::W1;
    constexpr int NFT = (TI + 1) * TJ, NFN = TI * (TJ + 1);
    constexpr int FPT = (NFT + NFN + NT - 1) / NT;
    static_assert(TJ % 32 == 0 && (TI * TJ) % NT == 0 && NT % 64 == 0, "tile shape");
    __shared__ S sm;
    const int tid = threadIdx.x;
    const bool last_stage = !TERMS && stage == 4;

    for (int tix = blockIdx.x; tix < ntiles; tix += gridDim.x) {
        const TileDesc tl = tiles[tix];
        const CaseMix &C = cases[tl.c];
        const MixConst &k = C.k;
        CaseCtl *ctl = C.ctl;
        const int ni = C.ni, nj = C.nj, pitch = C.pitch;
        const int plane = ni * pitch;
        const int i0 = tl.i0, j0 = tl.j0;
        const int status = *(volatile int *)&ctl->status;
        const long long step = *(volatile long long *)&ctl->step;
        const unsigned long long ek = *(volatile unsigned long long *)&ctl->err_key;
        const int rot = *(volatile int *)&ctl->rot;
        const long long seq = step * 5 + stage;
        if (!TERMS) {
            if (status != JST_RUN_ACTIVE) continue; // uniform per block
            if (ek != JST_ERR_NONE && (long long)(ek >> 32) < seq) continue; // failed earlier
        }
        const float *src = TERMS ? C.s32[rot] : C.s32[rot ^ (stage & 1)];
        float *sdst = C.s32[rot ^ ((stage & 1) ^ 1)];
        const double *former = C.u64[rot];
        double *u_next = C.u64[rot ^ 1];
        const float *ratio_now = C.ratio[rot];
        const double dt = ctl->dt;
        const double inv_cfl = 1.0 / k.k.cfl;

        for (int c = tid; c < W3; c += NT) {
            int jw = j0 - JST_H + c;
            while (jw < 0) jw += nj;
            while (jw >= nj) jw -= nj;
            sm.col[c] = jw;
        }
        __syncthreads();

        // ---- phase 1: R3 load + ghost synthesis (perturbation variables) ----
        for (int idx = tid; idx < H3 * W3; idx += NT) {
            const int r = idx / W3, c = idx - r * W3;
            const int i = i0 - JST_H + r; // >= -3
            float d[5], dp, du, dv, dT, nut;
            if (i < ni) {
                // interior row i, or row -1-i for the wall mirror (boundary.rs:30-51)
                const int q = cidx(i >= 0 ? i : -1 - i, sm.col[c], pitch);
#pragma unroll
                for (int m = 0; m < 5; ++m) d[m] = src[m * plane + q];
                prim_from_du(k, d, dp, du, dv, dT, nut);
                if (i < 0) { // scalars copied, u, v, nu~ negated
                    d[1] = -d[1] - 2.0f * k.m_inf;
                    d[2] = -d[2] - 2.0f * k.n_inf;
                    d[4] = -d[4] - 2.0f * k.rnu_inf;
                    du = -du - 2.0f * k.u_inf;
                    dv = -dv - 2.0f * k.v_inf;
                    nut = -nut;
                }
            } else if (i < ni + JST_H) {
                double s[7];
                far_state64(C, src, sm.col[c], s);
                const double *U = C.k.u64;
                d[0] = (float)(s[0] - U[0]);
                d[1] = (float)(s[0] * s[3] - U[1]);
                d[2] = (float)(s[0] * s[4] - U[2]);
                d[3] = (float)(s[0] * s[5] - U[3]);
                d[4] = (float)(s[0] * s[6] - U[4]);
                dp = (float)(s[1] - C.k.k.p_inf);
                du = (float)(s[3] - C.k.k.u_inf);
                dv = (float)(s[4] - C.k.k.v_inf);
                dT = (float)(s[2] - C.k.k.t_inf);
                nut = (float)s[6];
            } else {
#pragma unroll
                for (int m = 0; m < 5; ++m) d[m] = 0.0f;
                dp = du = dv = dT = 0.0f;
                nut = k.nut_inf;
            }
#pragma unroll
            for (int m = 0; m < 5; ++m) sm.a.u[m][r][c] = d[m];
            sm.a.p[r][c] = dp;
            sm.a.du[r][c] = du;
            sm.a.dv[r][c] = dv;
            sm.a.dT[r][c] = dT;
            sm.a.nu[r][c] = nut;
        }
        for (int idx = tid; idx < H1 * W1; idx += NT) {
            const int r = idx / W1, c = idx - r * W1;
            int i = i0 - 1 + r;
            i = i < 0 ? 0 : (i > ni - 1 ? ni - 1 : i);
            sm.ratio[r][c] = ratio_now[cidx(i, sm.col[c + 2], pitch)];
        }
        __syncthreads();

        // ---- phase 2: sensors, gradient + viscous tensor on R1, SA source ----
        const float pinf4 = 4.0f * k.p_inf;
        for (int idx = tid; idx < (TI + 4) * TJ; idx += NT) {
            const int r = idx / TJ, c = idx - r * TJ;
            sm.si[r][c] = sensor32(pinf4, sm.a.p[r][c + 3], sm.a.p[r + 1][c + 3], sm.a.p[r + 2][c + 3]);
        }
        for (int idx = tid; idx < TI * (TJ + 4); idx += NT) {
            const int r = idx / (TJ + 4), c = idx - r * (TJ + 4);
            sm.sj[r][c] = sensor32(pinf4, sm.a.p[r + 3][c], sm.a.p[r + 3][c + 1], sm.a.p[r + 3][c + 2]);
        }
        for (int idx = tid; idx < H1 * W1; idx += NT) {
            const int r = idx / W1, c = idx - r * W1;
            const int i = i0 - 1 + r, jj = j0 - 1 + c;
            const int R = r + 2, Q = c + 2;
            if (i > ni) continue;
            float g[8];
            if (i == ni) {
#pragma unroll
                for (int m = 0; m < 8; ++m) g[m] = 0.0f;
            } else {
                const int ie = i < 0 ? 0 : i;
                const int Re = i < 0 ? R + 1 : R;
                const int q0 = cidx(ie, sm.col[Q], pitch), qp = cidx(ie, sm.col[Q + 1], pitch);
                const float upx = C.tnx[q0 + pitch], upy = C.tny[q0 + pitch];
                const float dnx = C.tnx[q0], dny = C.tny[q0];
                const float lfx = C.nnx[q0], lfy = C.nny[q0];
                const float rtx = C.nnx[qp], rty = C.nny[qp];
                const float inv_v = C.inv_vol[q0];
#pragma unroll
                for (int f = 0; f < 4; ++f) {
                    const float(*fld)[W3] = f == 0 ? sm.a.du : (f == 1 ? sm.a.dv : (f == 2 ? sm.a.dT : sm.a.nu));
                    const float cc = fld[Re][Q];
                    const float f_up = 0.5f * (cc + fld[Re + 1][Q]);
                    const float f_dn = 0.5f * (cc + fld[Re - 1][Q]);
                    const float f_lf = 0.5f * (cc + fld[Re][Q - 1]);
                    const float f_rt = 0.5f * (cc + fld[Re][Q + 1]);
                    g[2 * f] = (f_up * upx - f_dn * dnx + f_rt * rtx - f_lf * lfx) * inv_v;
                    g[2 * f + 1] = (f_up * upy - f_dn * dny + f_rt * rty - f_lf * lfy) * inv_v;
                }
                if (i < 0) {
                    g[4] = 0.0f;
                    g[5] = 0.0f;
                }
            }
            float d[8], chi, fv1;
            diffusion_tensor32(k, sm.a.dT[R][Q], sm.a.u[4][R][Q], sm.a.du[R][Q], sm.a.dv[R][Q], g, d,
                               chi, fv1);
#pragma unroll
            for (int m = 0; m < 8; ++m) sm.d[m][r][c] = d[m];
            if (r >= 1 && r <= TI && c >= 1 && c <= TJ && i < ni && jj < nj) {
                const int q = cidx(i, jj, pitch);
                sm.src[r - 1][c - 1] = sa_source32(k, sm.a.u[0][R][Q], sm.a.u[4][R][Q], chi, fv1, g,
                                                   C.inv_wd2[q], C.vol[q]);
            }
        }
        __syncthreads();

        // ---- phase 3: every face of the tile once: G = dFc - Fv - Fd ----
        float G[FPT][5];
        int gsel[FPT];
#pragma unroll
        for (int q = 0; q < FPT; ++q) {
            const int idx = tid + q * NT;
            gsel[q] = -1;
            if (idx >= NFT + NFN) continue;
            float lam, e2, e4, nx, ny, dA[8], dB[8];
            float um2[5], um1[5], u0[5], up1[5];
            float s0, s1, s2, s3;
            if (idx < NFT) { // tau face (f, j): cells f-1 (below) and f
                const int rf = idx / TJ, c = idx - rf * TJ;
                const int f = i0 + rf, j = j0 + c;
                if (f > ni || j >= nj) continue;
                const int R = rf + 3, Q = c + 3;
                nx = C.tnx[cidx(f, j, pitch)];
                ny = C.tny[cidx(f, j, pitch)];
#pragma unroll
                for (int m = 0; m < 5; ++m) {
                    um2[m] = sm.a.u[m][R - 2][Q];
                    um1[m] = sm.a.u[m][R - 1][Q];
                    u0[m] = sm.a.u[m][R][Q];
                    up1[m] = sm.a.u[m][R + 1][Q];
                }
                s0 = sm.si[rf][c];
                s1 = sm.si[rf + 1][c];
                s2 = sm.si[rf + 2][c];
                s3 = sm.si[rf + 3][c];
                lam = sm.ratio[rf + 1][c + 1] + sm.ratio[rf][c + 1];
#pragma unroll
                for (int m = 0; m < 8; ++m) {
                    dA[m] = sm.d[m][rf][c + 1];
                    dB[m] = sm.d[m][rf + 1][c + 1];
                }
            } else { // n face (i, g): cells g-1 (left) and g
                const int id2 = idx - NFT;
                const int r = id2 / (TJ + 1), cg = id2 - r * (TJ + 1);
                const int i = i0 + r, gcol = j0 + cg;
                if (i >= ni || gcol > nj) continue;
                const int R = r + 3, Q = cg + 3;
                const int gw = sm.col[Q];
                nx = C.nnx[cidx(i, gw, pitch)];
                ny = C.nny[cidx(i, gw, pitch)];
#pragma unroll
                for (int m = 0; m < 5; ++m) {
                    um2[m] = sm.a.u[m][R][Q - 2];
                    um1[m] = sm.a.u[m][R][Q - 1];
                    u0[m] = sm.a.u[m][R][Q];
                    up1[m] = sm.a.u[m][R][Q + 1];
                }
                s0 = sm.sj[r][cg];
                s1 = sm.sj[r][cg + 1];
                s2 = sm.sj[r][cg + 2];
                s3 = sm.sj[r][cg + 3];
                lam = sm.ratio[r + 1][cg + 1] + sm.ratio[r + 1][cg];
#pragma unroll
                for (int m = 0; m < 8; ++m) {
                    dA[m] = sm.d[m][r + 1][cg];
                    dB[m] = sm.d[m][r + 1][cg + 1];
                }
            }
            gsel[q] = idx;
            lam *= k.cfl_half;
            const float mx = fmaxf(fmaxf(fmaxf(s0, s1), s2), s3);
            e2 = k.k2 * mx;
            e4 = fmaxf(k.k4 - e2, 0.0f);
            float fc[5];
            euler_flux_pert(k, um1, u0, nx, ny, fc);
            const float hx = 0.5f * nx, hy = 0.5f * ny;
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                const float d1 = u0[m] - um1[m];
                const float d3 = up1[m] - 3.0f * u0[m] + 3.0f * um1[m] - um2[m];
                const float fd = (d1 * e2 - d3 * e4) * lam;
                const float fv = m == 0 ? 0.0f : (dA[m - 1] + dB[m - 1]) * hx + (dA[m + 3] + dB[m + 3]) * hy;
                G[q][m] = fc[m] - fv - fd;
            }
        }
        __syncthreads(); // R3 arrays are dead from here on: the face arrays alias them
#pragma unroll
        for (int q = 0; q < FPT; ++q) {
            const int idx = gsel[q];
            if (idx < 0) continue;
            if (idx < NFT) {
                const int rf = idx / TJ, c = idx - rf * TJ;
#pragma unroll
                for (int m = 0; m < 5; ++m) sm.f.gt[m][rf][c] = G[q][m];
            } else {
                const int id2 = idx - NFT;
                const int r = id2 / (TJ + 1), cg = id2 - r * (TJ + 1);
#pragma unroll
                for (int m = 0; m < 5; ++m) sm.f.gn[m][r][cg] = G[q][m];
            }
        }
        __syncthreads();

        // ---- phase 4: residual and FP64 update of the tile's own cells ----
        for (int o = tid; o < TI * TJ; o += NT) { // uniform trip count (TI*TJ % NT == 0)
            const int ti = o / TJ, tj = o - ti * TJ;
            const int i = i0 + ti, j = j0 + tj;
            const bool valid = i < ni && j < nj;
            double d2 = 0.0, dmin = __builtin_huge_val();
            if (valid) {
                float R5[5];
#pragma unroll
                for (int m = 0; m < 5; ++m)
                    R5[m] = sm.f.gt[m][ti + 1][tj] - sm.f.gt[m][ti][tj] + sm.f.gn[m][ti][tj + 1] - sm.f.gn[m][ti][tj];
                const int q = cidx(i, j, pitch);
                if (TERMS) { // export R = Fc - Fv - Fd per component and S separately
#pragma unroll
                    for (int m = 0; m < 5; ++m) C.dbg[m * plane + q] = R5[m];
                    C.dbg[5 * plane + q] = sm.src[ti][tj];
                } else {
                    R5[4] -= sm.src[ti][tj];
                    const double a_dt = k.lts ? alpha * (double)C.localdt[rot][q] : alpha * dt;
                    const double sc = a_dt * (double)C.inv_vol[q];
                    double un[5];
#pragma unroll
                    for (int m = 0; m < 5; ++m) {
                        un[m] = former[m * plane + q] - (double)R5[m] * sc;
                        sdst[m * plane + q] = (float)(un[m] - k.u64[m]);
                    }
                    // NonPhysical test (state.rs:100-117) without a division: rho > 0 there,
                    // so p <= 1e-15 <=> p * rho <= 1e-15 * rho
                    const double prho = (k.k.gamma - 1.0) * (un[3] * un[0] - 0.5 * (un[1] * un[1] + un[2] * un[2]));
                    if (isnan(un[0]) || un[0] <= 1e-15 || isnan(prho) || prho <= 1e-15 * un[0])
                        atomicMin(&ctl->err_key, ((unsigned long long)seq << 32) |
                                                     (unsigned long long)((long long)i * nj + j));
                    if (last_stage) {
                        const double inv = 1.0 / un[0];
                        const double vx = un[1] * inv, vy = un[2] * inv;
                        const double p = prho * inv;
#pragma unroll
                        for (int m = 0; m < 5; ++m) u_next[m * plane + q] = un[m];
                        const double dr = un[0] - former[q];
                        d2 = dr * dr;
                        // next step's local dt and V/dt (jst-rs timestep.rs:24-47)
                        const int jp = j + 1 < nj ? j + 1 : 0;
                        const int qu = cidx(i + 1, j, pitch), qr = cidx(i, jp, pitch);
                        const double aa = 0.5 * ((double)C.tnx[q] + (double)C.tnx[qu]);
                        const double bb = 0.5 * ((double)C.tny[q] + (double)C.tny[qu]);
                        const double cc = 0.5 * ((double)C.nnx[q] + (double)C.nnx[qr]);
                        const double dd = 0.5 * ((double)C.nny[q] + (double)C.nny[qr]);
                        const double cs = sqrt(k.k.gamma * fmax(p, 0.0) * inv);
                        const double conv = fabs(vx * aa + vy * bb) + fabs(vx * cc + vy * dd);
                        const double acou = cs * (sqrt(aa * aa + bb * bb) + sqrt(cc * cc + dd * dd));
                        const double vol = (double)C.vol[q];
                        const double ldt = k.k.cfl * vol / (conv + acou);
                        C.localdt[rot ^ 1][q] = (float)ldt;
                        C.ratio[rot ^ 1][q] = (float)((conv + acou) * inv_cfl); // V / dt_local
                        dmin = ldt;
                    }
                }
            }
            if (last_stage) {
                // this row's 32-column chunk: fixed 32-lane butterflies (sum, min)
                for (int off = 16; off > 0; off >>= 1) {
                    d2 += __shfl_xor(d2, off, 32);
                    dmin = fmin(dmin, __shfl_xor(dmin, off, 32));
                }
                if ((tid & 31) == 0 && i < ni && j < nj) {
                    const int ch = i * C.nchunk + (j >> 5);
                    C.rpart[ch] = d2;
                    C.dpart[ch] = dmin;
                }
            }
        }
        __syncthreads(); // LDS is reused by the next tile of this block
    }
}

// ---------------------------------------------------------------------------------------
// dt refresh (case creation, residual-terms export): local dt and V/dt of the current state
// into buffer [rot], global minimum into ctl->dt.  Work items are fixed 1024-cell chunks.
// ---------------------------------------------------------------------------------------
// ---------------------------------------------------------------------------------------
// End of step, one block per case: residual sqrt(sum (rho^{n+1} - rho^n)^2 / NI / NJ)
// (jst-rs solver.rs:190-200) and the next global dt from the per-(row, 32-column) partials in
// a fixed order, totaltime += dt, step += 1, buffer rotation, and the stop decision
// (convergence / relative drop / step limit, or the non-physical record).
// ---------------------------------------------------------------------------------------
template <int NT>
__global__ __launch_bounds__(NT) void mix_end_kernel(const CaseMix *__restrict__ cases,
                                                     const int *__restrict__ slots)
{
    static_assert(NT == 256, "fixed reduction tree");
    __shared__ double red_s[NT / 64], red_m[NT / 64];
    const CaseMix &C = cases[slots[blockIdx.x]];
    CaseCtl *ctl = C.ctl;
    const int tid = threadIdx.x;
    if (*(volatile int *)&ctl->status != JST_RUN_ACTIVE) return;
    const int n = C.ni * C.nchunk;
    double s = 0.0, m = __builtin_huge_val();
    for (int t = tid; t < n; t += NT) {
        s += C.rpart[t];
        m = fmin(m, C.dpart[t]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off, 64);
        m = fmin(m, __shfl_xor(m, off, 64));
    }
    if ((tid & 63) == 0) {
        red_s[tid >> 6] = s;
        red_m[tid >> 6] = m;
    }
    __syncthreads();
    if (tid != 0) return;
    CaseCtl &v = *ctl;
    const PhysConst &k = C.k.k;
    if (v.err_key != JST_ERR_NONE) {
        v.status = JST_RUN_NONPHYSICAL;
        return;
    }
    const double tot = (red_s[0] + red_s[1]) + (red_s[2] + red_s[3]);
    const double res = sqrt(tot / (double)C.ni / (double)C.nj);
    const int rot = v.rot;
    v.totaltime = v.totaltime + v.dt;
    const long long st = v.step + 1;
    v.step = st;
    v.rot = rot ^ 1;
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
    v.dt = fmin(fmin(red_m[0], red_m[1]), fmin(red_m[2], red_m[3]));
}

template <int NT>
__global__ __launch_bounds__(NT) void mix_dt_kernel(const CaseMix *__restrict__ cases,
                                                    const PostDesc *__restrict__ posts)
{
    __shared__ double red_m[NT];
    __shared__ int last;
    const PostDesc pd = posts[blockIdx.x];
    const CaseMix &C = cases[pd.c];
    const PhysConst &k = C.k.k;
    CaseCtl *ctl = C.ctl;
    const int nj = C.nj, pitch = C.pitch;
    const int plane = C.ni * pitch;
    const int tid = threadIdx.x;
    const int rot = *(volatile int *)&ctl->rot;
    const double *u = C.u64[rot];
    double dmin = __builtin_huge_val();
    for (int o = tid; o < pd.count; o += NT) {
        const int cell = pd.first + o;
        const int i = cell / nj, j = cell - i * nj;
        const int c = cidx(i, j, pitch);
        const double rho = u[c], inv = 1.0 / rho;
        const double vx = u[plane + c] * inv, vy = u[2 * plane + c] * inv;
        const double p = (k.gamma - 1.0) * (u[3 * plane + c] - rho * (vx * vx + vy * vy) * 0.5);
        const double cs = sqrt(k.gamma * p * inv);
        const int jp = j + 1 < nj ? j + 1 : 0;
        const int cu = cidx(i + 1, j, pitch), cr = cidx(i, jp, pitch);
        const double aa = 0.5 * ((double)C.tnx[c] + (double)C.tnx[cu]);
        const double bb = 0.5 * ((double)C.tny[c] + (double)C.tny[cu]);
        const double ccv = 0.5 * ((double)C.nnx[c] + (double)C.nnx[cr]);
        const double ddv = 0.5 * ((double)C.nny[c] + (double)C.nny[cr]);
        const double conv = fabs(vx * aa + vy * bb) + fabs(vx * ccv + vy * ddv);
        const double acou = cs * (sqrt(aa * aa + bb * bb) + sqrt(ccv * ccv + ddv * ddv));
        const double vol = (double)C.vol[c];
        const double ldt = k.cfl * vol / (conv + acou);
        C.localdt[rot][c] = (float)ldt;
        C.ratio[rot][c] = (float)(vol / ldt);
        dmin = fmin(dmin, ldt);
    }
    red_m[tid] = dmin;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (tid < s) red_m[tid] = fmin(red_m[tid], red_m[tid + s]);
        __syncthreads();
    }
    if (tid == 0) {
        const double m = red_m[0];
        if (m == m) atomicMin(&ctl->dt_acc, (unsigned long long)__double_as_longlong(m));
        __threadfence();
        last = atomicAdd(&ctl->blocks_done, 1u) == (unsigned int)(C.npost - 1);
    }
    __syncthreads();
    if (!last || tid != 0) return;
    __threadfence();
    volatile CaseCtl *v = ctl;
    v->dt = __longlong_as_double((long long)v->dt_acc);
    v->dt_acc = JST_INF_BITS;
    v->blocks_done = 0u;
    __threadfence();
}

// FP64 geometry -> FP32 copies, and the FP32 shadow of the current FP64 state.
__global__ void to_f32_kernel(const double *__restrict__ a, float *__restrict__ b, size_t n)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        b[q] = (float)a[q];
}

__global__ void shadow_kernel(const CaseMix *__restrict__ cases, int ci)
{
    const CaseMix &C = cases[ci];
    const int rot = C.ctl->rot;
    const size_t plane = (size_t)C.ni * C.pitch;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < 5 * plane;
         q += (size_t)gridDim.x * blockDim.x) {
        const int m = (int)(q / plane);
        C.s32[rot][q] = (float)(C.u64[rot][q] - C.k.u64[m]);
    }
}

namespace jst {

template <int TI, int TJ, int NT>
static hipError_t stage_launch(const CaseMix *cases, const TileDesc *tiles, int ntiles, int grid,
                               int stage, double alpha, bool terms, hipStream_t s)
{
    if (grid <= 0 || grid > ntiles) grid = ntiles;
    if (terms)
        hipLaunchKernelGGL((mix_stage_kernel<TI, TJ, NT, true>), dim3(grid), dim3(NT), 0, s, cases,
                           tiles, ntiles, stage, alpha);
    else
        hipLaunchKernelGGL((mix_stage_kernel<TI, TJ, NT, false>), dim3(grid), dim3(NT), 0, s, cases,
                           tiles, ntiles, stage, alpha);
    return hipGetLastError();
}

const MixVariant MIX_VARIANTS[MIX_NVARIANTS] = {{16, 32, 256}, {8, 64, 256}, {8, 32, 256},
                                                 {16, 64, 512}};

hipError_t launch_stage_mixed(int variant, const CaseMix *cases, const TileDesc *tiles, int ntiles,
                              int grid, int stage, double alpha, bool terms, hipStream_t s)
{
    switch (variant) {
    case 0: return stage_launch<16, 32, 256>(cases, tiles, ntiles, grid, stage, alpha, terms, s);
    case 1: return stage_launch<8, 64, 256>(cases, tiles, ntiles, grid, stage, alpha, terms, s);
    case 2: return stage_launch<8, 32, 256>(cases, tiles, ntiles, grid, stage, alpha, terms, s);
    case 3: return stage_launch<16, 64, 512>(cases, tiles, ntiles, grid, stage, alpha, terms, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_end_mixed(const CaseMix *cases, const int *slots, int nslots, hipStream_t s)
{
    hipLaunchKernelGGL((mix_end_kernel<256>), dim3(nslots), dim3(256), 0, s, cases, slots);
    return hipGetLastError();
}

hipError_t launch_dt_mixed(const CaseMix *cases, const PostDesc *posts, int nposts, hipStream_t s)
{
    hipLaunchKernelGGL((mix_dt_kernel<256>), dim3(nposts), dim3(256), 0, s, cases, posts);
    return hipGetLastError();
}

hipError_t launch_to_f32(const double *a, float *b, size_t n, hipStream_t s)
{
    int grid = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(to_f32_kernel, dim3(grid > 0 ? grid : 1), dim3(256), 0, s, a, b, n);
    return hipGetLastError();
}

hipError_t launch_shadow(const CaseMix *cases, int ci, size_t n5, hipStream_t s)
{
    int grid = (int)std::min<size_t>((n5 + 255) / 256, 4096);
    hipLaunchKernelGGL(shadow_kernel, dim3(grid > 0 ? grid : 1), dim3(256), 0, s, cases, ci);
    return hipGetLastError();
}

} // namespace jst
