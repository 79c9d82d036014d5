// jst_mixed.hip — FP32-mixed path for gfx950 (batched: one launch covers many cases).
//
// The stage stencils read FP32 perturbations dU = U - U_inf and compute in FP32.  RK stages
// 1-4 write the FP32 stage state dU^(k) = dU^n - a_k dt R / V from the FP32 base dU^n; stage 5
// accumulates U^{n+1} = U^n - dt R / V in FP64 from the FP64 U^n and refreshes the FP32
// base.  Face fluxes are evaluated in perturbation form (F(U) - F(U_inf)) . n: the
// freestream part cancels exactly over a closed cell (metric closure, jst-rs
// geometry.rs:17-18) and so never enters a single-precision sum.  Every face flux is
// computed once per work item; results depend only on the case's own data, not on the
// launch grid or on the other cases of the launch.
//
// Scheme (same as the FP64 plain path): jst-rs boundary.rs (ghosts), convection.rs,
// gradient.rs, viscous.rs, dissipation.rs, source.rs, solver.rs:123-200 (update, residual).
#include <hip/hip_runtime.h>

#include <type_traits>

#include <algorithm>
#include <mutex>

#include "jst_dev.h"
#include "jst_launch.h"

namespace {

// 32-bit cell offsets: the host rejects cases whose planes exceed 2^31 / 8 elements
__device__ __forceinline__ unsigned cidx(int i, int j, int pitch) { return (unsigned)(i * pitch + j); }
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
        for (int q = 0; q < 5; ++q) u[q] = C.k.u64[q] + (double)(q < 4 ? S[4 * c + q] : S[4 * plane + c]);
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
// NEG: compiled with the negative-nu~ branch (taken where the case has it switched on; the
// wall ghosts included).
template <bool NEG>
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
    float mu_t = rnu * f1;
    float nu_coef = rnu;
    if (NEG && k.sa_neg) { // wave-uniform: extension spalart_allmaras.negative (ICCFD7-1902 eqs. 20-21)
        if (rnu < 0.0f) {
            mu_t = 0.0f;
            nu_coef = rnu * (16.0f + c3) * frcp(16.0f - c3);
        }
    }
    const float mu_eff = mu + mu_t;
    const float lam = mu * k.inv_pr + mu_t * k.inv_prt;
    const float txx = mu_eff * (4.0f / 3.0f * g[0] - 2.0f / 3.0f * g[3]);
    const float tyy = mu_eff * (4.0f / 3.0f * g[3] - 2.0f / 3.0f * g[0]);
    const float txy = mu_eff * (g[1] + g[2]);
    const float nu_diff = k.sigma * (mu + nu_coef);
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

// fw(r) of jst-rs source.rs:30-37 in FP64, as the plain path evaluates it.
__device__ __noinline__ double fw64(const PhysConst &k, double r)
{
    const double r2 = r * r, gg = r + k.cw2 * (r2 * r2 * r2 - r);
    const double g2 = gg * gg;
    return gg * cbrt(sqrt(k.one_plus_cw3_6 / (g2 * g2 * g2 + k.cw3_6)));
}

// jst-rs source.rs:45-88 in FP32 (x^(1/6) via exp2/log2).
template <bool NEG>
__device__ __forceinline__ float sa_source32(const MixConst &k, float drho, float drn, float nu,
                                             float chi, float fv1, const float g[8], float inv_d2,
                                             float vol)
{
    const float rho = k.rho_inf + drho;
    const float rnu = k.rnu_inf + drn;
    const float inv_kd2 = inv_d2 * k.inv_kappa2;
    const float ft2 = k.ct3 * __expf(-k.ct4 * chi * chi);
    const float fv2 = 1.0f - chi * frcp(1.0f + chi * fv1);
    const float s = k.fv3 * fabsf(g[2] - g[1]);
    const float st = fmaxf(s + nu * inv_kd2 * fv2, 1e-10f);
    float prod = k.cb1 * (1.0f - ft2) * st * rnu;
    const float r = fminf(nu * frcp(st) * inv_kd2, k.rmax);
    float fw;
    if (r >= -8.0f || (NEG && k.sa_neg)) { // with the negative branch, r < 0 lanes do not use fw
        // g^6 stays below 1e32 for r in [-8, rmax = 10]: FP32 is safe
        const float r2 = r * r, r6 = r2 * r2 * r2;
        const float gg = r + k.cw2 * (r6 - r);
        const float g2 = gg * gg, g6 = g2 * g2 * g2;
        const float q = k.one_plus_cw3_6 * frcp(g6 + k.cw3_6);
        fw = gg * __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(q) * (1.0f / 6.0f));
    } else {
        // negative nu~ under a clamped S~ gives a large negative r: r^6 and g^6 overflow
        // FP32 (inf * 0), so this rare case takes the FP64 formula of the plain path
        fw = (float)fw64(k.k, (double)r);
    }
    float destr = (k.cw1 * fw - k.cb1 * k.inv_kappa2 * ft2) * rho * (nu * nu * inv_d2);
    const float diff = k.cb2 * k.sigma * rho * (g[6] * g[6] + g[7] * g[7]);
    if (NEG && k.sa_neg) { // wave-uniform: extension spalart_allmaras.negative (ICCFD7-1902 eq. 19)
        if (rnu < 0.0f) {
            prod = k.cb1 * (1.0f - k.ct3) * s * rnu;
            destr = -k.cw1 * rho * (nu * nu * inv_d2);
        }
    }
    return (prod - destr + diff) * vol;
}

// Viscous spectral radius of a cell (extension simulation.viscous_time_step; Blazek,
// Computational Fluid Dynamics, 3rd ed., eqs. 6.21-6.22): max(4/3, gamma)/rho (mu/Pr +
// mu_t/Prt) (|A|^2 + |C|^2)/V, mu_t >= 0.  FP32 from the new state: rho, 1/rho, p, rho nu~.
template <bool NEG>
__device__ __forceinline__ float visc_radius32(const MixConst &k, float rho, float inv, float p, float rnu,
                                               float a, float b, float c, float d, float vol)
{
    const float t = p * inv * (1.0f / k.r_gas);
    const float r = t * k.inv_t_ref;
    const float mu = r * fsqrt(r) * k.suth_num * frcp(t + k.t_suth);
    const float chi = rnu * frcp(mu), c3 = chi * chi * chi;
    const float mut = (NEG && k.sa_neg && rnu < 0.0f) ? 0.0f : fmaxf(rnu * c3 * frcp(c3 + k.cv1_cubed), 0.0f);
    const float g = k.gamma > 4.0f / 3.0f ? k.gamma : 4.0f / 3.0f;
    return g * inv * (mu * k.inv_pr + mut * k.inv_prt) * (a * a + b * b + c * c + d * d) * frcp(vol);
}

// The same in FP64 (dt set-up from the FP64 state), in the operation order of the plain path.
__device__ __forceinline__ double visc_radius64(const PhysConst &k, double rho, double t, double rho_nu,
                                                double a, double b, double c, double d, double vol)
{
    const double r = t * k.inv_t_ref;
    const double mu = r * sqrt(r) * k.suth_num / (t + k.t_suth);
    const double chi = rho_nu / mu, c3 = chi * chi * chi;
    double mut = rho_nu * (c3 / (c3 + k.cv1_cubed));
    if (!(mut > 0.0) || (k.sa_neg && rho_nu < 0.0)) mut = 0.0;
    const double g = k.gamma > 4.0 / 3.0 ? k.gamma : 4.0 / 3.0;
    return g / rho * (mu * k.inv_pr + mut * k.inv_prt) * (a * a + b * b + c * c + d * d) / vol;
}

__device__ __forceinline__ float sensor32(float pinf4, float dm, float d0, float dp)
{
    // |p- - 2 p0 + p+| / (p- + 2 p0 + p+) with p = p_inf + dp (jst-rs dissipation.rs:29-31)
    return fabsf((dm - 2.0f * d0 + dp) * frcp(pinf4 + dm + 2.0f * d0 + dp));
}

// Lane l of a wave owns column J0 + l - 3 for 3 <= l < 61 (58 columns); lanes 0..2 and
// 61..63 carry the periodic halo (the widest stencil, a face's sensors, reaches 3 columns
// out).  Values of the j-neighbours move between lanes with DPP wave shifts, which read
// every lane, so they are only ever called from wave-uniform control flow.
// bound_ctrl: the edge lane with no source (lane 0, resp. 63: both halo lanes) reads 0, so
// the destination needs no initial value (no extra move per shift).
__device__ __forceinline__ float from_left(float x) // lane l receives lane l-1
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float from_right(float x) // lane l receives lane l+1
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x130, 0xf, 0xf, true));
}

// One face: perturbation Euler flux of the face average, JST dissipation with the 4-cell
// stencil um2 | um1 || u0 | up1 and the sensors of those cells, viscous flux from the
// diffusion tensors of the two cells (jst-rs convection.rs, dissipation.rs, viscous.rs).
// G = dFc - Fv - Fd.
__device__ __forceinline__ void face_flux(const MixConst &k, const float um2[5], const float um1[5],
                                          const float u0[5], const float up1[5], float s0, float s1,
                                          float s2, float s3, float lam, const float dA[8],
                                          const float dB[8], float nx, float ny, float G[5])
{
    lam *= k.cfl_half;
    const float mx = fmaxf(fmaxf(fmaxf(s0, s1), s2), s3);
    const float e2 = k.k2 * mx;
    const float e4 = fmaxf(k.k4 - e2, 0.0f);
    const float le2 = lam * e2, le4 = lam * e4;
    float fc[5];
    euler_flux_pert(k, um1, u0, nx, ny, fc);
    const float hx = 0.5f * nx, hy = 0.5f * ny;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        // d3 = up1 - 3 u0 + 3 um1 - um2 = (up1 - um2) - 3 d1; G = fc - (d1 le2 - d3 le4) - fv
        // as one chain of fused multiply-adds (18 fewer VALU instructions per row step)
        const float d1 = u0[m] - um1[m];
        const float d3 = __builtin_fmaf(-3.0f, d1, up1[m] - um2[m]);
        float g = __builtin_fmaf(d3, le4, __builtin_fmaf(-d1, le2, fc[m]));
        if (m > 0) { // the viscous flux: face average of the two cells' tensors . n
            g = __builtin_fmaf(-(dA[m - 1] + dB[m - 1]), hx, g);
            g = __builtin_fmaf(-(dA[m + 3] + dB[m + 3]), hy, g);
        }
        G[m] = g;
    }
}

enum { MIX_INNER = 0, MIX_LAST = 1, MIX_TERMS = 2 };

// Global-address-space views: loads and stores become global_* with a scalar base and a
// 32-bit lane offset instead of flat accesses with per-lane 64-bit addresses.
template <class T> using gptr = __attribute__((address_space(1))) T *;
template <class T> __device__ __forceinline__ gptr<const T> G(const T *p) { return (gptr<const T>)p; }
template <class T> __device__ __forceinline__ gptr<T> G(T *p) { return (gptr<T>)p; }
// Element q of a plane with a 32-bit byte offset (the host keeps every case below 4 GiB per
// array), so that the address is a scalar base plus a 32-bit lane offset.
template <class T> __device__ __forceinline__ T ldg(const T *p, unsigned q)
{
    return *(gptr<const T>)((gptr<const char>)p + q * (unsigned)sizeof(T));
}
template <class T> __device__ __forceinline__ void stg(T *p, unsigned q, T v)
{
    *(gptr<T>)((gptr<char>)p + q * (unsigned)sizeof(T)) = v;
}

// A multi-plane array through one buffer descriptor: every plane takes the same 32-bit lane
// offset and a scalar plane offset (soffset).  Five 64-bit bases per array held more scalar
// registers than the row loop has, and the loop restored the spilled ones with a lane read
// and a wait at every access.  The descriptor spans the array, raw addressing (stride 0),
// the gfx9 word 3.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t planes_rsrc(const void *p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float bld(__amdgpu_buffer_rsrc_t r, unsigned q, unsigned soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, q * 4u, soff, 0));
}
__device__ __forceinline__ double bld(__amdgpu_buffer_rsrc_t r, unsigned q, unsigned soff, double)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, q * 8u, soff, 0));
}
// Element q of an interleaved array of float pairs / quads (one wide load per lane).
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ f32x2 bld2(__amdgpu_buffer_rsrc_t r, unsigned q, unsigned soff)
{
    return __builtin_bit_cast(f32x2, __builtin_amdgcn_raw_buffer_load_b64(r, q * 8u, soff, 0));
}
__device__ __forceinline__ f32x4 bld4(__amdgpu_buffer_rsrc_t r, unsigned q, unsigned soff)
{
    return __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(r, q * 16u, soff, 0));
}
__device__ __forceinline__ void bst(__amdgpu_buffer_rsrc_t r, unsigned q, unsigned soff, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, q * 4u, soff, 0);
}
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void bst(__amdgpu_buffer_rsrc_t r, unsigned q, unsigned soff, double v)
{
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, q * 8u, soff, 0);
}
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
// An FP32 state array (dU, 5 components) of plane cells: components 0-3 of cell q as a quad
// at byte 16 q, component 4 in a plane after the quads, at byte 16 plane + 4 q.  A row's
// state is then one 16-byte and one 4-byte load (store) per lane instead of five.
__device__ __forceinline__ void sld5(__amdgpu_buffer_rsrc_t r, unsigned q, unsigned pb16, float v[5])
{
    const f32x4 a = bld4(r, q, 0u);
    v[0] = a.x;
    v[1] = a.y;
    v[2] = a.z;
    v[3] = a.w;
    v[4] = bld(r, q, pb16);
}
__device__ __forceinline__ void sst5(__amdgpu_buffer_rsrc_t r, unsigned q, unsigned pb16, const float v[5])
{
    const f32x4 a = {v[0], v[1], v[2], v[3]};
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, a), r, q * 16u, 0u, 0);
    bst(r, q, pb16, v[4]);
}

} // namespace

// ---------------------------------------------------------------------------------------
// FP32-mixed RK stage as a row march.  One wave per work item (case, 58-column group, row
// segment [is, ie)): lanes hold columns, the wave walks rows is-3 .. ie+2 keeping short
// windows of rows in registers (state, pressure, primitives, sensors, diffusion tensors,
// frozen V/dt), with the next row's loads issued one step ahead.  After loading row r it
//   - forms the primitives of row r (wall mirror / far-field Riemann ghosts synthesised),
//   - the i-sensor of row r-1 and the j-sensor of row r,
//   - the Green-Gauss gradient, diffusion tensor and SA source of row r-1,
//   - the tau face r-2 (between rows r-3 and r-2),
//   - the n faces and the residual of cell row r-3, and updates it.
// Every face is evaluated once; the right n face of a lane is its neighbour's left face.
// MIX_INNER (stages 1-4): U^(k) = U^n - a_k dt/V R in FP32 from the FP32 base dU^n.
// MIX_LAST (stage 5): FP64 U^{n+1} = U^n - dt/V R, the new FP32 base, the next frozen dt
// and V/dt, and the step-end partials per (row segment, 58-column group): summed down each
// column in row order, then across the group's columns by a fixed butterfly, so they depend
// on the case's segment length but never on the launch grid or on the other cases.
// MIX_TERMS: export R = Fc - Fv - Fd and S of the current state (no update).
// ---------------------------------------------------------------------------------------
template <int MODE, bool NEG>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void mix_march_kernel(const CaseMix *__restrict__ cases,
                                                        const TileDesc *__restrict__ items, int nitems,
                                                        int stage, double alpha)
{
    const int lane = threadIdx.x & 63;
    const int nwaves = gridDim.x * 4;
    // the work item is wave-uniform: keep it (and everything derived from it) in SGPRs
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    for (int it = blockIdx.x * 4 + wave; it < nitems; it += nwaves) {
        const TileDesc w = items[it];
        const CaseMix &C = cases[w.c];
        const MixConst &k = C.k;
        CaseCtl *ctl = C.ctl;
        const int ni = C.ni, nj = C.nj, pitch = C.pitch;
        const unsigned plane = (unsigned)(ni * pitch);
        const int is = w.i0, J0 = w.j0, ie = w.t;
        const int ngroups = (nj + MIX_COLS - 1) / MIX_COLS;
        // stage 5: this item's row segment (segments start at multiples of the case's length)
        const int seg = MODE == MIX_LAST ? is / C.seg_rows : 0;
        // control words, read once per item as they stand when the launch starts: the step end
        // that writes them is an earlier launch on the stream, and within this launch only
        // err_key changes (atomically, through the global pointer), whose new value no item
        // needs.  Through the constant address space they are scalar loads issued together
        // (one round trip instead of five dependent vector loads).
        const auto *cc = (const __attribute__((address_space(4))) CaseCtl *)ctl;
        const int status = cc->status;
        const long long step = cc->step;
        const unsigned long long ek = cc->err_key;
        const int rot = cc->rot;
        const double ctl_dt = cc->dt;
        const long long seq = step * 5 + stage;
        if (MODE != MIX_TERMS) {
            if (status != JST_RUN_ACTIVE) continue; // wave-uniform
            const int failed = __builtin_amdgcn_readfirstlane(
                (int)(ek != JST_ERR_NONE && (long long)(ek >> 32) < seq));
            if (failed) continue; // failed in an earlier stage
        }
        const float *src = (MODE == MIX_TERMS || stage == 0) ? C.sb : C.s32[(stage - 1) & 1];
        float *sdst = C.s32[stage & 1];
        // FP32 geometry, frozen V/dt and local dt: the case block holds them in the order tau2
        // (tau normal pairs), cell4 (n normal, V, 1/d^2 quads), ratio[0..1], localdt[0..1], so
        // one descriptor from tau2 with a byte offset per array covers them
        const char *g0 = (const char *)C.tau2;
        auto goff = [&](const float *p) { return (unsigned)((const char *)p - g0); };
        const __amdgpu_buffer_rsrc_t r_g = planes_rsrc(g0, goff(C.localdt[1]) + plane * 4u);
        const unsigned o_c4 = goff(C.cell4),
                       o_rt = goff(C.ratio[rot]), o_rt_next = goff(C.ratio[rot ^ 1]),
                       o_ldt = goff(C.localdt[rot]), o_ldt_next = goff(C.localdt[rot ^ 1]);
        const double a_dt = alpha * ctl_dt;

        const int jg = J0 - 3 + lane; // unwrapped column of this lane
        int jw = jg;
        while (jw < 0) jw += nj;
        while (jw >= nj) jw -= nj;
        const bool own = lane >= 3 && lane < 61 && jg < nj;

        // the 5-component state arrays (wave-uniform descriptors; the FP32 ones hold quads then a
        // plane of component 4, at byte 4 pb4; the FP64 ones five planes pb8 bytes apart)
        const unsigned pb4 = plane * 4u, pb8 = plane * 8u;
        const __amdgpu_buffer_rsrc_t r_src = planes_rsrc(src, 5u * pb4), r_sb = planes_rsrc(C.sb, 5u * pb4),
                                     r_dst = planes_rsrc(sdst, 5u * pb4),
                                     r_fm = planes_rsrc(C.u64[rot], 5u * pb8),
                                     r_un = planes_rsrc(C.u64[rot ^ 1], 5u * pb8);

        // register windows (index 0 = oldest row)
        float U[5][5];     // state dU, rows r-4 .. r
        float P[3];        // dp, rows r-2 .. r
        float PR[3][4];    // du dv dT nu~, rows r-2 .. r
        float SI[4];       // i-direction sensor, rows r-4 .. r-1
        float SJ[4];       // j-direction sensor, rows r-3 .. r
        float RT[4];       // frozen V/dt (row clamped), rows r-3 .. r
        float D[3][8];     // diffusion tensor, rows r-3 .. r-1
        float S[3];        // SA source, rows r-3 .. r-1
        float TN[4][2];    // tau-face normals, faces r-3 .. r
        float NN[3][2];    // n-face normals (left face of the lane's cell), rows r-3 .. r-1
        float IV[3], VOL[3]; // 1/V and V, rows r-3 .. r-1
        float GTp[5];      // tau face r-3 (previous step)
        float farc[10] = {}; // far-field ghost state (dU, dp, du, dv, dT, nu~)
        double racc = 0.0;       // stage 5: sum of (rho^{n+1} - rho^n)^2 over this row block
        float dacc = __builtin_huge_valf(); // stage 5: minimum next local dt over the block
        // bundle of row r, loaded during step r-1
        float bu[5] = {}, brt = 0.f, btn[2] = {}, bnn[2] = {}, bvol = 0.f, biw = 0.f;
        float bsb[5] = {}, bldt = 0.f;
        double bfm[5] = {}; // stage 5: U^n of the row the step updates
        // stages 1-4: the last updated row's stores, issued at the top of the next row step
        float pv[5] = {};
        unsigned pq = 0;
        bool pend = false;

        // Loads of row r's bundle.  In a steady step every condition below holds except that
        // row r may be the first far-field ghost row: its (unused) state load is clamped.
        auto load_bundle = [&](int r, auto steady) {
            constexpr bool ST = decltype(steady)::value;
            if (ST || r < ni) {
                const int rr = ST ? (r < ni ? r : ni - 1) : (r >= 0 ? r : -1 - r); // -1-r: wall mirror
                const unsigned q = cidx(rr, jw, pitch);
                sld5(r_src, q, 4u * pb4, bu);
            }
            // only what the segment's faces and cells use: V/dt of rows is-1 .. ie, tau
            // normals of faces is-1 .. ie+1, cell geometry of rows is-1 .. ie
            if (ST || (r >= is - 1 && r <= ie)) {
                const int rc = r < 0 ? 0 : (r > ni - 1 ? ni - 1 : r);
                brt = bld(r_g, cidx(rc, jw, pitch), o_rt);
            }
            if (ST || (r >= 0 && r <= ni && r >= is - 1 && r <= ie + 1)) {
                const f32x2 t = bld2(r_g, cidx(r, jw, pitch), 0u);
                btn[0] = t.x;
                btn[1] = t.y;
            }
            const int x = r - 1;
            if (ST || (x >= 0 && x < ni && x >= is - 1 && x <= ie)) {
                const f32x4 g = bld4(r_g, cidx(x, jw, pitch), o_c4);
                bnn[0] = g.x;
                bnn[1] = g.y;
                bvol = g.z;
                biw = g.w;
            }
            const int c = r - 3;
            if (MODE != MIX_TERMS && (ST || (c >= is && c < ie))) {
                const unsigned q = cidx(c, jw, pitch);
                if (MODE == MIX_INNER) {
                    sld5(r_sb, q, 4u * pb4, bsb);
                }
                if (MODE == MIX_LAST) {
#pragma unroll
                    for (int m = 0; m < 5; ++m) bfm[m] = bld(r_fm, q, m * pb8, 0.0);
                }
                if (k.lts) bldt = bld(r_g, q, o_ldt);
            }
        };

        // One row step.  steady = true: interior row with every stage of the pipeline active
        // (no ghost rows, no boundary special cases), compiled without any branch.
        auto row_step = [&](int r, auto steady) {
            constexpr bool ST = decltype(steady)::value;
            // ---- take the bundle of row r, prefetch row r+1 ----
            float un[5], rt_new = brt, tn_new[2] = {btn[0], btn[1]}, nn_new[2] = {bnn[0], bnn[1]};
            // 1/V from the FP32 volume: one reciprocal per row instead of 4 B/cell more to
            // read (the loads bound this kernel: 64-case batch 520 -> 505 us per step)
            const float iv_new = frcp(bvol), vol_new = bvol, iw_new = biw;
            float base[5];
            double former[5];
            float ldt_c = bldt;
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                un[m] = bu[m];
                base[m] = bsb[m];
                former[m] = bfm[m];
            }
            if (MODE == MIX_INNER) {
                // Wait for everything outstanding (the bundle loaded one row step ago, the
                // stores issued at the top of that step), then issue the previous row's stores
                // and this step's prefetch.  Without this wait, the wait the compiler places
                // for the previous bundle counts the stores too (loads and stores share one
                // counter) and lands after the new prefetch, waiting on it as well.
                __builtin_amdgcn_s_waitcnt(0x0F70); // vmcnt(0); expcnt, lgkmcnt untouched
                if (pend) {
                    sst5(r_dst, pq, 4u * pb4, pv);
                    pend = false;
                }
            }
            if (ST || r < ie + 2) load_bundle(r + 1, steady);
            // keep the prefetch and U^n loads ahead of the row's arithmetic (the scheduler
            // would otherwise sink some of them next to their use and wait on all loads there)
            __builtin_amdgcn_sched_barrier(0);
            // The row's arithmetic runs at raised wave priority: a wave that has its loads
            // issued and its operands arriving keeps the SIMD while the other resident wave
            // waits on memory, which spreads the two waves' load bursts apart in time
            // (64-case batch, 40-row segments: 567 -> 545 us per step).
            __builtin_amdgcn_s_setprio(1);

            // ---- row r: state and primitives, ghosts synthesised ----
            float dp, du, dv, dT, nut;
            if (ST || r < ni) {
                prim_from_du(k, un, dp, du, dv, dT, nut);
                if (!ST && r < 0) { // wall mirror (boundary.rs:30-51): scalars copied, u, v, nu~ negated
                    un[1] = -un[1] - 2.0f * k.m_inf;
                    un[2] = -un[2] - 2.0f * k.n_inf;
                    un[4] = -un[4] - 2.0f * k.rnu_inf;
                    du = -du - 2.0f * k.u_inf;
                    dv = -dv - 2.0f * k.v_inf;
                    nut = -nut;
                }
            } else {
                // far-field ghost rows ni .. ni+2 (boundary.rs:112-185): one state for all
                // three, formed when the march reaches row ni (always the first of them)
                if (r == ni) {
                    double fs[7];
                    far_state64(C, src, jw, fs);
                    const double *Ui = C.k.u64;
                    farc[0] = (float)(fs[0] - Ui[0]);
                    farc[1] = (float)(fs[0] * fs[3] - Ui[1]);
                    farc[2] = (float)(fs[0] * fs[4] - Ui[2]);
                    farc[3] = (float)(fs[0] * fs[5] - Ui[3]);
                    farc[4] = (float)(fs[0] * fs[6] - Ui[4]);
                    farc[5] = (float)(fs[1] - C.k.k.p_inf);
                    farc[6] = (float)(fs[3] - C.k.k.u_inf);
                    farc[7] = (float)(fs[4] - C.k.k.v_inf);
                    farc[8] = (float)(fs[2] - C.k.k.t_inf);
                    farc[9] = (float)fs[6];
                }
#pragma unroll
                for (int m = 0; m < 5; ++m) un[m] = farc[m];
                dp = farc[5];
                du = farc[6];
                dv = farc[7];
                dT = farc[8];
                nut = farc[9];
            }
#pragma unroll
            for (int m = 0; m < 5; ++m) U[4][m] = un[m];
            P[2] = dp;
            PR[2][0] = du;
            PR[2][1] = dv;
            PR[2][2] = dT;
            PR[2][3] = nut;
            RT[3] = rt_new;
            TN[3][0] = tn_new[0];
            TN[3][1] = tn_new[1];

            // ---- sensors: i-direction of row r-1, j-direction of row r ----
            const float pinf4 = 4.0f * k.p_inf;
            SI[3] = sensor32(pinf4, P[0], P[1], P[2]);
            SJ[3] = sensor32(pinf4, from_left(P[2]), P[2], from_right(P[2]));

            // ---- row x = r-1: gradient, diffusion tensor, SA source ----
            const int x = r - 1;
            NN[2][0] = nn_new[0];
            NN[2][1] = nn_new[1];
            IV[2] = iv_new;
            VOL[2] = vol_new;
            if (ST || (x >= (is - 1 > 0 ? is - 1 : 0) && x <= ie)) {
                float g[8];
                if (!ST && x == ni) {
#pragma unroll
                    for (int m = 0; m < 8; ++m) g[m] = 0.0f;
                } else {
                    const float upx = TN[3][0], upy = TN[3][1], dnx = TN[2][0], dny = TN[2][1];
                    const float lfx = NN[2][0], lfy = NN[2][1];
                    const float rtx = from_right(lfx), rty = from_right(lfy);
                    // sum over faces of (c + nb)/2 n / V = (c sum(n) + sum(nb n)) / (2V): the
                    // cell's own value enters once through the (vanishing) closure sum(n)
                    const float clx = upx - dnx + rtx - lfx, cly = upy - dny + rty - lfy;
                    const float hiv = 0.5f * IV[2];
#pragma unroll
                    for (int f = 0; f < 4; ++f) {
                        const float cc = PR[1][f], lf = from_left(cc), rt = from_right(cc);
                        const float sx = PR[2][f] * upx - PR[0][f] * dnx + rt * rtx - lf * lfx;
                        const float sy = PR[2][f] * upy - PR[0][f] * dny + rt * rty - lf * lfy;
                        g[2 * f] = (cc * clx + sx) * hiv;
                        g[2 * f + 1] = (cc * cly + sy) * hiv;
                    }
                }
                float chi, fv1;
                diffusion_tensor32<NEG>(k, PR[1][2], U[3][4], PR[1][0], PR[1][1], g, D[2], chi, fv1);
                if (ST || (x >= is && x < ie && x < ni))
                    S[2] = sa_source32<NEG>(k, U[3][0], U[3][4], PR[1][3], chi, fv1, g, iw_new, VOL[2]);
                if (!ST && x == 0) {
                    // ghost row -1 takes the gradient of row 0 without its temperature part
                    g[4] = 0.0f;
                    g[5] = 0.0f;
                    float chi_g, fv1_g;
                    diffusion_tensor32<NEG>(k, PR[0][2], U[2][4], PR[0][0], PR[0][1], g, D[1], chi_g, fv1_g);
                }
            }

            // ---- tau face f = r-2 (between rows f-1 and f) ----
            const int f = r - 2;
            float GT[5];
            if (ST || (f >= is && f <= ie))
                face_flux(k, U[0], U[1], U[2], U[3], SI[0], SI[1], SI[2], SI[3], RT[1] + RT[0], D[0], D[1],
                          TN[1][0], TN[1][1], GT);

            // ---- cell row c = r-3: n faces, residual, update ----
            const int c = r - 3;
            if (ST || (c >= is && c < ie)) {
                float um2[5], um1[5], up1[5], dA[8];
#pragma unroll
                for (int m = 0; m < 5; ++m) {
                    um1[m] = from_left(U[1][m]);
                    um2[m] = from_left(um1[m]);
                    up1[m] = from_right(U[1][m]);
                }
#pragma unroll
                for (int m = 0; m < 8; ++m) dA[m] = from_left(D[0][m]);
                const float s1 = from_left(SJ[0]);
                const float s0 = from_left(s1);
                const float s3 = from_right(SJ[0]);
                float GL[5], GR[5], R5[5];
                face_flux(k, um2, um1, U[1], up1, s0, s1, SJ[0], s3, RT[0] + from_left(RT[0]), dA, D[0],
                          NN[0][0], NN[0][1], GL);
#pragma unroll
                for (int m = 0; m < 5; ++m) {
                    GR[m] = from_right(GL[m]);
                    R5[m] = GT[m] - GTp[m] + GR[m] - GL[m];
                }
                const unsigned q = cidx(c, jw, pitch);
                if (MODE == MIX_TERMS) {
                    if (own) {
#pragma unroll
                        for (int m = 0; m < 5; ++m) stg(C.dbg + m * plane, q, R5[m]);
                        stg(C.dbg + 5 * plane, q, S[0]);
                    }
                } else if (MODE == MIX_INNER) {
                    R5[4] -= S[0];
                    const float sc = (float)(k.lts ? alpha * (double)ldt_c : a_dt) * IV[0];
                    float nw[5];
#pragma unroll
                    for (int m = 0; m < 5; ++m) nw[m] = base[m] - R5[m] * sc;
                    // p * rho = (gamma - 1) (E rho - |m|^2 / 2): the positivity test of p needs no
                    // division while rho > 0
                    const float rho = k.rho_inf + nw[0], mx = k.m_inf + nw[1], my = k.n_inf + nw[2];
                    const float prho = k.gm1 * ((k.e_inf + nw[3]) * rho - 0.5f * (mx * mx + my * my));
                    if (own) {
#pragma unroll
                        for (int m = 0; m < 5; ++m) pv[m] = nw[m];
                        pq = q;
                        pend = true;
                        // NonPhysical (solver.rs:151-185, state.rs:241-259), also catching NaN
                        if (!(rho > 1e-15f) || !(prho > 1e-15f * rho))
                            atomicMin(&ctl->err_key, ((unsigned long long)seq << 32) |
                                                         (unsigned long long)((long long)c * nj + jg));
                    }
                } else { // MIX_LAST
                    R5[4] -= S[0];
                    const double sc = (k.lts ? alpha * (double)ldt_c : a_dt) * (double)IV[0];
                    double u1[5];
#pragma unroll
                    for (int m = 0; m < 5; ++m) u1[m] = former[m] - (double)R5[m] * sc;
                    // NonPhysical test without a division: rho > 0 there, so
                    // p <= 1e-15 <=> p * rho <= 1e-15 * rho
                    const double prho = (k.k.gamma - 1.0) * (u1[3] * u1[0] - 0.5 * (u1[1] * u1[1] + u1[2] * u1[2]));
                    const bool bad = isnan(u1[0]) || u1[0] <= 1e-15 || isnan(prho) || prho <= 1e-15 * u1[0];
                    // next step's local dt and V/dt (timestep.rs:24-47), FP32
                    const float rho = (float)u1[0], inv = frcp(rho);
                    const float vx = (float)u1[1] * inv, vy = (float)u1[2] * inv;
                    const float pr = (float)prho * inv;
                    const float aa = 0.5f * (TN[0][0] + TN[1][0]), bb = 0.5f * (TN[0][1] + TN[1][1]);
                    const float cc = 0.5f * (NN[0][0] + from_right(NN[0][0]));
                    const float dd = 0.5f * (NN[0][1] + from_right(NN[0][1]));
                    const float cs = fsqrt(k.gamma * fmaxf(pr, 0.0f) * inv);
                    const float conv = fabsf(vx * aa + vy * bb) + fabsf(vx * cc + vy * dd);
                    const float acou = cs * (fsqrt(aa * aa + bb * bb) + fsqrt(cc * cc + dd * dd));
                    float lam = conv + acou;
                    if (k.visc_dt) // wave-uniform: the viscous spectral radius joins the local dt
                        lam += 4.0f * visc_radius32<NEG>(k, rho, inv, pr, (float)u1[4], aa, bb, cc, dd, VOL[0]);
                    const float ldt = k.cfl * VOL[0] / lam;
                    const double dr = u1[0] - former[0];
                    // this column's residual sum and dt minimum over the segment
                    racc += own ? dr * dr : 0.0;
                    dacc = own ? fminf(dacc, ldt) : dacc;
                    if (own) {
                        float nb[5];
#pragma unroll
                        for (int m = 0; m < 5; ++m) {
                            bst(r_un, q, m * pb8, u1[m]);
                            nb[m] = (float)(u1[m] - k.u64[m]);
                        }
                        sst5(r_sb, q, 4u * pb4, nb);
                        bst(r_g, q, o_ldt_next, ldt);
                        bst(r_g, q, o_rt_next, (conv + acou) * k.inv_cfl);
                        if (bad)
                            atomicMin(&ctl->err_key, ((unsigned long long)seq << 32) |
                                                         (unsigned long long)((long long)c * nj + jg));
                    }
                    if (c == ie - 1) { // segment end
                        // sum and minimum over the wave's columns: a fixed butterfly (every lane
                        // ends with the same value), then one partial per (segment, column group)
                        double rs = racc;
                        float dm = dacc;
#pragma unroll
                        for (int off = 32; off > 0; off >>= 1) {
                            rs += __shfl_xor(rs, off, 64);
                            dm = fminf(dm, __shfl_xor(dm, off, 64));
                        }
                        if (lane == 0) {
                            const unsigned ch = (unsigned)(seg * ngroups + J0 / MIX_COLS);
                            stg(C.rpart, ch, rs);
                            stg(C.dpart, ch, (double)dm);
                        }
                        racc = 0.0;
                        dacc = __builtin_huge_valf();
                    }
                }
            }

            __builtin_amdgcn_s_setprio(0);

            // ---- shift the windows ----
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                GTp[m] = GT[m];
#pragma unroll
                for (int q = 0; q < 4; ++q) U[q][m] = U[q + 1][m];
            }
            P[0] = P[1];
            P[1] = P[2];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                PR[0][m] = PR[1][m];
                PR[1][m] = PR[2][m];
            }
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                SI[q] = SI[q + 1];
                SJ[q] = SJ[q + 1];
                RT[q] = RT[q + 1];
                TN[q][0] = TN[q + 1][0];
                TN[q][1] = TN[q + 1][1];
            }
#pragma unroll
            for (int q = 0; q < 2; ++q) {
#pragma unroll
                for (int m = 0; m < 8; ++m) D[q][m] = D[q + 1][m];
                S[q] = S[q + 1];
                NN[q][0] = NN[q + 1][0];
                NN[q][1] = NN[q + 1][1];
                IV[q] = IV[q + 1];
                VOL[q] = VOL[q + 1];
            }
        };

        using steady_t = std::integral_constant<bool, true>;
        using edge_t = std::integral_constant<bool, false>;
        // steady rows: r >= 2 and r >= is + 3 (cell row r-3 is owned, no ghost rows behind),
        // r <= ie - 1 and r <= ni - 1 (the state of row r is interior)
        const int rs0 = is + 3 > 2 ? is + 3 : 2, rs1 = ie - 1 < ni - 1 ? ie - 1 : ni - 1;
        int r = is - 3;
        load_bundle(r, edge_t{});
        for (; r < rs0 && r <= ie + 2; ++r) row_step(r, edge_t{});
        // Two steady rows per trip, so that the row windows rename across the pair instead of
        // moving (stage 5: 820 instead of 889 instructions per row, 1.6 % faster; stages 1-4:
        // 609 instead of 666, and with the explicit wait and the stores one row step late,
        // 83.5 instead of 86.4 us per launch on the 64-case batch; pairs without them were
        // 10 % slower than single rows).
        if (MODE == MIX_LAST || MODE == MIX_INNER)
            for (; r + 1 <= rs1; r += 2) {
                row_step(r, steady_t{});
                row_step(r + 1, steady_t{});
            }
        for (; r <= rs1; ++r) row_step(r, steady_t{});
        for (; r <= ie + 2; ++r) row_step(r, edge_t{});
        if (MODE == MIX_INNER && pend) sst5(r_dst, pq, 4u * pb4, pv);
    }
}

// ---------------------------------------------------------------------------------------
// End of step, one block per case: residual sqrt(sum (rho^{n+1} - rho^n)^2 / NI / NJ)
// (jst-rs solver.rs:190-200) and the next global dt from the per-(row segment, column group)
// partials in a fixed order, totaltime += dt, step += 1, buffer rotation, and the stop decision
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
    // written by earlier launches: a scalar load through the constant address space
    if (((const __attribute__((address_space(4))) CaseCtl *)ctl)->status != JST_RUN_ACTIVE) return;
    const int n = ((C.ni + C.seg_rows - 1) / C.seg_rows) * ((C.nj + MIX_COLS - 1) / MIX_COLS);
    double s = 0.0, m = __builtin_huge_val();
    // sequential per lane (fixed order); unrolled so that the loads are in flight together
#pragma unroll 16
    for (int t = tid; t < n; t += NT) {
        s += ldg(C.rpart, (unsigned)t);
        m = fmin(m, ldg(C.dpart, (unsigned)t));
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
        const double aa = 0.5 * ((double)C.tau2[2 * c] + (double)C.tau2[2 * cu]);
        const double bb = 0.5 * ((double)C.tau2[2 * c + 1] + (double)C.tau2[2 * cu + 1]);
        const double ccv = 0.5 * ((double)C.cell4[4 * c] + (double)C.cell4[4 * cr]);
        const double ddv = 0.5 * ((double)C.cell4[4 * c + 1] + (double)C.cell4[4 * cr + 1]);
        const double conv = fabs(vx * aa + vy * bb) + fabs(vx * ccv + vy * ddv);
        const double acou = cs * (sqrt(aa * aa + bb * bb) + sqrt(ccv * ccv + ddv * ddv));
        const double vol = (double)C.cell4[4 * c + 2];
        const double dtc = k.cfl * vol / (conv + acou);
        double ldt = dtc;
        if (k.visc_dt) {
            const double t = p / (k.r_gas * rho);
            ldt = k.cfl * vol / (conv + acou + 4.0 * visc_radius64(k, rho, t, rho * (u[4 * plane + c] * inv),
                                                                   aa, bb, ccv, ddv, vol));
        }
        C.localdt[rot][c] = (float)ldt;
        C.ratio[rot][c] = (float)(vol / dtc);
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

__global__ void to_f32_strided_kernel(const double *__restrict__ a, float *__restrict__ b, size_t n,
                                      int stride, int lane)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        b[q * stride + lane] = (float)a[q];
}

__global__ void shadow_kernel(const CaseMix *__restrict__ cases, int ci)
{
    const CaseMix &C = cases[ci];
    const int rot = C.ctl->rot;
    const size_t plane = (size_t)C.ni * C.pitch;
    // the FP32 state layout: components 0-3 as quads per cell, component 4 as a plane after
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < 5 * plane;
         q += (size_t)gridDim.x * blockDim.x) {
        const int m = (int)(q / plane);
        const size_t c = q - (size_t)m * plane;
        C.sb[m < 4 ? 4 * c + m : 4 * plane + c] = (float)(C.u64[rot][q] - C.k.u64[m]);
    }
}

// ---------------------------------------------------------------------------------------
// Work-item list of one case in a fixed device array, for the entry points that allocate
// nothing.  The array holds JST_NOALLOC_MAX_ITEMS items (1 MiB) on every device the library
// runs on, for the life of the process; it is shared by every such call on that device, so
// calls that overlap on one device race.  Each call writes every item it then reads.
// ---------------------------------------------------------------------------------------
static __device__ __align__(256) TileDesc scratch[JST_NOALLOC_MAX_ITEMS];

// item k -> (case slot, column group k % ngroups, row segment k / ngroups of `rows` rows): the
// order of the batch's host-built lists
__global__ void make_items_kernel(TileDesc *__restrict__ items, int slot, int ni, int rows, int ngroups,
                                  int nitems)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nitems) return;
    const int grp = k % ngroups, seg = k / ngroups;
    const int i0 = seg * rows;
    items[k] = TileDesc{slot, i0, grp * MIX_COLS, min(i0 + rows, ni)};
}

namespace {
// device address of `scratch` per device, looked up once with that device current
constexpr int MAX_DEVICES = 64;
void *g_scratch_ptr[MAX_DEVICES] = {};
std::mutex g_scratch_mu;
} // namespace

namespace jst {

hipError_t noalloc_scratch(int device, TileDesc **ptr)
{
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) {
            const hipError_t e2 = hipSetDevice(cur);
            if (e == hipSuccess) e = e2;
        }
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *ptr = (TileDesc *)g_scratch_ptr[device];
    return hipSuccess;
}

hipError_t launch_make_items(TileDesc *items, int slot, int ni, int nj, int rows, int *nitems,
                             hipStream_t s)
{
    const int nseg = (ni + rows - 1) / rows, ngroups = (nj + MIX_COLS - 1) / MIX_COLS;
    const long n = (long)nseg * ngroups;
    if (n > JST_NOALLOC_MAX_ITEMS) return hipErrorInvalidValue;
    *nitems = (int)n;
    hipLaunchKernelGGL(make_items_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, items,
                       slot, ni, rows, ngroups, (int)n);
    return hipGetLastError();
}

const int MIX_SEGMENT_ROWS[MIX_NVARIANTS] = {16, 24, 32, 40, 48, 64, 8, 4, 2, 1};

template <bool NEG>
static void launch_stage_mixed_t(const CaseMix *cases, const TileDesc *items, int nitems, int grid,
                                 int stage, double alpha, bool terms, hipStream_t s)
{
    if (terms)
        hipLaunchKernelGGL((mix_march_kernel<MIX_TERMS, NEG>), dim3(grid), dim3(256), 0, s, cases, items,
                           nitems, stage, alpha);
    else if (stage == 4)
        hipLaunchKernelGGL((mix_march_kernel<MIX_LAST, NEG>), dim3(grid), dim3(256), 0, s, cases, items,
                           nitems, stage, alpha);
    else
        hipLaunchKernelGGL((mix_march_kernel<MIX_INNER, NEG>), dim3(grid), dim3(256), 0, s, cases, items,
                           nitems, stage, alpha);
}

hipError_t launch_stage_mixed(const CaseMix *cases, const TileDesc *items, int nitems, int grid,
                              int stage, double alpha, bool terms, bool neg, hipStream_t s)
{
    const int need = (nitems + 3) / 4; // four waves (work items) per workgroup
    if (grid <= 0 || grid > need) grid = need;
    if (grid <= 0) return hipSuccess;
    if (neg)
        launch_stage_mixed_t<true>(cases, items, nitems, grid, stage, alpha, terms, s);
    else
        launch_stage_mixed_t<false>(cases, items, nitems, grid, stage, alpha, terms, s);
    return hipGetLastError();
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

hipError_t launch_to_f32_strided(const double *a, float *b, size_t n, int stride, int lane, hipStream_t s)
{
    int grid = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(to_f32_strided_kernel, dim3(grid > 0 ? grid : 1), dim3(256), 0, s, a, b, n,
                       stride, lane);
    return hipGetLastError();
}

hipError_t launch_shadow(const CaseMix *cases, int ci, size_t n5, hipStream_t s)
{
    int grid = (int)std::min<size_t>((n5 + 255) / 256, 4096);
    hipLaunchKernelGGL(shadow_kernel, dim3(grid > 0 ? grid : 1), dim3(256), 0, s, cases, ci);
    return hipGetLastError();
}

} // namespace jst
