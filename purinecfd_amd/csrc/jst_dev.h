// jst_dev.h — structures shared by the host driver and the gfx950 kernels.
//
// One "case" is one O-mesh solve.  Its conserved state lives in HBM as structure-of-arrays
// planes [5][NI][pitch] with j (the circumferential index) fastest, so a wavefront's 64
// lanes read 64 consecutive cells of one ring (coalesced along the reference's fast axis,
// jst-rs field.rs:139-141).  No halo is stored: ghost cells (wall mirror, far-field Riemann
// state, periodic seam) are synthesised by the kernels' tile loaders from interior data.
#pragma once

#include <stdint.h>

#define JST_H 3                   // ghost layers (config.json IM; jst-rs config.rs:33-35)
#define JST_ERR_NONE 0xFFFFFFFFFFFFFFFFull
#define JST_INF_BITS 0x7FF0000000000000ull

// Per-case run control, resident in device memory and advanced by the kernels themselves,
// so that many steps can be queued (or replayed from a HIP graph) without host round trips.
struct CaseCtl {
    double dt;                    // frozen global dt of the step in progress (solver.rs:109-110)
    unsigned long long dt_acc;    // atomicMin accumulator of the next dt (bits of a positive double)
    double totaltime;             // accumulated physical time (solver.rs:111)
    long long step;               // completed steps
    long long stop_step;          // the device stops advancing when step reaches this
    double res_first;             // residual of step 1 (relative-drop stop criterion)
    double residual;              // residual of the last completed step
    unsigned long long err_key;   // (stage sequence << 32) | row-major cell index of first offender
    int status;                   // JST_RUN_*
    int rot;                      // buffer rotation 0..2 (U^n lives in ubuf[rot])
    unsigned int blocks_done;     // last-block-done counter of the step-end kernel
    int timing;                   // 1: ignore convergence / stop_step (benchmark timing)
};

enum { JST_RUN_ACTIVE = 0, JST_RUN_CONVERGED = 1, JST_RUN_STOPPED = 2, JST_RUN_NONPHYSICAL = 3 };

// Physical and scheme constants of a case (jst-rs config.rs Derived + inputs).
struct PhysConst {
    double gamma, r_gas, cfl, k2, k4;
    double t_inf, p_inf, rho_inf, u_inf, v_inf, c_inf, nut_inf;
    double suth_num, inv_t_ref, t_suth, inv_pr, inv_prt, cp, sigma, cv1_cubed;
    double cb1, cb2, ct3, ct4, cw1, cw2, cw3_6, one_plus_cw3_6, fv3, inv_kappa2, rmax;
    double targetres, residual_drop;
    int supersonic, lts;
    int visc_dt; // the local dt also holds the viscous spectral radius (default off)
    int sa_neg;  // negative-nu~ branch of the SA model where nu~ < 0 (default off)
};

// Device view of one case.  FP64 plain path: every array is double.
struct CaseDev {
    int ni, nj, pitch;            // pitch >= nj, multiple of 16 elements
    int ntiles;                   // tiles of this case in the tile list
    double *ubuf[3];              // 3 rotating buffers of 5 planes each
    const double *tnx, *tny;      // tau-face normals, (ni+1) x pitch
    const double *tmx, *tmy;      // tau-face midpoints, (ni+1) x pitch
    const double *nnx, *nny;      // n-face normals, ni x pitch
    const double *vol, *inv_vol, *inv_wd2, *cx, *cy;
    double *ratio;                // frozen V/dt_local per cell (jst-rs dissipation.rs:44-52), 2 planes
    double *localdt;              // frozen dt_local per cell (timestep.rs:24-47), 2 planes: the
                                  // current step reads plane (step & 1), stage 5 writes the other
    double *partial;              // step end: per-tile residual sums [ntiles] | tile min dt [ntiles]
    double *res_hist;             // residual ring buffer, res_cap entries
    int res_cap;
    CaseCtl *ctl;
    double *dbg;                  // residual-terms export planes (may be null)
    PhysConst k;
};

// Tile descriptor: x = case, y = i0, z = j0, w = tile index within the case.
struct TileDesc {
    int c, i0, j0, t;
};

// Export planes written by the residual-terms mode of the stage kernel.
enum {
    DBG_FC = 0, DBG_FV = 5, DBG_FD = 10, DBG_S = 15, DBG_GRAD = 16, DBG_CHI = 24, DBG_FV1 = 25,
    DBG_NPLANES = 26
};

// ---------------------------------------------------------------------------------------
// FP32-mixed path.  FP64: conserved state U^n / U^{n+1} (the RK accumulation), geometry
// precompute, far-field Riemann state, residual and force reductions.  FP32: a shadow of
// the perturbation dU = U - U_inf that the stage stencils read, and all stencil arithmetic
// (fluxes in perturbation form so that the freestream part, which cancels over a closed
// cell, never enters a single-precision sum).
// ---------------------------------------------------------------------------------------
struct MixConst {
    // FP32 scheme constants
    float gamma, gm1, r_gas, cfl, inv_cfl, cfl_half, k2, k4;
    float rho_inf, u_inf, v_inf, p_inf, t_inf, m_inf, n_inf, e_inf, h_inf, rnu_inf, nut_inf;
    float suth_num, inv_t_ref, t_suth, inv_pr, inv_prt, cp, sigma, cv1_cubed;
    float cb1, cb2, ct3, ct4, cw1, cw2, cw3_6, one_plus_cw3_6, fv3, inv_kappa2, rmax;
    int lts, visc_dt, sa_neg;
    // FP64 copies for the accumulation and the far-field state
    double u64[5];                // U_inf (conserved)
    PhysConst k;                  // full FP64 constants (far field, step control)
};

struct CaseMix {
    int ni, nj, pitch;
    int npost;                    // dt-refresh blocks of this case (fixed 1024-cell chunks)
    int seg_rows;                 // rows per row segment of the stage kernel's work items
    double *u64[2];               // U^n / U^{n+1}, 5 planes each
    // FP32 states, 5 components each: components 0-3 of cell c at [4c .. 4c+3], component 4
    // at [4 plane + c] (a row loads with one 16-byte and one 4-byte access per lane)
    float *sb;                    // FP32 base dU^n = U^n - U_inf
    float *s32[2];                // FP32 stage states dU^(k) (ping-pong)
    // FP32 geometry, interleaved so that the row march loads it with two wide loads per row:
    // tau2[2f + 0/1] = tau-normal (nx, ny) of face f, (ni+1) x pitch faces; cell4[4c + 0..3] =
    // n-normal (nx, ny), V and 1/d^2 of cell c, ni x pitch cells (1/V is formed from V)
    const float *tau2, *cell4;
    const double *tnx64, *tny64, *tmx64, *tmy64, *cx64, *cy64; // far-field support (FP64)
    float *ratio[2], *localdt[2]; // frozen per step; [rot] is read by the step, [rot^1] written
    double *rpart;                // residual partial sums per (row segment, 58-column group)
    double *dpart;                // next-dt minima per (row segment, 58-column group)
    double *res_hist;
    int res_cap;
    CaseCtl *ctl;
    float *dbg;                   // residual-terms export (may be null): 16 planes R0..R4,S,...
    MixConst k;
};

// dt-refresh work item: case, first cell (row-major), block index within the case.
struct PostDesc {
    int c, first, count, b;
};
#define JST_POST_CELLS 1024
