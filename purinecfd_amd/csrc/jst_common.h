// jst_common.h — host-side types and helpers shared by the single-case driver and the
// batched/serving driver: Config (jst-rs config.rs), Mesh (mesh.rs), error text.
#pragma once

#include <hip/hip_runtime_api.h>

#include <string>
#include <vector>

#include "../../include/jst_gpu.h"
#include "jst_dev.h"
#include "jst_launch.h"

namespace jsth {

struct Config {
    double gamma, r_gas, t_ref, t_suth, mu0;
    double aoa, mach, cfl;
    int halo;
    double t_inf, p_inf;
    long iteration = 10000;
    double targetres = 1e-10;
    double cv1, pr, prt, sigma, cb1, cb2, ct3, ct4, cw2, cw3, fv3, kappa, rmax;
    double k2, k4;
    // default-off extensions (absent from the reference; defaults reproduce it)
    double nut_ratio = 0.1, reynolds = 0.0, ref_length = 1.0, residual_drop = 0.0;
    bool lts = false;
    bool visc_dt = false; // simulation.viscous_time_step: viscous spectral radius in the local dt
    bool sa_neg = false;  // spalart_allmaras.negative: negative-nu~ branch of the SA model
};

struct Derived {
    double cv, cp, cw1, cv1_cubed, c_inf, u_inf, v_inf, rho_inf, mu_inf, nut_inf;
    bool supersonic;
    double suth_num, inv_t_ref, inv_pr, inv_prt, inv_kappa2, cw3_6, one_plus_cw3_6, mu0_eff;
};

struct Mesh {
    int n_rings = 0, n_theta = 0;
    std::vector<double> xy;
};

jst_status fail(jst_status st, const std::string &msg);
jst_status parse_config(const char *text, Config &c);
jst_status parse_mesh(const char *text, Mesh &m);
Derived finish(const Config &c);
PhysConst make_const(const Config &c, const Derived &d);
std::string rust_exp(double v, int prec);

// Wall distribution (jst_wall): per wall face j, Cp from the pressure the tau face (0, j)
// carries in its convective momentum flux, Cf from the tangential part of its viscous
// momentum flux.  d_terms is the wall-force kernel's scratch [4][nj] (Fc_x | Fc_y | Fv_x |
// Fv_y, device); g the case's FP64 geometry (device; row 0 of tnx/tny/tmx/tmy = the first nj
// entries).  Output arrays are host memory; any may be NULL.
jst_status wall_distribution(const double *d_terms, const jst::GeomPtrs &g, int nj,
                             const Config &c, const Derived &d, hipStream_t s, double *x,
                             double *y, double *cp, double *cf);

// Solver::unpack check of an injected state (solver.rs:151-171): the row-major first cell
// with rho or p NaN or <= 1e-15 fills err and fails with JST_ERR_NONPHYSICAL.  u_planes is
// 5 x NI x NJ, row-major per component.
jst_status check_injected_state(const double *u_planes, int ni, int nj, double gamma,
                                jst_nonphysical *err);

// single-case use of a batch (the FP32-mixed jst_ctx)
jst_status batch_forces(jst_batch *b, int case_id, double out[4]);
jst_status batch_wall(jst_batch *b, int case_id, double *x, double *y, double *cp, double *cf);
jst_status batch_run_case(jst_batch *b, int id, long limit, jst_step_cb cb, void *user,
                          jst_report *out, bool no_stop);

} // namespace jsth
