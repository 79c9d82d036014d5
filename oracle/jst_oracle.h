/*
 * jst_oracle.h — CPU oracle for the JST + Spalart-Allmaras residual-and-update loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64, loop-by-loop restatement of the
 * reference crate jst-rs (src/{config,mesh,geometry,field,state,boundary,timestep,
 * convection,gradient,viscous,dissipation,source,solver}.rs).  It is the checker that the
 * GPU library is compared against; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and never as the measured or shipped path.
 *
 * Layout mirrors jst-rs Field<T> (field.rs:143-192): every cell field is a row-major
 * array over the signed index space [-H, N+H), i slow, j fast, stride NJ+2H; Vec5 fields
 * are stored AoS (5 doubles per cell) like field.rs:37-39.  Face fields have halo 0.
 *
 * Pinning (see tests/test_oracle_golden.py, tests/test_oracle_properties.py):
 *   - every function up to residual/advance is pinned by the reference golden vectors
 *     tests/golden/cyl_10x12.json and cyl_17x40.json (meta, geometry, init, terms, final);
 *   - known answers from the reference unit tests (sensor, eps switch, fv1, fw, curl,
 *     Riemann free stream, Cw1 closure, mu(T0)=mu0);
 *   - invariants (metric closure, free-stream preservation, mass conservation,
 *     thread-count independence).
 *   - the extensions marked EXT below (nut_ratio, reynolds, local time stepping,
 *     residual_drop, viscous_time_step, sa_negative, forces) are not present in the
 *     reference; with their defaults the oracle is the reference's algorithm, and each is
 *     pinned by closed forms and invariants that a plausible mistake breaks
 *     (tests/test_oracle_properties.py, tests/test_oracle_sa_negative.py,
 *     tests/test_oracle_wall.py, tests/test_host_cpu.py).
 */
#ifndef JST_ORACLE_H
#define JST_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: mesh errors map 1:1 to MeshError (mesh.rs:18-25). */
enum {
    JO_OK = 0,
    JO_ERR_IO = 1,
    JO_ERR_BAD_HEADER = 2,
    JO_ERR_EMPTY = 3,
    JO_ERR_COUNT_MISMATCH = 4,
    JO_ERR_BAD_POINT = 5,
    JO_ERR_TOO_SMALL = 6,
    JO_ERR_NONPHYSICAL = 7,
    JO_ERR_CONFIG = 8
};

/* Config inputs (config.rs:10-98) plus default-off extensions. */
typedef struct {
    /* physics */
    double gamma, r_gas, t_ref, t_suth, mu0;
    /* simulation */
    double aoa_deg, mach, cfl;
    int halo;
    double t_inf, p_inf;
    /* solver */
    long iteration;
    double targetres;
    /* spalart_allmaras (sigma stores 1/sigma, config.rs:68-70) */
    double cv1, pr, prt, sigma, cb1, cb2, ct3, ct4, cw2, cw3, fv3, kappa, rmax;
    /* dissipation */
    double k2, k4;
    /* EXT: default-off extensions; defaults reproduce the reference exactly */
    double nut_ratio;        /* nu~_inf / nu_inf, reference hard-codes 0.1 (config.rs:196) */
    double reynolds;         /* <= 0: off; else mu0 rescaled so rho|V|L/mu_inf = reynolds */
    double ref_length;       /* L for Re and CL/CD (default 1) */
    int local_time_stepping; /* 0: global min dt (solver.rs:110,115) */
    double residual_drop;    /* decades below step-1 residual; <= 0: off */
    int viscous_time_step;   /* 0: convective spectral radius only (timestep.rs:24-47) */
    int sa_negative;         /* 0: fv1, source and nu~ diffusion as viscous.rs / source.rs for
                              * every sign of nu~; 1: the negative-nu~ branch of Allmaras,
                              * Johnson & Spalart (ICCFD7-1902, 2012) where nu~ < 0 */
} jo_config;

/* Derived constants (config.rs:114-146, finish 173-206). */
typedef struct {
    double cv, cp, cw1, cv1_cubed, c_inf, u_inf, v_inf, rho_inf, mu_inf, nut_inf;
    int supersonic;
    double suth_num, inv_t_ref, inv_pr, inv_prt, inv_kappa2, cw3_6, one_plus_cw3_6;
    double mu0_eff; /* mu0 after the EXT Reynolds rescale (== mu0 when off) */
} jo_derived;

typedef struct {
    int n_rings, n_theta;
    double *xy; /* 2 * n_rings * n_theta, ring-major */
} jo_mesh;

typedef struct jo_solver jo_solver;

/* ---- config ---- */
void jo_config_defaults_ext(jo_config *c);
void jo_config_finish(const jo_config *c, jo_derived *d);
double jo_mu(const jo_config *c, const jo_derived *d, double t);
double jo_sutherland(double mu0, double t, double t_ref, double t_suth);

/* ---- mesh (mesh.rs:109-189) ---- */
int jo_mesh_parse(const char *text, jo_mesh *out, long *err_a, long *err_b);
void jo_mesh_free(jo_mesh *m);

/* ---- solver ---- */
jo_solver *jo_solver_new(const jo_config *cfg, const jo_mesh *mesh);
void jo_solver_free(jo_solver *s);
int jo_ni(const jo_solver *s);
int jo_nj(const jo_solver *s);
int jo_halo(const jo_solver *s);
/* advance one step; returns JO_OK or JO_ERR_NONPHYSICAL and fills err[4] = {i,j,rho,p} */
int jo_advance(jo_solver *s, double *err);
double jo_residual(const jo_solver *s);
/* run up to max_steps (<=0: cfg.iteration); res_hist (may be NULL) receives each step's
 * residual; out[4] = {steps, residual, totaltime, converged} */
int jo_run(jo_solver *s, long max_steps, double *res_hist, double *out, double *err);
double jo_timestep(jo_solver *s);
void jo_residual_terms(jo_solver *s);
void jo_boundary_apply(jo_solver *s);
double jo_totaltime(const jo_solver *s);
long jo_step(const jo_solver *s);
void jo_set_uniform(jo_solver *s, double rho, double vx, double vy, double p, double nut);
/* unfused kernel entry points (used by the invariant tests) */
void jo_convection(jo_solver *s);
void jo_gradient(jo_solver *s);
void jo_viscous(jo_solver *s);
void jo_dissipation(jo_solver *s);
void jo_source(jo_solver *s);
/* copy the interior of a named field row-major into dst (NI*NJ, or (NI+1)*NJ for tau
 * faces); returns number of values or -1 for an unknown name.  Names:
 *   rho p T u v E H c nut U0..U4 Fc0..4 Fv0..4 Fd0..4 S localdt grad0..7 mu chi fv1
 *   vol cx cy wall_dist inv_vol inv_wd2 tau_nx tau_ny tau_mx tau_my n_nx n_ny n_mx n_my
 *   tau_flux0..4 tau_diff0..4 n_flux0..4 spec_ratio */
long jo_get(const jo_solver *s, const char *name, double *dst);
/* overwrite the interior conserved state (5 x NI x NJ, plane-major) and unpack */
int jo_set_state(jo_solver *s, const double *u_planes, double *err);
/* EXT: wall forces -> {Fx, Fy, CL, CD} from the current state (runs boundary+fluxes) */
void jo_forces(jo_solver *s, double *out4);
/* EXT: wall distribution per wall face j (NJ values each): face midpoint x, y, pressure
 * coefficient Cp and skin friction Cf along the unit tangent of increasing j (runs
 * boundary+fluxes) */
void jo_wall(jo_solver *s, double *x, double *y, double *cp, double *cf);
/* number of OpenMP threads the oracle will use */
int jo_threads(void);
/* set the OpenMP thread count (n <= 0: unchanged) */
void jo_set_threads(int n);
/* advance() split for tests: step-start snapshot + frozen global dt (returned), then one RK
 * stage U = U_former - alpha*dt*R/V (alpha*localdt with local time stepping) + unpack */
double jo_begin_step(jo_solver *s);
int jo_stage(jo_solver *s, double alpha, double dt, double *err);

#ifdef __cplusplus
}
#endif
#endif
