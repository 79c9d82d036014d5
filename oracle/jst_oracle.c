/*
 * jst_oracle.c — CPU oracle (TEST INFRASTRUCTURE ONLY; see jst_oracle.h).
 *
 * A plain-C FP64 restatement of jst-rs, kept deliberately unfused and in the reference's
 * operation order so that it reproduces the reference golden vectors.  Build with
 * -ffp-contract=off (Rust never contracts a*b+c into an FMA).  Parallel loops are OpenMP
 * row loops in which every output element is written exactly once, so results do not
 * depend on the thread count (jst-rs lib.rs:18-23, field.rs:286-314).
 */
#include "jst_oracle.h"

#include <ctype.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_D 3.14159265358979323846

/* ------------------------------------------------------------------------------------ */
/* config (jst-rs config.rs)                                                             */
/* ------------------------------------------------------------------------------------ */

void jo_config_defaults_ext(jo_config *c)
{
    c->nut_ratio = 0.1; /* config.rs:196 hard-codes 0.1 */
    c->reynolds = 0.0;
    c->ref_length = 1.0;
    c->local_time_stepping = 0;
    c->residual_drop = 0.0;
    c->viscous_time_step = 0;
    c->sa_negative = 0;
}

/* config.rs:153-156: mu0 * (r*sqrt r) * (T0+Ts) / (T+Ts) */
double jo_sutherland(double mu0, double t, double t_ref, double t_suth)
{
    double r = t / t_ref;
    return mu0 * (r * sqrt(r)) * (t_ref + t_suth) / (t + t_suth);
}

/* LLVM's constant-exponent powi expansion (square-and-multiply, low bit first), which is
 * what Rust's f64::powi lowers to at -O3. */
static double powi_c(double x, int n)
{
    double res = 0.0, sq = x;
    int have = 0;
    while (n) {
        if (n & 1) {
            if (have) res = res * sq;
            else { res = sq; have = 1; }
        }
        sq = sq * sq;
        n >>= 1;
    }
    return have ? res : 1.0;
}

/* config.rs:173-206 (Config::finish).  EXT: Reynolds rescale of mu0 and nut_ratio. */
void jo_config_finish(const jo_config *c, jo_derived *d)
{
    double cv = c->r_gas / (c->gamma - 1.0);
    double c_inf = sqrt(c->gamma * c->r_gas * c->t_inf);
    double aoa = c->aoa_deg * (PI_D / 180.0); /* f64::to_radians */
    double rho_inf = c->p_inf / (c->r_gas * c->t_inf);
    double mu0 = c->mu0;
    double u_inf = c_inf * c->mach * cos(aoa);
    double v_inf = c_inf * c->mach * sin(aoa);
    if (c->reynolds > 0.0) {
        /* EXT: scale mu0 so that rho_inf*|V_inf|*L/mu_inf == reynolds (mu_inf is linear in
         * mu0).  Default off: the reference has no Reynolds input. */
        double mu_ref = jo_sutherland(mu0, c->t_inf, c->t_ref, c->t_suth);
        double re_now = rho_inf * hypot(u_inf, v_inf) * c->ref_length / mu_ref;
        mu0 = mu0 * (re_now / c->reynolds);
    }
    double mu_inf = jo_sutherland(mu0, c->t_inf, c->t_ref, c->t_suth);
    d->mu0_eff = mu0;
    d->cv = cv;
    d->cp = c->gamma * cv;
    d->cw1 = c->cb1 / (c->kappa * c->kappa) + (1.0 + c->cb2) * c->sigma;
    d->cv1_cubed = powi_c(c->cv1, 3);
    d->c_inf = c_inf;
    d->u_inf = u_inf;
    d->v_inf = v_inf;
    d->rho_inf = rho_inf;
    d->mu_inf = mu_inf;
    d->nut_inf = c->nut_ratio * mu_inf / rho_inf;
    d->supersonic = hypot(u_inf, v_inf) >= c_inf;
    d->suth_num = mu0 * (c->t_ref + c->t_suth);
    d->inv_t_ref = 1.0 / c->t_ref;
    d->inv_pr = 1.0 / c->pr;
    d->inv_prt = 1.0 / c->prt;
    d->inv_kappa2 = 1.0 / (c->kappa * c->kappa);
    d->cw3_6 = powi_c(c->cw3, 6);
    d->one_plus_cw3_6 = 1.0 + powi_c(c->cw3, 6);
}

/* config.rs:210-213 (Config::mu) */
double jo_mu(const jo_config *c, const jo_derived *d, double t)
{
    double r = t * d->inv_t_ref;
    return r * sqrt(r) * d->suth_num / (t + c->t_suth);
}

/* ------------------------------------------------------------------------------------ */
/* mesh (jst-rs mesh.rs:109-189)                                                         */
/* ------------------------------------------------------------------------------------ */

/* Rust `usize::from_str`: optional '+', then at least one ASCII digit, nothing else. */
static int parse_usize(const char *s, size_t n, long *out)
{
    size_t k = 0;
    long v = 0;
    if (n > 0 && s[0] == '+') k = 1;
    if (k >= n) return 0;
    for (; k < n; ++k) {
        if (s[k] < '0' || s[k] > '9') return 0;
        if (v > (long)1e15) return 0;
        v = v * 10 + (s[k] - '0');
    }
    *out = v;
    return 1;
}

/* Rust `f64::from_str` accepts decimal/exponent forms and inf/infinity/nan; it rejects
 * hex floats and trailing junk, which strtod would accept. */
static int parse_f64(const char *s, size_t n, double *out)
{
    char buf[128];
    char *end;
    size_t k;
    if (n == 0 || n >= sizeof buf) return 0;
    memcpy(buf, s, n);
    buf[n] = 0;
    for (k = 0; k < n; ++k) {
        char ch = buf[k];
        if (ch == 'x' || ch == 'X' || ch == '(' || ch == ')' || ch == 'p' || ch == 'P')
            return 0;
    }
    *out = strtod(buf, &end);
    return end == buf + n;
}

/* next whitespace-delimited token within [p, e) */
static int next_tok(const char **p, const char *e, const char **ts, size_t *tn)
{
    const char *q = *p;
    while (q < e && isspace((unsigned char)*q)) ++q;
    if (q >= e) { *p = q; return 0; }
    *ts = q;
    while (q < e && !isspace((unsigned char)*q)) ++q;
    *tn = (size_t)(q - *ts);
    *p = q;
    return 1;
}

int jo_mesh_parse(const char *text, jo_mesh *out, long *err_a, long *err_b)
{
    const char *p = text, *le;
    const char *ts;
    size_t tn;
    long ni, nj, cap, count = 0, lineno = 1;
    double *xy = NULL;
    *err_a = *err_b = 0;
    out->xy = NULL;
    out->n_rings = out->n_theta = 0;
    if (*p == 0) return JO_ERR_EMPTY; /* lines().next() is None */
    le = strchr(p, '\n');
    if (!le) le = p + strlen(p);
    {
        const char *q = p;
        int ok = next_tok(&q, le, &ts, &tn) && parse_usize(ts, tn, &ni);
        ok = ok && next_tok(&q, le, &ts, &tn) && parse_usize(ts, tn, &nj);
        if (!ok) return JO_ERR_BAD_HEADER;
    }
    cap = ni * nj;
    xy = (double *)malloc(sizeof(double) * 2 * (size_t)(cap > 0 ? cap : 1));
    p = (*le == '\n') ? le + 1 : le;
    while (*p) {
        const char *q;
        double x, y;
        le = strchr(p, '\n');
        if (!le) le = p + strlen(p);
        ++lineno;
        q = p;
        if (next_tok(&q, le, &ts, &tn)) {
            int ok = parse_f64(ts, tn, &x);
            const char *t2;
            size_t n2;
            ok = ok && next_tok(&q, le, &t2, &n2) && parse_f64(t2, n2, &y);
            if (!ok) {
                free(xy);
                *err_a = lineno;
                return JO_ERR_BAD_POINT;
            }
            if (count < cap) {
                xy[2 * count] = x;
                xy[2 * count + 1] = y;
            }
            ++count;
        }
        p = (*le == '\n') ? le + 1 : le;
    }
    if (count == 0) { free(xy); return JO_ERR_EMPTY; }
    if (count != cap) {
        free(xy);
        *err_a = cap;
        *err_b = count;
        return JO_ERR_COUNT_MISMATCH;
    }
    /* trim_closing_duplicate (mesh.rs:169-189) */
    if (nj >= 2) {
        int closed = 1;
        long r;
        for (r = 0; r < ni && closed; ++r) {
            double ax = xy[2 * (r * nj)], ay = xy[2 * (r * nj) + 1];
            double bx = xy[2 * (r * nj + nj - 1)], by = xy[2 * (r * nj + nj - 1) + 1];
            closed = fabs(ax - bx) <= 1e-10 && fabs(ay - by) <= 1e-10;
        }
        if (closed) {
            long nn = nj - 1, k;
            for (r = 0; r < ni; ++r)
                for (k = 0; k < nn; ++k) {
                    xy[2 * (r * nn + k)] = xy[2 * (r * nj + k)];
                    xy[2 * (r * nn + k) + 1] = xy[2 * (r * nj + k) + 1];
                }
            nj = nn;
        }
    }
    if (ni < 4 || nj < 8) {
        free(xy);
        *err_a = ni;
        *err_b = nj;
        return JO_ERR_TOO_SMALL;
    }
    out->n_rings = (int)ni;
    out->n_theta = (int)nj;
    out->xy = xy;
    return JO_OK;
}

void jo_mesh_free(jo_mesh *m)
{
    free(m->xy);
    m->xy = NULL;
}

/* ------------------------------------------------------------------------------------ */
/* solver state (state.rs:67-336)                                                        */
/* ------------------------------------------------------------------------------------ */

struct jo_solver {
    jo_config cfg;
    jo_derived d;
    int ni, nj, h, stride, rows;
    size_t ncell;
    /* geometry.rs:44-63 */
    double *vol, *cx, *cy, *wd, *inv_vol, *inv_wd2;
    double *tau; /* (ni+1) x nj x {nx,ny,mx,my} */
    double *nrm; /* ni x nj x 4 */
    /* Cells */
    double *u, *u_former; /* x5 */
    double *rho, *p, *t, *vx, *vy, *e, *hh, *c, *nut;
    double *grad; /* x8: dudx dudy dvdx dvdy dtdx dtdy dnutdx dnutdy */
    double *aux;  /* x3: mu chi fv1 */
    double *diff; /* x10: x[0..5] y[0..5] */
    double *fc, *fv, *fd; /* x5 */
    double *src, *localdt;
    double *dtc; /* convective local dt, the JST spectral radius (== localdt unless EXT viscous dt) */
    /* Faces: tau (ni+1) x nj, nrm ni x nj, halo 0 */
    double *tflux, *tdiff, *tdiss, *tlam, *teps;
    double *nflux, *ndiff, *ndiss, *nlam, *neps;
    double *sensor_i, *sensor_j, *spec_ratio;
    double *rho_prev;
    double totaltime;
    long step;
    double res_first;
};

#define CI(s, i, j) ((size_t)((i) + (s)->h) * (size_t)(s)->stride + (size_t)((j) + (s)->h))
#define FI(s, i, j) ((size_t)(i) * (size_t)(s)->nj + (size_t)(j))

static double *zalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

int jo_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- geometry.rs:76-216 ---- */
static void mesh_node(const jo_mesh *m, int ring, int theta, double *x, double *y)
{
    size_t k = (size_t)ring * m->n_theta + (size_t)(theta % m->n_theta);
    *x = m->xy[2 * k];
    *y = m->xy[2 * k + 1];
}

static void build_geometry(jo_solver *s, const jo_mesh *m)
{
    int ni = s->ni, nj = s->nj, i, j;
    /* build_cells 109-145 */
    for (i = 0; i < ni; ++i)
        for (j = 0; j < nj; ++j) {
            double ax, ay, bx, by, cx, cy, dx, dy;
            mesh_node(m, i, j, &ax, &ay);
            mesh_node(m, i + 1, j, &bx, &by);
            mesh_node(m, i + 1, j + 1, &cx, &cy);
            mesh_node(m, i, j + 1, &dx, &dy);
            double d1x = cx - ax, d1y = cy - ay;
            double d2x = bx - dx, d2y = by - dy;
            double vol = 0.5 * fabs(d1x * d2y - d1y * d2x);
            double cr0 = ax * by - bx * ay;
            double cr1 = bx * cy - cx * by;
            double cr2 = cx * dy - dx * cy;
            double cr3 = dx * ay - ax * dy;
            double sgn = 0.5 * (cr0 + cr1 + cr2 + cr3);
            double sx = (ax + bx) * cr0 + (bx + cx) * cr1 + (cx + dx) * cr2 + (dx + ax) * cr3;
            double sy = (ay + by) * cr0 + (by + cy) * cr1 + (cy + dy) * cr2 + (dy + ay) * cr3;
            s->vol[CI(s, i, j)] = vol;
            if (fabs(sgn) > 1e-30) {
                s->cx[CI(s, i, j)] = sx / (6.0 * sgn);
                s->cy[CI(s, i, j)] = sy / (6.0 * sgn);
            }
        }
    /* build_faces 148-185 */
    for (i = 0; i <= ni; ++i)
        for (j = 0; j < nj; ++j) {
            double ax, ay, bx, by;
            mesh_node(m, i, j, &ax, &ay);
            mesh_node(m, i, j + 1, &bx, &by);
            double dx = bx - ax, dy = by - ay;
            double *f = s->tau + 4 * FI(s, i, j);
            f[0] = dy;
            f[1] = -dx;
            f[2] = 0.5 * (ax + bx);
            f[3] = 0.5 * (ay + by);
        }
    for (i = 0; i < ni; ++i)
        for (j = 0; j < nj; ++j) {
            double ax, ay, bx, by;
            mesh_node(m, i, j, &ax, &ay);
            mesh_node(m, i + 1, j, &bx, &by);
            double dx = bx - ax, dy = by - ay;
            double *f = s->nrm + 4 * FI(s, i, j);
            f[0] = -dy;
            f[1] = dx;
            f[2] = 0.5 * (ax + bx);
            f[3] = 0.5 * (ay + by);
        }
    /* build_wall_distance 192-216: window = min(max(15, nj/5), nj/2) */
    {
        int w = nj / 5 > 15 ? nj / 5 : 15;
        if (w > nj / 2) w = nj / 2;
#pragma omp parallel for schedule(static)
        for (i = 0; i < ni; ++i) {
            int jj;
            for (jj = 0; jj < nj; ++jj) {
                double px = s->cx[CI(s, i, jj)], py = s->cy[CI(s, i, jj)];
                double best = INFINITY;
                int dk;
                for (dk = -w; dk <= w; ++dk) {
                    int k = ((jj + dk) % nj + nj) % nj; /* rem_euclid */
                    const double *f = s->tau + 4 * FI(s, 0, k);
                    double ex = px - f[2], ey = py - f[3];
                    double dd = sqrt(ex * ex + ey * ey);
                    if (dd < best) best = dd;
                }
                s->wd[CI(s, i, jj)] = best;
            }
        }
    }
    for (i = 0; i < ni; ++i)
        for (j = 0; j < nj; ++j) {
            s->inv_vol[CI(s, i, j)] = 1.0 / s->vol[CI(s, i, j)];
            double dd = s->wd[CI(s, i, j)];
            s->inv_wd2[CI(s, i, j)] = 1.0 / (dd * dd);
        }
}

/* state.rs:142-156 Cells::pack */
static void pack(jo_solver *s, int i, int j)
{
    size_t k = CI(s, i, j);
    double rho = s->rho[k];
    double *u = s->u + 5 * k;
    u[0] = rho;
    u[1] = rho * s->vx[k];
    u[2] = rho * s->vy[k];
    u[3] = rho * s->e[k];
    u[4] = rho * s->nut[k];
}

/* state.rs:222-237 PrimState::from_primitives, then write_state */
static void write_prim(jo_solver *s, int i, int j, double rho, double vx, double vy, double p,
                       double nut)
{
    double gamma = s->cfg.gamma;
    double t = p / (s->cfg.r_gas * rho);
    double e = p / (rho * (gamma - 1.0)) + 0.5 * (vx * vx + vy * vy);
    size_t k = CI(s, i, j);
    s->rho[k] = rho;
    s->p[k] = p;
    s->t[k] = t;
    s->vx[k] = vx;
    s->vy[k] = vy;
    s->e[k] = e;
    s->hh[k] = e + p / rho;
    s->c[k] = sqrt(gamma * s->cfg.r_gas * t);
    s->nut[k] = nut;
    pack(s, i, j);
}

/* ---- boundary.rs:23-185 ---- */
static void wall_mirror(jo_solver *s)
{
    int k, j;
    for (k = 0; k < s->h; ++k) {
        int g = -1 - k;
        for (j = 0; j < s->nj; ++j) {
            size_t a = CI(s, g, j), b = CI(s, k, j);
            s->rho[a] = s->rho[b];
            s->p[a] = s->p[b];
            s->t[a] = s->t[b];
            s->e[a] = s->e[b];
            s->hh[a] = s->hh[b];
            s->c[a] = s->c[b];
            s->vx[a] = -s->vx[b];
            s->vy[a] = -s->vy[b];
            s->nut[a] = -s->nut[b];
            pack(s, g, j);
        }
    }
}

static void periodic_seam(jo_solver *s)
{
    int i, k, w;
    int nj = s->nj;
    for (i = 0; i < s->ni; ++i)
        for (k = 1; k <= s->h; ++k)
            for (w = 0; w < 2; ++w) {
                int g = w == 0 ? -k : nj - 1 + k;
                int src = w == 0 ? nj - k : k - 1;
                size_t a = CI(s, i, g), b = CI(s, i, src);
                s->rho[a] = s->rho[b];
                s->p[a] = s->p[b];
                s->t[a] = s->t[b];
                s->e[a] = s->e[b];
                s->hh[a] = s->hh[b];
                s->c[a] = s->c[b];
                s->vx[a] = s->vx[b];
                s->vy[a] = s->vy[b];
                s->nut[a] = s->nut[b];
                pack(s, i, g);
            }
}

/* boundary.rs:158-185 */
static double extrapolate_nut(const jo_solver *s, int j)
{
    const double FLOOR = 1e-10;
    int ni = s->ni, k;
    const double *f = s->tau + 4 * FI(s, ni, j);
    double dist[3], nut[3];
    for (k = 0; k < 3; ++k) {
        int i = ni - 1 - k;
        dist[k] = hypot(s->cx[CI(s, i, j)] - f[2], s->cy[CI(s, i, j)] - f[3]);
        nut[k] = s->nut[CI(s, i, j)];
    }
    double l1 = dist[0], l2 = dist[1], l3 = dist[2];
    double d12 = l1 - l2, d13 = l1 - l3, d23 = l2 - l3;
    if (fabs(d12) < 1e-14 || fabs(d13) < 1e-14 || fabs(d23) < 1e-14)
        return fmax(nut[0], FLOOR);
    double v = l2 * l3 / (d12 * d13) * nut[0] + l1 * l3 / (-d12 * d23) * nut[1] +
               l1 * l2 / (d23 * -d13) * nut[2];
    return v > FLOOR ? v : FLOOR;
}

/* boundary.rs:112-155 riemann_face -> {rho,p,t,vx,vy,e,nut} */
static void riemann_face(const jo_solver *s, int j, double *st)
{
    int ni = s->ni;
    const double *f = s->tau + 4 * FI(s, ni, j);
    double len = hypot(f[0], f[1]);
    double nx = f[0], ny = f[1];
    double gamma = s->cfg.gamma;
    const jo_derived *d = &s->d;
    double vn_inf = (d->u_inf * nx + d->v_inf * ny) / len;
    double vt_inf = (d->u_inf * ny - d->v_inf * nx) / len;
    size_t kin = CI(s, ni - 1, j);
    double uin = s->vx[kin], vin = s->vy[kin];
    double vn_in = (uin * nx + vin * ny) / len;
    double vt_in = (uin * ny - vin * nx) / len;
    double c_face, vn_face, vt_face;
    int inflow;
    if (!d->supersonic) {
        double r_in = vn_in + 2.0 * s->c[kin] / (gamma - 1.0);
        double r_inf = vn_inf - 2.0 * d->c_inf / (gamma - 1.0);
        double vn = 0.5 * (r_in + r_inf);
        double c = (gamma - 1.0) / 4.0 * (r_in - r_inf);
        inflow = vn <= 0.0;
        c_face = c;
        vn_face = vn;
        vt_face = inflow ? vt_inf : vt_in;
    } else if (vn_inf <= 0.0) {
        c_face = d->c_inf;
        vn_face = vn_inf;
        vt_face = vt_inf;
        inflow = 1;
    } else {
        c_face = s->c[kin];
        vn_face = vn_in;
        vt_face = vt_in;
        inflow = 0;
    }
    double t = c_face * c_face / (s->cfg.r_gas * gamma);
    double rho = d->rho_inf * pow(t / s->cfg.t_inf, 1.0 / (gamma - 1.0));
    double p = rho * s->cfg.r_gas * t;
    double vx = (vt_face * ny + vn_face * nx) / len;
    double vy = (-vt_face * nx + vn_face * ny) / len;
    double e = p / (rho * (gamma - 1.0)) + 0.5 * (vx * vx + vy * vy);
    double nut = inflow ? d->nut_inf : extrapolate_nut(s, j);
    st[0] = rho;
    st[1] = p;
    st[2] = t;
    st[3] = vx;
    st[4] = vy;
    st[5] = e;
    st[6] = nut;
}

static void far_field_riemann(jo_solver *s)
{
    int j, k;
    for (j = 0; j < s->nj; ++j) {
        double st[7];
        riemann_face(s, j, st);
        double c = sqrt(s->cfg.gamma * s->cfg.r_gas * st[2]);
        double h = st[5] + st[1] / st[0];
        for (k = 0; k < s->h; ++k) {
            int g = s->ni + k;
            size_t a = CI(s, g, j);
            s->rho[a] = st[0];
            s->p[a] = st[1];
            s->t[a] = st[2];
            s->e[a] = st[5];
            s->hh[a] = h;
            s->c[a] = c;
            s->vx[a] = st[3];
            s->vy[a] = st[4];
            s->nut[a] = st[6];
            pack(s, g, j);
        }
    }
}

void jo_boundary_apply(jo_solver *s)
{
    wall_mirror(s);
    far_field_riemann(s);
    periodic_seam(s);
}

/* ---- timestep.rs:24-57 ---- */
double jo_timestep(jo_solver *s)
{
    int ni = s->ni, nj = s->nj, i;
    double cfl = s->cfg.cfl;
#pragma omp parallel for schedule(static)
    for (i = 0; i < ni; ++i) {
        int j;
        for (j = 0; j < nj; ++j) {
            int jp1 = j + 1 < nj ? j + 1 : 0;
            const double *f0 = s->tau + 4 * FI(s, i, j), *f1 = s->tau + 4 * FI(s, i + 1, j);
            const double *g0 = s->nrm + 4 * FI(s, i, j), *g1 = s->nrm + 4 * FI(s, i, jp1);
            double a = 0.5 * (f0[0] + f1[0]);
            double b = 0.5 * (f0[1] + f1[1]);
            double cc = 0.5 * (g0[0] + g1[0]);
            double dd = 0.5 * (g0[1] + g1[1]);
            size_t k = CI(s, i, j);
            double u = s->vx[k], v = s->vy[k];
            double conv = fabs(u * a + v * b) + fabs(u * cc + v * dd);
            double acou = s->c[k] * (sqrt(a * a + b * b) + sqrt(cc * cc + dd * dd));
            double dtc = cfl * s->vol[k] / (conv + acou);
            s->dtc[k] = dtc;
            if (s->cfg.viscous_time_step) {
                /* EXT: the viscous spectral radius of the cell, both directions (Blazek,
                 * Computational Fluid Dynamics, 3rd ed., eqs. 6.21-6.22, with C = 4):
                 * max(4/3, gamma)/rho (mu/Pr + mu_t/Prt) (|A|^2 + |C|^2)/V, mu_t >= 0 */
                double rho = s->rho[k], t = s->t[k];
                double mu = jo_mu(&s->cfg, &s->d, t);
                double rnu = rho * s->nut[k];
                double chi = rnu / mu, c3 = chi * chi * chi;
                double mut = rnu * (c3 / (c3 + s->d.cv1_cubed));
                if (!(mut > 0.0) || (s->cfg.sa_negative && rnu < 0.0)) mut = 0.0;
                double g = s->cfg.gamma > 4.0 / 3.0 ? s->cfg.gamma : 4.0 / 3.0;
                double lv = g / rho * (mu * s->d.inv_pr + mut * s->d.inv_prt) *
                            (a * a + b * b + cc * cc + dd * dd) / s->vol[k];
                s->localdt[k] = cfl * s->vol[k] / (conv + acou + 4.0 * lv);
            } else {
                s->localdt[k] = dtc;
            }
        }
    }
    double m = INFINITY;
    for (i = 0; i < ni; ++i) {
        int j;
        for (j = 0; j < nj; ++j) m = fmin(m, s->localdt[CI(s, i, j)]);
    }
    return m;
}

/* ---- convection.rs:25-80 ---- */
static void euler_flux(const double *fu, const double *face, double gamma, double *out)
{
    double rho = fu[0];
    double inv_rho = 1.0 / rho;
    double u = fu[1] * inv_rho;
    double v = fu[2] * inv_rho;
    double rho_e = fu[3];
    double p = (gamma - 1.0) * (rho_e - rho * (u * u + v * v) * 0.5);
    double vn = face[0] * u + face[1] * v;
    out[0] = rho * vn;
    out[1] = fu[1] * vn + p * face[0];
    out[2] = fu[2] * vn + p * face[1];
    out[3] = (rho_e + p) * vn;
    out[4] = fu[4] * vn;
}

void jo_convection(jo_solver *s)
{
    int ni = s->ni, nj = s->nj, i;
    double gamma = s->cfg.gamma;
#pragma omp parallel for schedule(static)
    for (i = 0; i <= ni; ++i) {
        int j, k;
        for (j = 0; j < nj; ++j) {
            double fu[5];
            const double *a = s->u + 5 * CI(s, i - 1, j), *b = s->u + 5 * CI(s, i, j);
            for (k = 0; k < 5; ++k) fu[k] = 0.5 * (a[k] + b[k]);
            euler_flux(fu, s->tau + 4 * FI(s, i, j), gamma, s->tflux + 5 * FI(s, i, j));
        }
    }
#pragma omp parallel for schedule(static)
    for (i = 0; i < ni; ++i) {
        int j, k;
        for (j = 0; j < nj; ++j) {
            double fu[5];
            const double *a = s->u + 5 * CI(s, i, j - 1), *b = s->u + 5 * CI(s, i, j);
            for (k = 0; k < 5; ++k) fu[k] = 0.5 * (a[k] + b[k]);
            euler_flux(fu, s->nrm + 4 * FI(s, i, j), gamma, s->nflux + 5 * FI(s, i, j));
        }
    }
#pragma omp parallel for schedule(static)
    for (i = 0; i < ni; ++i) {
        int j, k;
        for (j = 0; j < nj; ++j) {
            int jp1 = j + 1 < nj ? j + 1 : 0;
            const double *t1 = s->tflux + 5 * FI(s, i + 1, j), *t0 = s->tflux + 5 * FI(s, i, j);
            const double *n1 = s->nflux + 5 * FI(s, i, jp1), *n0 = s->nflux + 5 * FI(s, i, j);
            double *o = s->fc + 5 * CI(s, i, j);
            for (k = 0; k < 5; ++k) o[k] = t1[k] - t0[k] + n1[k] - n0[k];
        }
    }
}

/* ---- gradient.rs:22-102 ---- */
static void gg(const jo_solver *s, const double *src, int i, int j, const double *up,
               const double *dn, const double *lf, const double *rt, double inv_v, double *gx,
               double *gy)
{
    double c = src[CI(s, i, j)];
    double f_up = 0.5 * (c + src[CI(s, i + 1, j)]);
    double f_dn = 0.5 * (c + src[CI(s, i - 1, j)]);
    double f_lf = 0.5 * (c + src[CI(s, i, j - 1)]);
    double f_rt = 0.5 * (c + src[CI(s, i, j + 1)]);
    *gx = (f_up * up[0] - f_dn * dn[0] + f_rt * rt[0] - f_lf * lf[0]) * inv_v;
    *gy = (f_up * up[1] - f_dn * dn[1] + f_rt * rt[1] - f_lf * lf[1]) * inv_v;
}

void jo_gradient(jo_solver *s)
{
    int ni = s->ni, nj = s->nj, i, j;
#pragma omp parallel for schedule(static)
    for (i = 0; i < ni; ++i) {
        int jj;
        for (jj = 0; jj < nj; ++jj) {
            const double *up = s->tau + 4 * FI(s, i + 1, jj), *dn = s->tau + 4 * FI(s, i, jj);
            const double *lf = s->nrm + 4 * FI(s, i, jj);
            const double *rt = s->nrm + 4 * FI(s, i, jj + 1 < nj ? jj + 1 : 0);
            double inv_v = s->inv_vol[CI(s, i, jj)];
            double *g = s->grad + 8 * CI(s, i, jj);
            gg(s, s->vx, i, jj, up, dn, lf, rt, inv_v, &g[0], &g[1]);
            gg(s, s->vy, i, jj, up, dn, lf, rt, inv_v, &g[2], &g[3]);
            gg(s, s->t, i, jj, up, dn, lf, rt, inv_v, &g[4], &g[5]);
            gg(s, s->nut, i, jj, up, dn, lf, rt, inv_v, &g[6], &g[7]);
        }
    }
    /* fill_ghost_gradients 87-102 */
    for (j = 0; j < nj; ++j) {
        double *w = s->grad + 8 * CI(s, -1, j), *g0 = s->grad + 8 * CI(s, 0, j);
        memcpy(w, g0, 8 * sizeof(double));
        w[4] = 0.0;
        w[5] = 0.0;
        memset(s->grad + 8 * CI(s, ni, j), 0, 8 * sizeof(double));
    }
    for (i = 0; i < ni; ++i) {
        memcpy(s->grad + 8 * CI(s, i, -1), s->grad + 8 * CI(s, i, nj - 1), 8 * sizeof(double));
        memcpy(s->grad + 8 * CI(s, i, nj), s->grad + 8 * CI(s, i, 0), 8 * sizeof(double));
    }
}

/* ---- viscous.rs:27-164 ---- */
static void diffusion_tensor(const jo_solver *s, double t, double rho_nu, double vx, double vy,
                             const double *g, double *aux, double *dt)
{
    const jo_derived *d = &s->d;
    double mu = jo_mu(&s->cfg, d, t);
    double chi = rho_nu / mu;
    double c3 = chi * chi * chi;
    double f1 = c3 / (c3 + d->cv1_cubed);
    double mu_t = rho_nu * f1;
    double nu_coef = rho_nu;
    if (s->cfg.sa_negative && rho_nu < 0.0) {
        /* EXT (ICCFD7-1902 eqs. 20-21): no eddy viscosity, and the nu~ diffusion
         * coefficient mu + rho nu~ fn with fn = (cn1 + chi^3) / (cn1 - chi^3), cn1 = 16.
         * It applies to the wall ghosts too (nu~ = -nu~ of the first cell), where the
         * reference's fv1 has its pole at chi = +cv1 of the first cell. */
        mu_t = 0.0;
        nu_coef = rho_nu * ((16.0 + c3) / (16.0 - c3));
    }
    double mu_eff = mu + mu_t;
    double lam_eff = mu * d->inv_pr + mu_t * d->inv_prt;
    double txx = mu_eff * (4.0 / 3.0 * g[0] - 2.0 / 3.0 * g[3]);
    double tyy = mu_eff * (4.0 / 3.0 * g[3] - 2.0 / 3.0 * g[0]);
    double txy = mu_eff * (g[1] + g[2]);
    double qx = lam_eff * d->cp * g[4];
    double qy = lam_eff * d->cp * g[5];
    double nu_diff = s->cfg.sigma * (mu + nu_coef);
    aux[0] = mu;
    aux[1] = chi;
    aux[2] = f1;
    dt[0] = 0.0;
    dt[1] = txx;
    dt[2] = txy;
    dt[3] = vx * txx + vy * txy + qx;
    dt[4] = nu_diff * g[6];
    dt[5] = 0.0;
    dt[6] = txy;
    dt[7] = tyy;
    dt[8] = vx * txy + vy * tyy + qy;
    dt[9] = nu_diff * g[7];
}

static void eval_tensor(jo_solver *s, int i, int j)
{
    size_t k = CI(s, i, j);
    diffusion_tensor(s, s->t[k], s->u[5 * k + 4], s->vx[k], s->vy[k], s->grad + 8 * k,
                     s->aux + 3 * k, s->diff + 10 * k);
}

static void face_diff(const double *a, const double *b, const double *g, double *out)
{
    int k;
    for (k = 0; k < 5; ++k)
        out[k] = (a[k] + b[k]) * (0.5 * g[0]) + (a[5 + k] + b[5 + k]) * (0.5 * g[1]);
}

void jo_viscous(jo_solver *s)
{
    int ni = s->ni, nj = s->nj, i, j;
#pragma omp parallel for schedule(static)
    for (i = 0; i < ni; ++i) {
        int jj;
        for (jj = 0; jj < nj; ++jj) eval_tensor(s, i, jj);
    }
    for (j = 0; j < nj; ++j) {
        eval_tensor(s, -1, j);
        eval_tensor(s, ni, j);
    }
    for (i = 0; i < ni; ++i) {
        eval_tensor(s, i, -1);
        eval_tensor(s, i, nj);
    }
#pragma omp parallel for schedule(static)
    for (i = 0; i <= ni; ++i) {
        int jj;
        for (jj = 0; jj < nj; ++jj)
            face_diff(s->diff + 10 * CI(s, i - 1, jj), s->diff + 10 * CI(s, i, jj),
                      s->tau + 4 * FI(s, i, jj), s->tdiff + 5 * FI(s, i, jj));
    }
#pragma omp parallel for schedule(static)
    for (i = 0; i < ni; ++i) {
        int jj;
        for (jj = 0; jj < nj; ++jj)
            face_diff(s->diff + 10 * CI(s, i, jj - 1), s->diff + 10 * CI(s, i, jj),
                      s->nrm + 4 * FI(s, i, jj), s->ndiff + 5 * FI(s, i, jj));
    }
#pragma omp parallel for schedule(static)
    for (i = 0; i < ni; ++i) {
        int jj, k;
        for (jj = 0; jj < nj; ++jj) {
            int jp1 = jj + 1 < nj ? jj + 1 : 0;
            const double *t1 = s->tdiff + 5 * FI(s, i + 1, jj), *t0 = s->tdiff + 5 * FI(s, i, jj);
            const double *n1 = s->ndiff + 5 * FI(s, i, jp1), *n0 = s->ndiff + 5 * FI(s, i, jj);
            double *o = s->fv + 5 * CI(s, i, jj);
            for (k = 0; k < 5; ++k) o[k] = t1[k] - t0[k] + n1[k] - n0[k];
        }
    }
}

/* ---- dissipation.rs:29-185 ---- */
static double pressure_sensor(double pm, double p0, double pp)
{
    return fabs((pm - 2.0 * p0 + pp) / (pm + 2.0 * p0 + pp));
}

static void face_diss(const double *um2, const double *um1, const double *u0, const double *up1,
                      double lam, const double *eps, double *out)
{
    int k;
    for (k = 0; k < 5; ++k) {
        double d1 = u0[k] - um1[k];
        double d3 = up1[k] - u0[k] * 3.0 + um1[k] * 3.0 - um2[k];
        out[k] = (d1 * eps[0] - d3 * eps[1]) * lam;
    }
}

void jo_dissipation(jo_solver *s)
{
    int ni = s->ni, nj = s->nj, i, j;
    double cfl = s->cfg.cfl, k2 = s->cfg.k2, k4 = s->cfg.k4;
    /* spectral_radii 44-70 */
#pragma omp parallel for schedule(static)
    for (i = 0; i < ni; ++i) {
        int jj;
        for (jj = 0; jj < nj; ++jj)
            s->spec_ratio[CI(s, i, jj)] = s->vol[CI(s, i, jj)] / s->dtc[CI(s, i, jj)];
    }
#pragma omp parallel for schedule(static)
    for (i = 0; i <= ni; ++i) {
        int a = i < 0 ? 0 : (i > ni - 1 ? ni - 1 : i);
        int b = i - 1 < 0 ? 0 : (i - 1 > ni - 1 ? ni - 1 : i - 1);
        int jj;
        for (jj = 0; jj < nj; ++jj)
            s->tlam[FI(s, i, jj)] =
                0.5 * cfl * (s->spec_ratio[CI(s, a, jj)] + s->spec_ratio[CI(s, b, jj)]);
    }
#pragma omp parallel for schedule(static)
    for (i = 0; i < ni; ++i) {
        int jj;
        for (jj = 0; jj < nj; ++jj) {
            int jm = jj > 0 ? jj - 1 : nj - 1;
            s->nlam[FI(s, i, jj)] =
                0.5 * cfl * (s->spec_ratio[CI(s, i, jj)] + s->spec_ratio[CI(s, i, jm)]);
        }
    }
    /* sensors 73-90 (serial in the reference) */
    for (i = -2; i <= ni + 1; ++i)
        for (j = 0; j < nj; ++j)
            s->sensor_i[CI(s, i, j)] =
                pressure_sensor(s->p[CI(s, i - 1, j)], s->p[CI(s, i, j)], s->p[CI(s, i + 1, j)]);
    for (i = 0; i < ni; ++i)
        for (j = -2; j <= nj; ++j)
            s->sensor_j[CI(s, i, j)] =
                pressure_sensor(s->p[CI(s, i, j - 1)], s->p[CI(s, i, j)], s->p[CI(s, i, j + 1)]);
    /* coefficients 93-126 */
#pragma omp parallel for schedule(static)
    for (i = 0; i <= ni; ++i) {
        int jj;
        for (jj = 0; jj < nj; ++jj) {
            double m = fmax(fmax(fmax(s->sensor_i[CI(s, i - 2, jj)], s->sensor_i[CI(s, i - 1, jj)]),
                                 s->sensor_i[CI(s, i, jj)]),
                            s->sensor_i[CI(s, i + 1, jj)]);
            double e2 = k2 * m;
            s->teps[2 * FI(s, i, jj)] = e2;
            s->teps[2 * FI(s, i, jj) + 1] = fmax(k4 - e2, 0.0);
        }
    }
#pragma omp parallel for schedule(static)
    for (i = 0; i < ni; ++i) {
        int jj;
        for (jj = 0; jj < nj; ++jj) {
            double m = fmax(fmax(fmax(s->sensor_j[CI(s, i, jj - 2)], s->sensor_j[CI(s, i, jj - 1)]),
                                 s->sensor_j[CI(s, i, jj)]),
                            s->sensor_j[CI(s, i, jj + 1)]);
            double e2 = k2 * m;
            s->neps[2 * FI(s, i, jj)] = e2;
            s->neps[2 * FI(s, i, jj) + 1] = fmax(k4 - e2, 0.0);
        }
    }
    /* face_dissipation 129-166 */
#pragma omp parallel for schedule(static)
    for (i = 0; i <= ni; ++i) {
        int jj;
        for (jj = 0; jj < nj; ++jj)
            face_diss(s->u + 5 * CI(s, i - 2, jj), s->u + 5 * CI(s, i - 1, jj), s->u + 5 * CI(s, i, jj),
                      s->u + 5 * CI(s, i + 1, jj), s->tlam[FI(s, i, jj)], s->teps + 2 * FI(s, i, jj),
                      s->tdiss + 5 * FI(s, i, jj));
    }
#pragma omp parallel for schedule(static)
    for (i = 0; i < ni; ++i) {
        int jj;
        for (jj = 0; jj < nj; ++jj)
            face_diss(s->u + 5 * CI(s, i, jj - 2), s->u + 5 * CI(s, i, jj - 1), s->u + 5 * CI(s, i, jj),
                      s->u + 5 * CI(s, i, jj + 1), s->nlam[FI(s, i, jj)], s->neps + 2 * FI(s, i, jj),
                      s->ndiss + 5 * FI(s, i, jj));
    }
    /* assemble 176-185: tau(i+1) + n(j+1) - tau(i) - n(j) */
#pragma omp parallel for schedule(static)
    for (i = 0; i < ni; ++i) {
        int jj, k;
        for (jj = 0; jj < nj; ++jj) {
            int jp1 = jj + 1 < nj ? jj + 1 : 0;
            const double *t1 = s->tdiss + 5 * FI(s, i + 1, jj), *t0 = s->tdiss + 5 * FI(s, i, jj);
            const double *n1 = s->ndiss + 5 * FI(s, i, jp1), *n0 = s->ndiss + 5 * FI(s, i, jj);
            double *o = s->fd + 5 * CI(s, i, jj);
            for (k = 0; k < 5; ++k) o[k] = t1[k] + n1[k] - t0[k] - n0[k];
        }
    }
}

/* ---- source.rs:23-88 ---- */
static double fw_fn(double r, double cw2, double cw3_6, double one_plus_cw3_6)
{
    double g = r + cw2 * (powi_c(r, 6) - r);
    return g * cbrt(sqrt(one_plus_cw3_6 / (powi_c(g, 6) + cw3_6)));
}

void jo_source(jo_solver *s)
{
    int ni = s->ni, nj = s->nj, i;
    const jo_config *sa = &s->cfg;
    const jo_derived *d = &s->d;
    double cb1_over_kappa2 = sa->cb1 * d->inv_kappa2;
    double cb2_sigma = sa->cb2 * sa->sigma;
#pragma omp parallel for schedule(static)
    for (i = 0; i < ni; ++i) {
        int j;
        for (j = 0; j < nj; ++j) {
            size_t k = CI(s, i, j);
            const double *uc = s->u + 5 * k;
            double rho = uc[0], rho_nu = uc[4];
            double nu_tilde = rho_nu / rho;
            double inv_dist2 = s->inv_wd2[k];
            double inv_kd2 = inv_dist2 * d->inv_kappa2;
            const double *a = s->aux + 3 * k;
            const double *g = s->grad + 8 * k;
            double ft2 = sa->ct3 * exp(-sa->ct4 * a[1] * a[1]);
            double fv2 = 1.0 - a[1] / (1.0 + a[1] * a[2]);
            double sv = sa->fv3 * fabs(g[2] - g[1]);
            double s_tilde = fmax(sv + nu_tilde * inv_kd2 * fv2, 1e-10);
            double prod = sa->cb1 * (1.0 - ft2) * s_tilde * rho_nu;
            double r = fmin(nu_tilde / s_tilde * inv_kd2, sa->rmax);
            double destr = (d->cw1 * fw_fn(r, sa->cw2, d->cw3_6, d->one_plus_cw3_6) -
                            cb1_over_kappa2 * ft2) *
                           rho * (nu_tilde * nu_tilde * inv_dist2);
            double diff = cb2_sigma * rho * (g[6] * g[6] + g[7] * g[7]);
            if (sa->sa_negative && rho_nu < 0.0) {
                /* EXT (ICCFD7-1902 eq. 19): production cb1 (1 - ct3) S rho nu~ on the
                 * vorticity S, destruction -cw1 rho (nu~/d)^2 */
                prod = sa->cb1 * (1.0 - sa->ct3) * sv * rho_nu;
                destr = -d->cw1 * rho * (nu_tilde * nu_tilde * inv_dist2);
            }
            s->src[k] = (prod - destr + diff) * s->vol[k];
        }
    }
}

/* ---- solver.rs ---- */
void jo_residual_terms(jo_solver *s)
{
    jo_boundary_apply(s);
    jo_convection(s);
    jo_gradient(s);
    jo_viscous(s);
    jo_dissipation(s);
    jo_source(s);
}

/* solver.rs:123-145; EXT local_time_stepping uses alpha*localdt(i,j) */
static void update(jo_solver *s, double alpha, double dt)
{
    int ni = s->ni, nj = s->nj, i;
    int lts = s->cfg.local_time_stepping;
#pragma omp parallel for schedule(static)
    for (i = 0; i < ni; ++i) {
        int j, k;
        for (j = 0; j < nj; ++j) {
            size_t c = CI(s, i, j);
            double r[5];
            const double *fc = s->fc + 5 * c, *fv = s->fv + 5 * c, *fd = s->fd + 5 * c;
            for (k = 0; k < 5; ++k) r[k] = fc[k] - fv[k] - fd[k];
            r[4] -= s->src[c];
            double a = lts ? alpha * s->localdt[c] : alpha * dt;
            double sc = a * s->inv_vol[c];
            const double *uf = s->u_former + 5 * c;
            double *u = s->u + 5 * c;
            for (k = 0; k < 5; ++k) u[k] = uf[k] - r[k] * sc;
        }
    }
}

/* solver.rs:151-185 */
static int unpack(jo_solver *s, double *err)
{
    double gamma = s->cfg.gamma, r_gas = s->cfg.r_gas;
    int i, j;
    for (i = 0; i < s->ni; ++i)
        for (j = 0; j < s->nj; ++j) {
            size_t k = CI(s, i, j);
            const double *u = s->u + 5 * k;
            double rho = u[0];
            if (isnan(rho) || rho <= 1e-15) {
                if (err) { err[0] = i; err[1] = j; err[2] = rho; err[3] = NAN; }
                return JO_ERR_NONPHYSICAL;
            }
            double inv = 1.0 / rho;
            double vx = u[1] * inv, vy = u[2] * inv;
            double e = u[3] * inv;
            double p = (gamma - 1.0) * (u[3] - rho * (vx * vx + vy * vy) * 0.5);
            if (isnan(p) || p <= 1e-15) {
                if (err) { err[0] = i; err[1] = j; err[2] = rho; err[3] = p; }
                return JO_ERR_NONPHYSICAL;
            }
            double t = p / (r_gas * rho);
            s->rho[k] = rho;
            s->vx[k] = vx;
            s->vy[k] = vy;
            s->nut[k] = u[4] * inv;
            s->e[k] = e;
            s->p[k] = p;
            s->hh[k] = e + p * inv;
            s->t[k] = t;
            s->c[k] = sqrt(r_gas * gamma * t);
        }
    return JO_OK;
}

/* Test hooks: the two halves of advance() (solver.rs:97-120) so that a single RK stage can
 * be observed.  jo_begin_step takes the step-start snapshot and the frozen dt; jo_stage runs
 * one stage (residual terms at the current state, update from the snapshot, unpack). */
double jo_begin_step(jo_solver *s)
{
    memcpy(s->u_former, s->u, sizeof(double) * 5 * s->ncell);
    memcpy(s->rho_prev, s->rho, sizeof(double) * s->ncell);
    return jo_timestep(s);
}

int jo_stage(jo_solver *s, double alpha, double dt, double *err)
{
    jo_residual_terms(s);
    update(s, alpha, dt);
    return unpack(s, err);
}

void jo_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int jo_advance(jo_solver *s, double *err)
{
    static const double RK_ALPHA[5] = {0.25, 1.0 / 6.0, 0.375, 0.5, 1.0}; /* config.rs:217 */
    int st, k;
    memcpy(s->u_former, s->u, sizeof(double) * 5 * s->ncell);
    memcpy(s->rho_prev, s->rho, sizeof(double) * s->ncell);
    double dt = jo_timestep(s);
    s->totaltime += dt;
    for (k = 0; k < 5; ++k) {
        jo_residual_terms(s);
        update(s, RK_ALPHA[k], dt);
        st = unpack(s, err);
        if (st != JO_OK) return st;
    }
    s->step += 1;
    return JO_OK;
}

/* solver.rs:190-200 */
double jo_residual(const jo_solver *s)
{
    double acc = 0.0;
    int i, j;
    for (i = 0; i < s->ni; ++i)
        for (j = 0; j < s->nj; ++j) {
            double dd = s->rho[CI(s, i, j)] - s->rho_prev[CI(s, i, j)];
            acc += dd * dd;
        }
    return sqrt(acc / (double)s->ni / (double)s->nj);
}

/* solver.rs:203-226; EXT residual_drop stops at res < res(step 1) * 10^-drop */
int jo_run(jo_solver *s, long max_steps, double *res_hist, double *out, double *err)
{
    long limit = max_steps > 0 ? max_steps : s->cfg.iteration;
    double target = s->cfg.targetres;
    double residual = NAN;
    long done = 0, k;
    int conv = 0;
    for (k = 1; k <= limit; ++k) {
        int st = jo_advance(s, err);
        if (st != JO_OK) {
            if (out) { out[0] = (double)done; out[1] = residual; out[2] = s->totaltime; out[3] = 0; }
            return st;
        }
        residual = jo_residual(s);
        if (s->step == 1) s->res_first = residual;
        done = k;
        if (res_hist) res_hist[k - 1] = residual;
        conv = residual < target;
        if (s->cfg.residual_drop > 0.0 && s->res_first > 0.0)
            conv = conv || residual < s->res_first * pow(10.0, -s->cfg.residual_drop);
        if (conv) break;
    }
    if (out) {
        out[0] = (double)done;
        out[1] = residual;
        out[2] = s->totaltime;
        out[3] = conv ? 1.0 : 0.0;
    }
    return JO_OK;
}

/* ---- construction (solver.rs:47-62, state.rs:174-190) ---- */
jo_solver *jo_solver_new(const jo_config *cfg, const jo_mesh *mesh)
{
    jo_solver *s = (jo_solver *)calloc(1, sizeof(jo_solver));
    int i, j;
    s->cfg = *cfg;
    jo_config_finish(cfg, &s->d);
    s->ni = mesh->n_rings - 1;
    s->nj = mesh->n_theta;
    s->h = cfg->halo;
    s->stride = s->nj + 2 * s->h;
    s->rows = s->ni + 2 * s->h;
    s->ncell = (size_t)s->rows * (size_t)s->stride;
    size_t nc = s->ncell, nt = (size_t)(s->ni + 1) * s->nj, nn = (size_t)s->ni * s->nj;
    s->vol = zalloc(nc); s->cx = zalloc(nc); s->cy = zalloc(nc); s->wd = zalloc(nc);
    s->inv_vol = zalloc(nc); s->inv_wd2 = zalloc(nc);
    s->tau = zalloc(4 * nt); s->nrm = zalloc(4 * nn);
    s->u = zalloc(5 * nc); s->u_former = zalloc(5 * nc);
    s->rho = zalloc(nc); s->p = zalloc(nc); s->t = zalloc(nc); s->vx = zalloc(nc);
    s->vy = zalloc(nc); s->e = zalloc(nc); s->hh = zalloc(nc); s->c = zalloc(nc);
    s->nut = zalloc(nc);
    s->grad = zalloc(8 * nc); s->aux = zalloc(3 * nc); s->diff = zalloc(10 * nc);
    s->fc = zalloc(5 * nc); s->fv = zalloc(5 * nc); s->fd = zalloc(5 * nc);
    s->src = zalloc(nc); s->localdt = zalloc(nc); s->dtc = zalloc(nc);
    s->tflux = zalloc(5 * nt); s->tdiff = zalloc(5 * nt); s->tdiss = zalloc(5 * nt);
    s->tlam = zalloc(nt); s->teps = zalloc(2 * nt);
    s->nflux = zalloc(5 * nn); s->ndiff = zalloc(5 * nn); s->ndiss = zalloc(5 * nn);
    s->nlam = zalloc(nn); s->neps = zalloc(2 * nn);
    s->sensor_i = zalloc(nc); s->sensor_j = zalloc(nc); s->spec_ratio = zalloc(nc);
    s->rho_prev = zalloc(nc);
    build_geometry(s, mesh);
    /* Cells::initialize */
    {
        const jo_derived *d = &s->d;
        double rho = cfg->p_inf / (cfg->r_gas * cfg->t_inf);
        double mu = jo_mu(cfg, d, cfg->t_inf);
        for (i = 0; i < s->ni; ++i)
            for (j = 0; j < s->nj; ++j) {
                write_prim(s, i, j, rho, d->u_inf, d->v_inf, cfg->p_inf, d->nut_inf);
                double *a = s->aux + 3 * CI(s, i, j);
                a[0] = mu;
                a[1] = 0.0;
                a[2] = 0.0;
            }
    }
    jo_boundary_apply(s);
    return s;
}

void jo_solver_free(jo_solver *s)
{
    if (!s) return;
    double *all[] = {s->vol, s->cx, s->cy, s->wd, s->inv_vol, s->inv_wd2, s->tau, s->nrm,
                     s->u, s->u_former, s->rho, s->p, s->t, s->vx, s->vy, s->e, s->hh, s->c,
                     s->nut, s->grad, s->aux, s->diff, s->fc, s->fv, s->fd, s->src, s->localdt,
                     s->tflux, s->tdiff, s->tdiss, s->tlam, s->teps, s->nflux, s->ndiff,
                     s->ndiss, s->nlam, s->neps, s->sensor_i, s->sensor_j, s->spec_ratio,
                     s->rho_prev, s->dtc};
    size_t k;
    for (k = 0; k < sizeof all / sizeof all[0]; ++k) free(all[k]);
    free(s);
}

int jo_ni(const jo_solver *s) { return s->ni; }
int jo_nj(const jo_solver *s) { return s->nj; }
int jo_halo(const jo_solver *s) { return s->h; }
double jo_totaltime(const jo_solver *s) { return s->totaltime; }
long jo_step(const jo_solver *s) { return s->step; }

/* state.rs:198-203 Cells::set_uniform: every addressable cell incl. halo and corners */
void jo_set_uniform(jo_solver *s, double rho, double vx, double vy, double p, double nut)
{
    int i, j;
    for (i = -s->h; i < s->ni + s->h; ++i)
        for (j = -s->h; j < s->nj + s->h; ++j) write_prim(s, i, j, rho, vx, vy, p, nut);
}

int jo_set_state(jo_solver *s, const double *u_planes, double *err)
{
    int i, j, k;
    size_t n = (size_t)s->ni * s->nj;
    for (i = 0; i < s->ni; ++i)
        for (j = 0; j < s->nj; ++j)
            for (k = 0; k < 5; ++k)
                s->u[5 * CI(s, i, j) + k] = u_planes[k * n + (size_t)i * s->nj + j];
    return unpack(s, err);
}

/* EXT (not in the reference): wall force from the wall tau-face momentum fluxes,
 * F_body = -sum_j (Fc_tau(0,j) - Fv_tau(0,j))[mx,my], summed in j order; JST dissipation is
 * excluded.  CD/CL rotate by the angle of attack over q_inf * L. */
void jo_forces(jo_solver *s, double *out4)
{
    int j;
    double fx = 0.0, fy = 0.0;
    jo_boundary_apply(s);
    jo_convection(s);
    jo_gradient(s);
    jo_viscous(s);
    for (j = 0; j < s->nj; ++j) {
        const double *f = s->tflux + 5 * FI(s, 0, j), *v = s->tdiff + 5 * FI(s, 0, j);
        fx -= f[1] - v[1];
        fy -= f[2] - v[2];
    }
    double aoa = s->cfg.aoa_deg * (PI_D / 180.0);
    double ca = cos(aoa), sa = sin(aoa);
    double q = 0.5 * s->d.rho_inf * (s->d.u_inf * s->d.u_inf + s->d.v_inf * s->d.v_inf);
    double ql = q * s->cfg.ref_length;
    out4[0] = fx;
    out4[1] = fy;
    out4[2] = (-fx * sa + fy * ca) / ql;
    out4[3] = (fx * ca + fy * sa) / ql;
}

/* EXT (not in the reference): wall distribution over the wall tau faces (0, j), j in
 * [0, NJ), from the same face fluxes as jo_forces.  n = tau normal of the face (into the
 * fluid, |n| = edge length), s = (-n_y, n_x)/|n| its unit tangent in the direction of
 * increasing j.  x, y = face midpoint; p_w = (Fc_x n_x + Fc_y n_y)/|n|^2, the pressure the
 * face's convective momentum flux carries (the mirrored ghost cancels the face momentum,
 * so Fc[mx,my] = p_w n); Cp = (p_w - p_inf)/q_inf; Cf = (Fv[mx,my] . s)/(|n| q_inf), the
 * tangential traction the fluid exerts on the body. */
void jo_wall(jo_solver *s, double *x, double *y, double *cp, double *cf)
{
    int j;
    jo_boundary_apply(s);
    jo_convection(s);
    jo_gradient(s);
    jo_viscous(s);
    {
        double q = 0.5 * s->d.rho_inf * (s->d.u_inf * s->d.u_inf + s->d.v_inf * s->d.v_inf);
        for (j = 0; j < s->nj; ++j) {
            const double *g = s->tau + 4 * FI(s, 0, j);
            const double *f = s->tflux + 5 * FI(s, 0, j), *v = s->tdiff + 5 * FI(s, 0, j);
            const double n2 = g[0] * g[0] + g[1] * g[1];
            const double pw = (f[1] * g[0] + f[2] * g[1]) / n2;
            const double tau = (v[1] * (-g[1]) + v[2] * g[0]) / n2;
            x[j] = g[2];
            y[j] = g[3];
            cp[j] = (pw - s->cfg.p_inf) / q;
            cf[j] = tau / q;
        }
    }
}

/* ---- field export ---- */
static long copy_cells(const jo_solver *s, const double *f, int ncomp, int comp, double *dst)
{
    int i, j;
    long n = 0;
    for (i = 0; i < s->ni; ++i)
        for (j = 0; j < s->nj; ++j) dst[n++] = f[(size_t)ncomp * CI(s, i, j) + comp];
    return n;
}

static long copy_faces(const double *f, int nrows, int nj, int ncomp, int comp, double *dst)
{
    long n = 0, k;
    for (k = 0; k < (long)nrows * nj; ++k) dst[n++] = f[(size_t)ncomp * k + comp];
    return n;
}

long jo_get(const jo_solver *s, const char *name, double *dst)
{
    struct { const char *n; const double *f; int nc; } cells[] = {
        {"rho", s->rho, 1}, {"p", s->p, 1}, {"T", s->t, 1}, {"u", s->vx, 1}, {"v", s->vy, 1},
        {"E", s->e, 1}, {"H", s->hh, 1}, {"c", s->c, 1}, {"nut", s->nut, 1},
        {"S", s->src, 1}, {"localdt", s->localdt, 1}, {"vol", s->vol, 1}, {"cx", s->cx, 1},
        {"cy", s->cy, 1}, {"wall_dist", s->wd, 1}, {"inv_vol", s->inv_vol, 1},
        {"inv_wd2", s->inv_wd2, 1}, {"spec_ratio", s->spec_ratio, 1}};
    struct { const char *n; const double *f; int nc; } multi[] = {
        {"U", s->u, 5}, {"Fc", s->fc, 5}, {"Fv", s->fv, 5}, {"Fd", s->fd, 5}, {"grad", s->grad, 8}};
    size_t k;
    for (k = 0; k < sizeof cells / sizeof cells[0]; ++k)
        if (strcmp(name, cells[k].n) == 0) return copy_cells(s, cells[k].f, 1, 0, dst);
    if (strcmp(name, "mu") == 0) return copy_cells(s, s->aux, 3, 0, dst);
    if (strcmp(name, "chi") == 0) return copy_cells(s, s->aux, 3, 1, dst);
    if (strcmp(name, "fv1") == 0) return copy_cells(s, s->aux, 3, 2, dst);
    for (k = 0; k < sizeof multi / sizeof multi[0]; ++k) {
        size_t l = strlen(multi[k].n);
        if (strncmp(name, multi[k].n, l) == 0 && isdigit((unsigned char)name[l]) && name[l + 1] == 0) {
            int c = name[l] - '0';
            if (c >= multi[k].nc) return -1;
            return copy_cells(s, multi[k].f, multi[k].nc, c, dst);
        }
    }
    {
        static const char *gk[] = {"nx", "ny", "mx", "my"};
        int c;
        for (c = 0; c < 4; ++c) {
            char buf[16];
            snprintf(buf, sizeof buf, "tau_%s", gk[c]);
            if (strcmp(name, buf) == 0) return copy_faces(s->tau, s->ni + 1, s->nj, 4, c, dst);
            snprintf(buf, sizeof buf, "n_%s", gk[c]);
            if (strcmp(name, buf) == 0) return copy_faces(s->nrm, s->ni, s->nj, 4, c, dst);
        }
        for (c = 0; c < 5; ++c) {
            char buf[16];
            snprintf(buf, sizeof buf, "tau_flux%d", c);
            if (strcmp(name, buf) == 0) return copy_faces(s->tflux, s->ni + 1, s->nj, 5, c, dst);
            snprintf(buf, sizeof buf, "tau_diff%d", c);
            if (strcmp(name, buf) == 0) return copy_faces(s->tdiff, s->ni + 1, s->nj, 5, c, dst);
            snprintf(buf, sizeof buf, "n_flux%d", c);
            if (strcmp(name, buf) == 0) return copy_faces(s->nflux, s->ni, s->nj, 5, c, dst);
        }
    }
    return -1;
}
