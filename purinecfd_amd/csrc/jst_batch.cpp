// jst_batch.cpp — batched FP32-mixed driver and the serving loop (include/jst_gpu.h).
//
// A batch keeps every admitted case resident in one HBM arena and advances all running cases
// with one launch per RK stage plus a small step-end launch: the stage kernel walks a tile list
// spanning every case, so small and large meshes share the GPU.  One step of the whole batch
// is captured once into a HIP graph and replayed; the case table, tile list and per-case
// control blocks live in device memory, so the graph stays valid while cases advance,
// converge or fail independently.  Membership changes rebuild the lists and the graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/jst_gpu.h"
#include "jst_common.h"
#include "jst_dev.h"
#include "jst_launch.h"

using namespace jsth;

#define HIPCHK(x)                                                                            \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess)                                                                \
            return fail(JST_ERR_HIP, std::string("HIP error: ") + hipGetErrorString(e_) +    \
                                         " at " #x);                                         \
    } while (0)

namespace {

const double RK_ALPHA[5] = {0.25, 1.0 / 6.0, 0.375, 0.5, 1.0}; // jst-rs config.rs:217
const int GRAPH_STEPS = 16;

inline size_t r256(size_t n) { return (n + 255) & ~(size_t)255; }

// First-fit arena over one device allocation; blocks are whole cases.
struct Arena {
    char *base = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> free_; // offset -> length
    long long alloc(size_t n)
    {
        n = r256(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                size_t off = it->first, len = it->second;
                free_.erase(it);
                if (len > n) free_[off + n] = len - n;
                return (long long)off;
            }
        }
        return -1;
    }
    void release(size_t off, size_t n)
    {
        n = r256(n);
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

// Byte layout of one case inside its arena block.
struct CaseLayout {
    size_t u64[2], sb, s32[2], g64[14], tau2, cell4, ratio[2], localdt[2], rpart, dpart, res_hist, posts, scratch, dbg;
    size_t total;
};

CaseLayout layout_for(int ni, int nj, int pitch, int npost, int res_cap)
{
    CaseLayout L{};
    const size_t plane = (size_t)ni * pitch, fplane = (size_t)(ni + 1) * pitch;
    size_t o = 0;
    auto carve = [&](size_t b) {
        size_t a = o;
        o += r256(b);
        return a;
    };
    for (int b = 0; b < 2; ++b) L.u64[b] = carve(5 * plane * 8);
    L.sb = carve(5 * plane * 4);
    for (int b = 0; b < 2; ++b) L.s32[b] = carve(5 * plane * 4);
    // FP64 geometry: vol cx cy wd inv_vol inv_wd2 (plane) | tnx tny tmx tmy (fplane) | nnx nny nmx nmy
    for (int q = 0; q < 6; ++q) L.g64[q] = carve(plane * 8);
    for (int q = 6; q < 10; ++q) L.g64[q] = carve(fplane * 8);
    for (int q = 10; q < 14; ++q) L.g64[q] = carve(plane * 8);
    // FP32 geometry: tau normals (nx, ny) per face (fplane pairs), then (n-normal nx, ny, V,
    // 1/d^2) per cell (plane quads), then V/dt and the local dt.  The stage kernels reach all
    // of them through one buffer descriptor that starts at tau2 and ends with localdt[1]:
    // keep tau2 first and localdt[1] last.
    L.tau2 = carve(fplane * 8);
    L.cell4 = carve(plane * 16);
    for (int b = 0; b < 2; ++b) L.ratio[b] = carve(plane * 4);
    for (int b = 0; b < 2; ++b) L.localdt[b] = carve(plane * 4);
    // one partial per (row segment, column group): at most ni segments (1-row segments)
    const size_t nparts = (size_t)ni * ((nj + MIX_COLS - 1) / MIX_COLS);
    L.rpart = carve(nparts * 8);
    L.dpart = carve(nparts * 8);
    L.res_hist = carve((size_t)res_cap * 8);
    L.posts = carve((size_t)npost * sizeof(PostDesc));
    L.scratch = carve((4 * (size_t)nj + 2) * 8); // wall-force terms [4][nj] + 2
    L.dbg = carve(6 * plane * 4);
    L.total = o;
    return L;
}

struct HostCase {
    int slot = -1;
    Config cfg{};
    Derived d{};
    int ni = 0, nj = 0, pitch = 0, npost = 0;
    size_t plane = 0, fplane = 0, off = 0, bytes = 0;
    long max_steps = 0;
    CaseMix dev{};
    jst::GeomPtrs g64{};
    PostDesc *d_posts = nullptr;
    double *scratch = nullptr;
    std::vector<PostDesc> posts;
};

MixConst make_mix_const(const Config &c, const Derived &d)
{
    MixConst m{};
    m.k = make_const(c, d);
    const double rho = d.rho_inf, u = d.u_inf, v = d.v_inf, p = c.p_inf;
    const double e = p / (rho * (c.gamma - 1.0)) + 0.5 * (u * u + v * v);
    m.u64[0] = rho;
    m.u64[1] = rho * u;
    m.u64[2] = rho * v;
    m.u64[3] = rho * e;
    m.u64[4] = rho * d.nut_inf;
    m.gamma = (float)c.gamma;
    m.gm1 = (float)(c.gamma - 1.0);
    m.r_gas = (float)c.r_gas;
    m.cfl = (float)c.cfl;
    m.inv_cfl = (float)(1.0 / c.cfl);
    m.cfl_half = (float)(0.5 * c.cfl);
    m.k2 = (float)c.k2;
    m.k4 = (float)c.k4;
    m.rho_inf = (float)rho;
    m.u_inf = (float)u;
    m.v_inf = (float)v;
    m.p_inf = (float)p;
    m.t_inf = (float)c.t_inf;
    m.m_inf = (float)m.u64[1];
    m.n_inf = (float)m.u64[2];
    m.e_inf = (float)m.u64[3];
    m.h_inf = (float)(m.u64[3] + p);
    m.rnu_inf = (float)m.u64[4];
    m.nut_inf = (float)d.nut_inf;
    m.suth_num = (float)d.suth_num;
    m.inv_t_ref = (float)d.inv_t_ref;
    m.t_suth = (float)c.t_suth;
    m.inv_pr = (float)d.inv_pr;
    m.inv_prt = (float)d.inv_prt;
    m.cp = (float)d.cp;
    m.sigma = (float)c.sigma;
    m.cv1_cubed = (float)d.cv1_cubed;
    m.cb1 = (float)c.cb1;
    m.cb2 = (float)c.cb2;
    m.ct3 = (float)c.ct3;
    m.ct4 = (float)c.ct4;
    m.cw1 = (float)d.cw1;
    m.cw2 = (float)c.cw2;
    m.cw3_6 = (float)d.cw3_6;
    m.one_plus_cw3_6 = (float)d.one_plus_cw3_6;
    m.fv3 = (float)c.fv3;
    m.inv_kappa2 = (float)d.inv_kappa2;
    m.rmax = (float)c.rmax;
    m.lts = c.lts ? 1 : 0;
    m.visc_dt = c.visc_dt ? 1 : 0;
    m.sa_neg = c.sa_neg ? 1 : 0;
    return m;
}

} // namespace

struct jst_batch {
    jst_precision prec = JST_MIXED;
    hipStream_t stream = nullptr;
    Arena arena;
    int max_cases = 0;
    std::vector<std::unique_ptr<HostCase>> cases; // by slot
    CaseMix *d_cases = nullptr;
    CaseCtl *d_ctl = nullptr;
    CaseCtl *h_ctl = nullptr; // pinned mirror
    TileDesc *d_tiles = nullptr;
    size_t tiles_cap = 0;
    int ntiles = 0;
    int *d_slots = nullptr; // slots of the batch's cases (one step-end block each)
    int nslots = 0;
    bool neg = false;       // some case of the lists runs the negative-nu~ SA branch
    int variant = 0, grid = 0;
    int fit_rows = 24; // segment length under MIX_VARIANT_FIT, refitted by rebuild_lists
    bool dirty = true;
    hipGraph_t graph = nullptr;      // one step
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph_k = nullptr;    // GRAPH_STEPS steps (host launch cost amortised)
    hipGraphExec_t gexec_k = nullptr;
    bool use_graph = true;
    int res_cap = 4096;
    double *d_scratch_out = nullptr;
    CaseDev *d_view = nullptr; // FP64 view of one case for the wall-force kernel
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::vector<int> timing_slots; // cases whose stops jst_batch_time_begin switched off
    int device = 0;                // the HIP device the batch was created on
    TileDesc *noalloc_items = nullptr; // that device's fixed work-item array (no allocation)

    ~jst_batch()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (gexec_k) (void)hipGraphExecDestroy(gexec_k);
        if (graph_k) (void)hipGraphDestroy(graph_k);
        for (void *p : {(void *)arena.base, (void *)d_cases, (void *)d_ctl, (void *)d_tiles,
                        (void *)d_slots, (void *)d_scratch_out, (void *)d_view})
            if (p) (void)hipFree(p);
        if (h_ctl) (void)hipHostFree(h_ctl);
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

jst_status drop_graph(jst_batch *b)
{
    if (b->gexec) HIPCHK(hipGraphExecDestroy(b->gexec));
    if (b->graph) HIPCHK(hipGraphDestroy(b->graph));
    if (b->gexec_k) HIPCHK(hipGraphExecDestroy(b->gexec_k));
    if (b->graph_k) HIPCHK(hipGraphDestroy(b->graph_k));
    b->gexec = nullptr;
    b->graph = nullptr;
    b->gexec_k = nullptr;
    b->graph_k = nullptr;
    return JST_OK;
}


// A case's own row-segment length, from its size alone: the shortest segment whose work items
// (58-column groups x segments) all fit in the 2048 waves the GPU keeps resident at once (two
// per SIMD).  Short segments give a case the most waves (lowest latency while the GPU is not
// full); a longer one would only re-read fewer halo rows, a shorter one leaves waves queued
// behind the resident ones (the 1024 x 512-cell case: 172 / 147 / 185 us per step at 8 / 5 /
// 4 rows).
int own_segment_rows(int ni, int nj)
{
    const long groups = (nj + MIX_COLS - 1) / MIX_COLS;
    for (int rows = 1; rows < MIX_MAX_ROWS; ++rows)
        if (groups * ((ni + rows - 1) / rows) <= 2048) return rows;
    return MIX_MAX_ROWS;
}

// One segment length for the whole batch: the one that minimises the estimated launch time,
// in row steps of the busiest wave, ceil(items / 2048) x (L + 6) (the items are dealt round-
// robin over the 2048 resident waves, and an item of L rows takes L + 6 row steps: 3 to fill
// the row pipeline, 3 to drain it); ties go to the shorter length.  This is the shortest
// length whose items all fit in one round when such a length exists: the 64-case NACA batch gets 37 rows (2019 items), against 526 / 540 / 545 us
// per step at 37 / 38 / 40 rows and 700 us at 36 rows (2172 items, a second round).
int fit_segment_rows(const jst_batch *b, double *row_steps = nullptr)
{
    int best = 24;
    double best_t = 1e300;
    for (int rows = 1; rows <= MIX_MAX_ROWS; ++rows) {
        long items = 0;
        for (auto &hc : b->cases)
            if (hc) items += (long)((hc->nj + MIX_COLS - 1) / MIX_COLS) * ((hc->ni + rows - 1) / rows);
        if (items == 0) break;
        const double t = (double)((items + 2047) / 2048) * (rows + 6);
        if (t < best_t) {
            best_t = t;
            best = rows;
        }
    }
    if (row_steps) *row_steps = best_t;
    return best;
}

// Row-segment length of a case under the batch's tiling: the variant's length, under
// MIX_VARIANT_PER_CASE the case's own, which depends on nothing else in the batch, or under
// MIX_VARIANT_FIT the batch's fitted length as of the last list rebuild.
int case_rows(const jst_batch *b, const HostCase &hc);

int case_rows(const jst_batch *b, const HostCase &hc)
{
    if (b->variant >= MIX_VARIANT_ROWS) return b->variant - MIX_VARIANT_ROWS;
    if (b->variant == MIX_VARIANT_FIT) return b->fit_rows;
    return b->variant < MIX_NVARIANTS ? jst::MIX_SEGMENT_ROWS[b->variant] : own_segment_rows(hc.ni, hc.nj);
}

// Work items of one case for the row march: row segments of `rows` rows x 58-column groups,
// the column groups of a segment consecutive.  Four consecutive items share a workgroup and
// march the same rows at the same time, so the three halo columns each wave reads on either
// side are its neighbours' own columns, loaded at the same moment (L2 hits): against column
// groups outermost, 2-3 % less time per inner and stage-5 launch of the 64-case batch.
void case_items(const HostCase &hc, int rows, std::vector<TileDesc> &out)
{
    for (int i0 = 0; i0 < hc.ni; i0 += rows)
        for (int j0 = 0; j0 < hc.nj; j0 += MIX_COLS)
            out.push_back(TileDesc{hc.slot, i0, j0, std::min(i0 + rows, hc.ni)});
}

// Rebuild the work-item list and the step-end slot list.  XCD-aware: workgroup w (four
// consecutive items) runs on XCD w % 8, so runs of consecutive items of a case - neighbouring
// column groups and row segments, which share halo columns and rows - are dealt to one XCD.
jst_status rebuild_lists(jst_batch *b)
{
    std::vector<TileDesc> nat;
    std::vector<int> slots;
    bool neg = false;
    if (b->variant == MIX_VARIANT_FIT) b->fit_rows = fit_segment_rows(b);
    for (auto &hc : b->cases) {
        if (!hc) continue;
        const int rows = case_rows(b, *hc);
        case_items(*hc, rows, nat);
        slots.push_back(hc->slot);
        neg = neg || hc->cfg.sa_neg;
        // stage 5 leaves one step-end partial per (segment, column group): the segment length
        // is read by stage 5 and the step end
        if (hc->dev.seg_rows != rows) {
            hc->dev.seg_rows = rows;
            HIPCHK(hipMemcpyAsync(&b->d_cases[hc->slot].seg_rows, &hc->dev.seg_rows, sizeof(int),
                                  hipMemcpyHostToDevice, b->stream));
        }
    }
    const int nu = ((int)nat.size() + 3) / 4, per = (nu + 7) / 8; // units of 4 items
    std::vector<TileDesc> tiles;
    tiles.reserve(nat.size());
    for (int q = 0; q < per; ++q)
        for (int x = 0; x < 8; ++x) {
            const int u = x * per + q;
            for (int e = 0; e < 4 && u < nu && 4 * u + e < (int)nat.size(); ++e) tiles.push_back(nat[4 * u + e]);
        }
    if (tiles.size() > b->tiles_cap) {
        if (b->d_tiles) HIPCHK(hipFree(b->d_tiles));
        b->tiles_cap = std::max<size_t>(tiles.size() * 2, 1024);
        HIPCHK(hipMalloc(&b->d_tiles, b->tiles_cap * sizeof(TileDesc)));
    }
    if (!tiles.empty())
        HIPCHK(hipMemcpyAsync(b->d_tiles, tiles.data(), tiles.size() * sizeof(TileDesc),
                              hipMemcpyHostToDevice, b->stream));
    if (!slots.empty())
        HIPCHK(hipMemcpyAsync(b->d_slots, slots.data(), slots.size() * sizeof(int),
                              hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    b->ntiles = (int)tiles.size();
    b->nslots = (int)slots.size();
    b->neg = neg;
    jst_status st = drop_graph(b);
    if (st != JST_OK) return st;
    b->dirty = false;
    return JST_OK;
}

// One RK step of every running case: five stage launches (the fifth also writes the step-end
// partials) and the one-block-per-case step end.
jst_status launch_one_step(jst_batch *b)
{
    for (int k = 0; k < 5; ++k)
        HIPCHK(jst::launch_stage_mixed(b->d_cases, b->d_tiles, b->ntiles, b->grid, k, RK_ALPHA[k], false,
                                       b->neg, b->stream));
    HIPCHK(jst::launch_end_mixed(b->d_cases, b->d_slots, b->nslots, b->stream));
    return JST_OK;
}

// Capture `steps` consecutive steps of the whole batch into one executable graph.
jst_status capture_steps(jst_batch *b, int steps, hipGraph_t *g, hipGraphExec_t *x)
{
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    jst_status st = JST_OK;
    for (int q = 0; q < steps && st == JST_OK; ++q) st = launch_one_step(b);
    hipGraph_t gg = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &gg);
    if (st != JST_OK) return st;
    HIPCHK(e);
    *g = gg;
    HIPCHK(hipGraphInstantiate(x, gg, nullptr, nullptr, 0));
    return JST_OK;
}

// Capture (without running) the graphs that queueing nsteps steps will use.
jst_status prepare_graphs(jst_batch *b, int nsteps)
{
    jst_status st;
    if (!b->use_graph) return JST_OK;
    if (nsteps >= GRAPH_STEPS && !b->gexec_k &&
        (st = capture_steps(b, GRAPH_STEPS, &b->graph_k, &b->gexec_k)) != JST_OK)
        return st;
    if (nsteps % GRAPH_STEPS && !b->gexec && (st = capture_steps(b, 1, &b->graph, &b->gexec)) != JST_OK)
        return st;
    return JST_OK;
}

// Queue nsteps steps: graphs of GRAPH_STEPS steps, then single-step graphs for the rest.
// Steps past a case's stop are no-ops on the device, so queueing never overshoots a case.
jst_status launch_steps(jst_batch *b, int nsteps)
{
    if (b->dirty) {
        jst_status st = rebuild_lists(b);
        if (st != JST_OK) return st;
    }
    if (b->ntiles == 0 || nsteps <= 0) return JST_OK;
    if (!b->use_graph) {
        for (int s = 0; s < nsteps; ++s) {
            jst_status st = launch_one_step(b);
            if (st != JST_OK) return st;
        }
        return JST_OK;
    }
    jst_status st = prepare_graphs(b, nsteps);
    if (st != JST_OK) return st;
    for (int s = 0; s < nsteps / GRAPH_STEPS; ++s) HIPCHK(hipGraphLaunch(b->gexec_k, b->stream));
    for (int s = 0; s < nsteps % GRAPH_STEPS; ++s) HIPCHK(hipGraphLaunch(b->gexec, b->stream));
    return JST_OK;
}

jst_status fetch_ctl(jst_batch *b)
{
    HIPCHK(hipMemcpyAsync(b->h_ctl, b->d_ctl, sizeof(CaseCtl) * (size_t)b->max_cases,
                          hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return JST_OK;
}

template <class T> jst_status set_ctl_field(jst_batch *b, int slot, T CaseCtl::*field, T value)
{
    CaseCtl *h = &b->h_ctl[slot];
    h->*field = value;
    const size_t off = (size_t)((char *)&(h->*field) - (char *)h);
    HIPCHK(hipMemcpyAsync((char *)&b->d_ctl[slot] + off, &(h->*field), sizeof(T),
                          hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return JST_OK;
}

// The same field of several cases: the copies are queued together and waited for once.
template <class T>
jst_status set_ctl_fields(jst_batch *b, const std::vector<int> &slots, T CaseCtl::*field, T value)
{
    for (int slot : slots) {
        CaseCtl *h = &b->h_ctl[slot];
        h->*field = value;
        const size_t off = (size_t)((char *)&(h->*field) - (char *)h);
        HIPCHK(hipMemcpyAsync((char *)&b->d_ctl[slot] + off, &(h->*field), sizeof(T),
                              hipMemcpyHostToDevice, b->stream));
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    return JST_OK;
}

HostCase *get_case(const jst_batch *b, int id)
{
    if (!b || id < 0 || id >= (int)b->cases.size()) return nullptr;
    return b->cases[id].get();
}

jst_status download(jst_batch *b, const void *d, size_t bytes, void *h)
{
    HIPCHK(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return JST_OK;
}

} // namespace


namespace jsth {

// The FP64 wall-force kernel on an FP64 view of one case (its U^n buffers and FP64
// geometry): per-face wall terms into the case's scratch, {Fx, Fy} into d_scratch_out.
jst_status launch_wall_terms(jst_batch *b, HostCase *hc, int case_id)
{
    CaseDev v{};
    v.ni = hc->ni;
    v.nj = hc->nj;
    v.pitch = hc->pitch;
    v.ubuf[0] = hc->dev.u64[0];
    v.ubuf[1] = hc->dev.u64[1];
    v.tnx = hc->g64.tnx;
    v.tny = hc->g64.tny;
    v.tmx = hc->g64.tmx;
    v.tmy = hc->g64.tmy;
    v.nnx = hc->g64.nnx;
    v.nny = hc->g64.nny;
    v.vol = hc->g64.vol;
    v.inv_vol = hc->g64.inv_vol;
    v.inv_wd2 = hc->g64.inv_wd2;
    v.cx = hc->g64.cx;
    v.cy = hc->g64.cy;
    v.ctl = b->d_ctl + case_id;
    v.k = hc->dev.k.k;
    HIPCHK(hipMemcpyAsync(b->d_view, &v, sizeof(CaseDev), hipMemcpyHostToDevice, b->stream));
    HIPCHK(jst::launch_forces_fp64(b->d_view, 0, hc->scratch, b->d_scratch_out, b->stream));
    return JST_OK;
}

// Wall force {Fx, Fy, CL, CD} of one case: forces are an FP64 reduction.
jst_status batch_forces(jst_batch *b, int case_id, double out[4])
{
    HostCase *hc = get_case(b, case_id);
    if (!hc) return fail(JST_ERR_ARG, "no such case");
    jst_status st = launch_wall_terms(b, hc, case_id);
    if (st != JST_OK) return st;
    double f[2];
    st = download(b, b->d_scratch_out, sizeof f, f);
    if (st != JST_OK) return st;
    const Config &c = hc->cfg;
    const double aoa = c.aoa * (M_PI / 180.0), ca = std::cos(aoa), sa = std::sin(aoa);
    const double q = 0.5 * hc->d.rho_inf * (hc->d.u_inf * hc->d.u_inf + hc->d.v_inf * hc->d.v_inf);
    const double ql = q * c.ref_length;
    out[0] = f[0];
    out[1] = f[1];
    out[2] = (-f[0] * sa + f[1] * ca) / ql;
    out[3] = (f[0] * ca + f[1] * sa) / ql;
    return JST_OK;
}

// Wall distribution of one case from the same FP64 wall terms as its forces.
jst_status batch_wall(jst_batch *b, int case_id, double *x, double *y, double *cp, double *cf)
{
    HostCase *hc = get_case(b, case_id);
    if (!hc) return fail(JST_ERR_ARG, "no such case");
    jst_status st = launch_wall_terms(b, hc, case_id);
    if (st != JST_OK) return st;
    return wall_distribution(hc->scratch, hc->g64, hc->nj, hc->cfg, hc->d, b->stream, x, y, cp,
                             cf);
}

// Solver::run (jst-rs solver.rs:203-226) for one case of a batch: the device stops by itself
// on convergence or after `limit` steps (unless no_stop, the advance() semantics); residuals
// are read back per synchronised batch of steps and fed to cb in order.
jst_status batch_run_case(jst_batch *b, int id, long limit, jst_step_cb cb, void *user,
                          jst_report *out, bool no_stop)
{
    HostCase *hc = get_case(b, id);
    if (!hc) return fail(JST_ERR_ARG, "no such case");
    jst_status st = fetch_ctl(b);
    if (st != JST_OK) return st;
    if (b->h_ctl[id].err_key != JST_ERR_NONE) return fail(JST_ERR_NONPHYSICAL, "non-physical state");
    const long long start = b->h_ctl[id].step;
    if ((st = set_ctl_field(b, id, &CaseCtl::timing, no_stop ? 1 : 0)) != JST_OK) return st;
    if ((st = set_ctl_field(b, id, &CaseCtl::stop_step, (long long)(start + limit))) != JST_OK) return st;
    if ((st = set_ctl_field(b, id, &CaseCtl::status, (int)JST_RUN_ACTIVE)) != JST_OK) return st;
    long done = 0;
    int batch = 8;
    std::vector<double> hist((size_t)b->res_cap);
    while (done < limit) {
        const int n = (int)std::min<long>(batch, limit - done);
        if ((st = launch_steps(b, n)) != JST_OK) return st;
        if ((st = fetch_ctl(b)) != JST_OK) return st;
        const CaseCtl &c = b->h_ctl[id];
        const long nd = (long)(c.step - start) - done;
        if (nd > 0) {
            if ((st = download(b, hc->dev.res_hist, hist.size() * 8, hist.data())) != JST_OK) return st;
            for (long q = 0; q < nd; ++q)
                if (cb) cb(done + q + 1, hist[(size_t)((start + done + q) % b->res_cap)], user);
            done += nd;
        }
        if (c.err_key != JST_ERR_NONE) {
            if (out) *out = jst_report{done, c.residual, c.totaltime, 0};
            const long long cell = (long long)(c.err_key & 0xFFFFFFFFull);
            char buf[160];
            const long long seq = (long long)(c.err_key >> 32);
            snprintf(buf, sizeof buf, "non-physical state at cell (%lld, %lld) in RK stage %lld of step %lld",
                     cell / hc->nj, cell % hc->nj, seq % 5 + 1, seq / 5 + 1);
            set_ctl_field(b, id, &CaseCtl::timing, 0);
            return fail(JST_ERR_NONPHYSICAL, buf);
        }
        if (c.status != JST_RUN_ACTIVE) break;
        if (nd <= 0) return fail(JST_ERR_HIP, "the device completed no step of an active case");
        batch = std::min(batch * 2, std::min(512, b->res_cap));
    }
    if ((st = set_ctl_field(b, id, &CaseCtl::timing, 0)) != JST_OK) return st;
    const CaseCtl &c = b->h_ctl[id];
    if (out) {
        out->steps = done;
        out->residual = done > 0 ? c.residual : NAN;
        out->totaltime = c.totaltime;
        out->converged = c.status == JST_RUN_CONVERGED ? 1 : 0;
    }
    return JST_OK;
}

} // namespace jsth

extern "C" {

jst_status jst_case_footprint(const char *config_json, const char *mesh_text, jst_precision prec,
                              size_t *bytes)
{
    if (!bytes) return fail(JST_ERR_ARG, "bytes is NULL");
    Config c;
    Mesh m;
    jst_status st = parse_config(config_json, c);
    if (st != JST_OK) return st;
    st = parse_mesh(mesh_text, m);
    if (st != JST_OK) return st;
    if (prec != JST_MIXED) return fail(JST_ERR_ARG, "batches run the FP32-mixed path");
    const int ni = m.n_rings - 1, nj = m.n_theta, pitch = (nj + 31) / 32 * 32;
    const int npost = (int)(((size_t)ni * nj + JST_POST_CELLS - 1) / JST_POST_CELLS);
    *bytes = layout_for(ni, nj, pitch, npost, 4096).total + r256(m.xy.size() * 8);
    return JST_OK;
}

jst_status jst_batch_create(jst_precision prec, size_t hbm_bytes, int max_cases, jst_batch **out)
{
    if (!out || max_cases <= 0 || hbm_bytes == 0) return fail(JST_ERR_ARG, "bad arguments");
    if (prec != JST_MIXED) return fail(JST_ERR_ARG, "batches run the FP32-mixed path");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(JST_ERR_HIP, "no HIP device available");
    std::unique_ptr<jst_batch> b(new jst_batch());
    b->prec = prec;
    b->max_cases = max_cases;
    b->cases.resize(max_cases);
    HIPCHK(hipGetDevice(&b->device));
    // set-up of the allocation-free entry points: the device address of this device's fixed
    // work-item array (looked up once per device)
    HIPCHK(jst::noalloc_scratch(b->device, &b->noalloc_items));
    HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&b->ev[0]));
    HIPCHK(hipEventCreate(&b->ev[1]));
    HIPCHK(hipMalloc((void **)&b->arena.base, hbm_bytes));
    b->arena.size = hbm_bytes;
    b->arena.free_[0] = hbm_bytes;
    HIPCHK(hipMalloc((void **)&b->d_cases, sizeof(CaseMix) * (size_t)max_cases));
    HIPCHK(hipMalloc((void **)&b->d_ctl, sizeof(CaseCtl) * (size_t)max_cases));
    HIPCHK(hipMalloc((void **)&b->d_slots, sizeof(int) * (size_t)max_cases));
    HIPCHK(hipMalloc((void **)&b->d_scratch_out, 64));
    HIPCHK(hipMalloc((void **)&b->d_view, sizeof(CaseDev)));
    HIPCHK(hipHostMalloc((void **)&b->h_ctl, sizeof(CaseCtl) * (size_t)max_cases, hipHostMallocDefault));
    memset(b->h_ctl, 0, sizeof(CaseCtl) * (size_t)max_cases);
    for (int q = 0; q < max_cases; ++q) {
        b->h_ctl[q].status = JST_RUN_STOPPED;
        b->h_ctl[q].err_key = JST_ERR_NONE;
        b->h_ctl[q].dt_acc = JST_INF_BITS;
    }
    HIPCHK(hipMemcpy(b->d_ctl, b->h_ctl, sizeof(CaseCtl) * (size_t)max_cases, hipMemcpyHostToDevice));
    *out = b.release();
    return JST_OK;
}

void jst_batch_destroy(jst_batch *b) { delete b; }

jst_status jst_batch_add(jst_batch *b, const char *config_json, const char *mesh_text,
                         long max_steps, int *case_id)
{
    if (!b || !case_id) return fail(JST_ERR_ARG, "NULL argument");
    std::unique_ptr<HostCase> hc(new HostCase());
    Mesh mesh;
    jst_status st = parse_config(config_json, hc->cfg);
    if (st != JST_OK) return st;
    st = parse_mesh(mesh_text, mesh);
    if (st != JST_OK) return st;
    if (hc->cfg.halo < JST_H) return fail(JST_ERR_CONFIG, "unsupported halo: the stencils need IM >= 3");
    if (mesh.n_rings - 1 < JST_H || mesh.n_theta < 4)
        return fail(JST_ERR_ARG, "the mixed path needs at least 3 cell rings and 4 cells per ring");
    // 32-bit byte offsets: every array of a case (at most 6 planes of (NI+1) x pitch, 8
    // bytes), a 5-plane FP64 state (40 B per cell) and the FP32 geometry / V/dt / local dt
    // span the stage kernels address through one descriptor (40 B per cell) stay below 4 GiB
    if ((size_t)mesh.n_rings * ((mesh.n_theta + 31) / 32 * 32) * 6 * 8 >= ((size_t)1 << 32))
        return fail(JST_ERR_ARG, "mesh too large for 32-bit array offsets");
    int slot = -1;
    for (int q = 0; q < b->max_cases; ++q)
        if (!b->cases[q]) {
            slot = q;
            break;
        }
    if (slot < 0) return fail(JST_ERR_NOSPACE, "case table full");
    hc->slot = slot;
    hc->d = finish(hc->cfg);
    hc->ni = mesh.n_rings - 1;
    hc->nj = mesh.n_theta;
    hc->pitch = (hc->nj + 31) / 32 * 32;
    hc->plane = (size_t)hc->ni * hc->pitch;
    hc->fplane = (size_t)(hc->ni + 1) * hc->pitch;
    hc->npost = (int)(((size_t)hc->ni * hc->nj + JST_POST_CELLS - 1) / JST_POST_CELLS);
    hc->max_steps = max_steps > 0 ? max_steps : hc->cfg.iteration;
    const CaseLayout L = layout_for(hc->ni, hc->nj, hc->pitch, hc->npost, b->res_cap);
    const size_t xy_bytes = r256(mesh.xy.size() * 8);
    long long off = b->arena.alloc(L.total + xy_bytes);
    if (off < 0) return fail(JST_ERR_NOSPACE, "HBM arena cannot hold the case now");
    hc->off = (size_t)off;
    hc->bytes = L.total + xy_bytes;
    char *base = b->arena.base + off;
    auto D = [&](size_t o) { return (double *)(base + o); };
    auto F = [&](size_t o) { return (float *)(base + o); };
    CaseMix &cm = hc->dev;
    cm.ni = hc->ni;
    cm.nj = hc->nj;
    cm.pitch = hc->pitch;
    cm.npost = hc->npost;
    cm.seg_rows = case_rows(b, *hc);
    cm.sb = F(L.sb);
    for (int q = 0; q < 2; ++q) {
        cm.u64[q] = D(L.u64[q]);
        cm.s32[q] = F(L.s32[q]);
    }
    jst::GeomPtrs &g = hc->g64;
    double **g64p[14] = {&g.vol, &g.cx, &g.cy, &g.wd, &g.inv_vol, &g.inv_wd2, &g.tnx, &g.tny,
                         &g.tmx, &g.tmy, &g.nnx, &g.nny, &g.nmx, &g.nmy};
    for (int q = 0; q < 14; ++q) *g64p[q] = D(L.g64[q]);
    float *tau2 = F(L.tau2), *cell4 = F(L.cell4);
    cm.tau2 = tau2;
    cm.cell4 = cell4;
    cm.tnx64 = g.tnx;
    cm.tny64 = g.tny;
    cm.tmx64 = g.tmx;
    cm.tmy64 = g.tmy;
    cm.cx64 = g.cx;
    cm.cy64 = g.cy;
    for (int q = 0; q < 2; ++q) {
        cm.ratio[q] = F(L.ratio[q]);
        cm.localdt[q] = F(L.localdt[q]);
    }
    cm.rpart = D(L.rpart);
    cm.dpart = D(L.dpart);
    cm.res_hist = D(L.res_hist);
    cm.res_cap = b->res_cap;
    cm.ctl = b->d_ctl + slot;
    cm.dbg = F(L.dbg);
    cm.k = make_mix_const(hc->cfg, hc->d);
    hc->d_posts = (PostDesc *)(base + L.posts);
    hc->scratch = D(L.scratch);
    for (int p = 0; p < hc->npost; ++p) {
        const int first = p * JST_POST_CELLS;
        const int count = std::min(JST_POST_CELLS, hc->ni * hc->nj - first);
        hc->posts.push_back(PostDesc{slot, first, count, p});
    }
    hipStream_t s = b->stream;
    // zero the block (ghost-free planes, pitch padding), geometry in FP64, FP32 copies
    HIPCHK(hipMemsetAsync(base, 0, L.total, s));
    double *d_xy = D(L.total);
    HIPCHK(hipMemcpyAsync(d_xy, mesh.xy.data(), mesh.xy.size() * 8, hipMemcpyHostToDevice, s));
    HIPCHK(jst::launch_geometry(d_xy, mesh.n_theta, hc->ni, hc->nj, hc->pitch, g, s));
    HIPCHK(jst::launch_to_f32_strided(g.tnx, tau2, hc->fplane, 2, 0, s));
    HIPCHK(jst::launch_to_f32_strided(g.tny, tau2, hc->fplane, 2, 1, s));
    HIPCHK(jst::launch_to_f32_strided(g.nnx, cell4, hc->plane, 4, 0, s));
    HIPCHK(jst::launch_to_f32_strided(g.nny, cell4, hc->plane, 4, 1, s));
    HIPCHK(jst::launch_to_f32_strided(g.vol, cell4, hc->plane, 4, 2, s));
    HIPCHK(jst::launch_to_f32_strided(g.inv_wd2, cell4, hc->plane, 4, 3, s));
    HIPCHK(hipMemcpyAsync(hc->d_posts, hc->posts.data(), hc->posts.size() * sizeof(PostDesc),
                          hipMemcpyHostToDevice, s));
    // freestream start (jst-rs state.rs:174-190), packed
    {
        std::vector<double> h(5 * hc->plane, 0.0);
        for (int m = 0; m < 5; ++m)
            for (int i = 0; i < hc->ni; ++i)
                for (int j = 0; j < hc->nj; ++j) h[m * hc->plane + (size_t)i * hc->pitch + j] = cm.k.u64[m];
        HIPCHK(hipMemcpyAsync(cm.u64[0], h.data(), h.size() * 8, hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    CaseCtl c0{};
    c0.dt_acc = JST_INF_BITS;
    c0.residual = NAN;
    c0.err_key = JST_ERR_NONE;
    c0.status = JST_RUN_ACTIVE;
    c0.stop_step = hc->max_steps;
    b->h_ctl[slot] = c0;
    HIPCHK(hipMemcpyAsync(b->d_ctl + slot, &b->h_ctl[slot], sizeof(CaseCtl), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(b->d_cases + slot, &cm, sizeof(CaseMix), hipMemcpyHostToDevice, s));
    HIPCHK(jst::launch_shadow(b->d_cases, slot, 5 * hc->plane, s));
    HIPCHK(jst::launch_dt_mixed(b->d_cases, hc->d_posts, hc->npost, s));
    HIPCHK(hipStreamSynchronize(s));
    // the node buffer is only needed for the geometry precompute
    b->arena.release(hc->off + L.total, xy_bytes);
    hc->bytes = L.total;
    b->cases[slot] = std::move(hc);
    b->dirty = true;
    *case_id = slot;
    return JST_OK;
}

jst_status jst_batch_remove(jst_batch *b, int case_id)
{
    HostCase *hc = get_case(b, case_id);
    if (!hc) return fail(JST_ERR_ARG, "no such case");
    HIPCHK(hipStreamSynchronize(b->stream));
    jst_status st = set_ctl_field(b, case_id, &CaseCtl::status, (int)JST_RUN_STOPPED);
    if (st != JST_OK) return st;
    b->arena.release(hc->off, hc->bytes);
    b->cases[case_id].reset();
    b->dirty = true;
    return JST_OK;
}

jst_status jst_batch_step(jst_batch *b, int nsteps, int sync)
{
    if (!b || nsteps < 0) return fail(JST_ERR_ARG, "bad arguments");
    jst_status st = launch_steps(b, nsteps);
    if (st != JST_OK) return st;
    return sync ? fetch_ctl(b) : JST_OK;
}

jst_status jst_batch_info(const jst_batch *b, int case_id, jst_case_info *out)
{
    HostCase *hc = get_case(b, case_id);
    if (!hc || !out) return fail(JST_ERR_ARG, "no such case");
    const CaseCtl &c = b->h_ctl[case_id];
    out->step = (long)c.step;
    out->status = c.err_key != JST_ERR_NONE ? JST_CASE_NONPHYSICAL : c.status;
    out->residual = c.residual;
    out->res_first = c.res_first;
    out->totaltime = c.totaltime;
    out->cells = (long)hc->ni * hc->nj;
    return JST_OK;
}

jst_status jst_batch_counts(const jst_batch *b, int *active_cases, long *active_cells)
{
    if (!b) return fail(JST_ERR_ARG, "NULL batch");
    int n = 0;
    long cells = 0;
    for (auto &hc : b->cases)
        if (hc && b->h_ctl[hc->slot].status == JST_RUN_ACTIVE && b->h_ctl[hc->slot].err_key == JST_ERR_NONE) {
            ++n;
            cells += (long)hc->ni * hc->nj;
        }
    if (active_cases) *active_cases = n;
    if (active_cells) *active_cells = cells;
    return JST_OK;
}

jst_status jst_batch_history(jst_batch *b, int case_id, long first, long n, double *res)
{
    HostCase *hc = get_case(b, case_id);
    if (!hc || !res || first < 1 || n < 0) return fail(JST_ERR_ARG, "bad arguments");
    std::vector<double> h((size_t)b->res_cap);
    jst_status st = download(b, hc->dev.res_hist, h.size() * 8, h.data());
    if (st != JST_OK) return st;
    for (long q = 0; q < n; ++q) res[q] = h[(size_t)((first - 1 + q) % b->res_cap)];
    return JST_OK;
}

jst_status jst_batch_wall(jst_batch *b, int case_id, double *x, double *y, double *cp, double *cf)
{
    if (!b) return fail(JST_ERR_ARG, "batch is NULL");
    return batch_wall(b, case_id, x, y, cp, cf);
}

jst_status jst_batch_set_state(jst_batch *b, int case_id, const double *u_planes,
                               jst_nonphysical *err)
{
    if (!b || !u_planes) return fail(JST_ERR_ARG, "NULL argument");
    HostCase *hc = get_case(b, case_id);
    if (!hc) return fail(JST_ERR_ARG, "no such case");
    // the unpack check first: a rejected state leaves the case as it was
    jst_status st = check_injected_state(u_planes, hc->ni, hc->nj, hc->cfg.gamma, err);
    if (st != JST_OK) return st;
    if ((st = fetch_ctl(b)) != JST_OK) return st;
    // U^n into the current FP64 state, then the FP32 base and the frozen dt / V/dt of it
    const int rot = b->h_ctl[hc->slot].rot;
    const size_t ncell = (size_t)hc->ni * hc->nj;
    std::vector<double> h(5 * hc->plane, 0.0);
    for (int m = 0; m < 5; ++m)
        for (int i = 0; i < hc->ni; ++i)
            for (int j = 0; j < hc->nj; ++j)
                h[m * hc->plane + (size_t)i * hc->pitch + j] = u_planes[m * ncell + (size_t)i * hc->nj + j];
    HIPCHK(hipMemcpyAsync(hc->dev.u64[rot], h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                          b->stream));
    if ((st = set_ctl_field(b, hc->slot, &CaseCtl::err_key, (unsigned long long)JST_ERR_NONE)) != JST_OK)
        return st;
    if ((st = set_ctl_field(b, hc->slot, &CaseCtl::status, (int)JST_RUN_ACTIVE)) != JST_OK) return st;
    HIPCHK(jst::launch_shadow(b->d_cases, hc->slot, 5 * hc->plane, b->stream));
    HIPCHK(jst::launch_dt_mixed(b->d_cases, hc->d_posts, hc->npost, b->stream));
    return fetch_ctl(b);
}

jst_status jst_batch_forces(jst_batch *b, int case_id, double *cl, double *cd)
{
    double f[4];
    jst_status st = jsth::batch_forces(b, case_id, f);
    if (st != JST_OK) return st;
    if (cl) *cl = f[2];
    if (cd) *cd = f[3];
    return JST_OK;
}

// The frozen dt, local dt and V/dt of the current state are always current on this path (set
// at admission, then by stage 5 of every step), so the export evaluates the terms with them
// and leaves the trajectory untouched.
jst_status jst_batch_residual_terms(jst_batch *b, int case_id, double *dt_out)
{
    HostCase *hc = get_case(b, case_id);
    if (!hc) return fail(JST_ERR_ARG, "no such case");
    std::vector<TileDesc> t;
    case_items(*hc, case_rows(b, *hc), t);
    TileDesc *dt = nullptr;
    HIPCHK(hipMalloc(&dt, t.size() * sizeof(TileDesc)));
    std::unique_ptr<TileDesc, void (*)(TileDesc *)> guard(dt, [](TileDesc *p) { (void)hipFree(p); });
    HIPCHK(hipMemcpyAsync(dt, t.data(), t.size() * sizeof(TileDesc), hipMemcpyHostToDevice, b->stream));
    HIPCHK(jst::launch_stage_mixed(b->d_cases, dt, (int)t.size(), 0, 0, RK_ALPHA[0], true,
                                   hc->cfg.sa_neg, b->stream));
    jst_status st = fetch_ctl(b); // synchronises the stream before the list is freed
    if (st != JST_OK) return st;
    if (dt_out) *dt_out = b->h_ctl[case_id].dt;
    return JST_OK;
}

// The same export without any device allocation: the case's work-item list is generated on
// the device into the library's fixed per-device array instead of a fresh buffer.  Same
// kernels, same items (the case's segment length under the batch's tiling), so the exported
// terms are bitwise those of jst_batch_residual_terms.
jst_status jst_batch_residual_terms_noalloc(jst_batch *b, int case_id, double *dt_out)
{
    HostCase *hc = get_case(b, case_id);
    if (!hc) return fail(JST_ERR_ARG, "no such case");
    const int rows = case_rows(b, *hc);
    const long nitems = (long)((hc->ni + rows - 1) / rows) * ((hc->nj + MIX_COLS - 1) / MIX_COLS);
    if (nitems > JST_NOALLOC_MAX_ITEMS) {
        char buf[200];
        snprintf(buf, sizeof buf,
                 "case needs %ld work items; the allocation-free export holds at most %d", nitems,
                 JST_NOALLOC_MAX_ITEMS);
        return fail(JST_ERR_ARG, buf);
    }
    int cur = -1;
    HIPCHK(hipGetDevice(&cur));
    if (cur != b->device) return fail(JST_ERR_ARG, "the batch's device is not the current device");
    int n = 0;
    HIPCHK(jst::launch_make_items(b->noalloc_items, hc->slot, hc->ni, hc->nj, rows, &n, b->stream));
    HIPCHK(jst::launch_stage_mixed(b->d_cases, b->noalloc_items, n, 0, 0, RK_ALPHA[0], true,
                                   hc->cfg.sa_neg, b->stream));
    jst_status st = fetch_ctl(b); // waits for the stream
    if (st != JST_OK) return st;
    if (dt_out) *dt_out = b->h_ctl[case_id].dt;
    return JST_OK;
}

jst_status jst_batch_get_field(jst_batch *b, int case_id, const char *name, double *dst, size_t n)
{
    HostCase *hc = get_case(b, case_id);
    if (!hc || !name || !dst) return fail(JST_ERR_ARG, "bad arguments");
    const int ni = hc->ni, nj = hc->nj, pitch = hc->pitch;
    std::string s(name);
    jst_status st = fetch_ctl(b);
    if (st != JST_OK) return st;
    const jst::GeomPtrs &g = hc->g64;
    struct G {
        const char *n;
        const double *p;
        bool face;
    } geo[] = {{"vol", g.vol, false}, {"cx", g.cx, false}, {"cy", g.cy, false},
               {"wall_dist", g.wd, false}, {"inv_vol", g.inv_vol, false}, {"inv_wd2", g.inv_wd2, false},
               {"tau_nx", g.tnx, true}, {"tau_ny", g.tny, true}, {"tau_mx", g.tmx, true},
               {"tau_my", g.tmy, true}, {"n_nx", g.nnx, false}, {"n_ny", g.nny, false},
               {"n_mx", g.nmx, false}, {"n_my", g.nmy, false}};
    for (auto &gd : geo)
        if (s == gd.n) {
            const int rows = gd.face ? ni + 1 : ni;
            if (n < (size_t)rows * nj) return fail(JST_ERR_ARG, "destination too small");
            std::vector<double> h((size_t)rows * pitch);
            if ((st = download(b, gd.p, h.size() * 8, h.data())) != JST_OK) return st;
            for (int i = 0; i < rows; ++i)
                for (int j = 0; j < nj; ++j) dst[(size_t)i * nj + j] = h[(size_t)i * pitch + j];
            return JST_OK;
        }
    if (n < (size_t)ni * nj) return fail(JST_ERR_ARG, "destination too small");
    auto plane32 = [&](const float *p) -> jst_status {
        std::vector<float> h(hc->plane);
        jst_status q = download(b, p, h.size() * 4, h.data());
        if (q != JST_OK) return q;
        for (int i = 0; i < ni; ++i)
            for (int j = 0; j < nj; ++j) dst[(size_t)i * nj + j] = h[(size_t)i * pitch + j];
        return JST_OK;
    };
    if (s.size() == 2 && s[0] == 'R' && s[1] >= '0' && s[1] <= '4')
        return plane32(hc->dev.dbg + (size_t)(s[1] - '0') * hc->plane);
    if (s == "S") return plane32(hc->dev.dbg + 5 * hc->plane);
    if (s == "localdt") return plane32(hc->dev.localdt[b->h_ctl[case_id].rot]);
    if (s == "ratio") return plane32(hc->dev.ratio[b->h_ctl[case_id].rot]);
    std::vector<double> h(5 * hc->plane);
    if ((st = download(b, hc->dev.u64[b->h_ctl[case_id].rot], h.size() * 8, h.data())) != JST_OK)
        return st;
    if (s.size() == 2 && s[0] == 'U' && s[1] >= '0' && s[1] <= '4') {
        const size_t base = (size_t)(s[1] - '0') * hc->plane;
        for (int i = 0; i < ni; ++i)
            for (int j = 0; j < nj; ++j) dst[(size_t)i * nj + j] = h[base + (size_t)i * pitch + j];
        return JST_OK;
    }
    static const char *prims[] = {"rho", "p", "T", "u", "v", "E", "H", "c", "nut"};
    int which = -1;
    for (int q = 0; q < 9; ++q)
        if (s == prims[q]) which = q;
    if (which < 0) return fail(JST_ERR_ARG, "unknown field " + s);
    const Config &c = hc->cfg;
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j) {
            double u[5];
            for (int m = 0; m < 5; ++m) u[m] = h[m * hc->plane + (size_t)i * pitch + j];
            const double rho = u[0], inv = 1.0 / rho, vx = u[1] * inv, vy = u[2] * inv, e = u[3] * inv;
            const double p = (c.gamma - 1.0) * (u[3] - rho * (vx * vx + vy * vy) * 0.5);
            const double t = p / (c.r_gas * rho);
            const double v[9] = {rho, p, t, vx, vy, e, e + p * inv, std::sqrt(c.r_gas * c.gamma * t), u[4] * inv};
            dst[(size_t)i * nj + j] = v[which];
        }
    return JST_OK;
}

jst_status jst_batch_segment_rows(const jst_batch *b, int case_id, int *rows)
{
    if (!rows) return fail(JST_ERR_ARG, "bad arguments");
    const HostCase *hc = get_case(b, case_id);
    if (!hc) return fail(JST_ERR_ARG, "no such case");
    // under the fitted variant the lists may not have been rebuilt since the last change of
    // the batch: report the length the next step will use
    *rows = b->variant == MIX_VARIANT_FIT ? fit_segment_rows(b) : case_rows(b, *hc);
    return JST_OK;
}

jst_status jst_batch_set_tiling(jst_batch *b, int variant, int grid)
{
    const bool rows_ok = variant > MIX_VARIANT_ROWS && variant <= MIX_VARIANT_ROWS + MIX_MAX_ROWS;
    if (!b || variant < 0 || (variant > MIX_VARIANT_FIT && !rows_ok) || grid < 0)
        return fail(JST_ERR_ARG, "bad tiling");
    HIPCHK(hipStreamSynchronize(b->stream));
    if (variant != b->variant) b->dirty = true;
    if (grid != b->grid) {
        jst_status st = drop_graph(b);
        if (st != JST_OK) return st;
    }
    b->variant = variant;
    b->grid = grid;
    return JST_OK;
}

jst_status jst_batch_time_begin(jst_batch *b, int nsteps)
{
    if (!b || nsteps <= 0) return fail(JST_ERR_ARG, "bad arguments");
    jst_status st = fetch_ctl(b);
    if (st != JST_OK) return st;
    b->timing_slots.clear();
    for (auto &hc : b->cases)
        if (hc && b->h_ctl[hc->slot].status == JST_RUN_ACTIVE) b->timing_slots.push_back(hc->slot);
    if ((st = set_ctl_fields(b, b->timing_slots, &CaseCtl::timing, 1)) != JST_OK) return st;
    if (b->dirty && (st = rebuild_lists(b)) != JST_OK) return st;
    // build the graphs that queueing nsteps steps uses and launch each once, so that neither
    // capture, instantiation nor a graph's first upload falls inside the timed steps
    if ((st = prepare_graphs(b, nsteps)) != JST_OK) return st;
    if (nsteps >= GRAPH_STEPS && (st = launch_steps(b, GRAPH_STEPS)) != JST_OK) return st;
    if ((nsteps % GRAPH_STEPS || nsteps < GRAPH_STEPS) && (st = launch_steps(b, 1)) != JST_OK) return st;
    HIPCHK(hipStreamSynchronize(b->stream));
    return JST_OK;
}

jst_status jst_batch_time_end(jst_batch *b)
{
    if (!b) return fail(JST_ERR_ARG, "bad arguments");
    const std::vector<int> touched = b->timing_slots;
    b->timing_slots.clear();
    jst_status st = set_ctl_fields(b, touched, &CaseCtl::timing, 0);
    if (st != JST_OK) return st;
    if ((st = fetch_ctl(b)) != JST_OK) return st;
    // a diverged case stops its own kernels early: a timing that includes one is not valid
    for (int slot : touched)
        if (b->h_ctl[slot].err_key != JST_ERR_NONE) {
            const long long cell = (long long)(b->h_ctl[slot].err_key & 0xFFFFFFFFull);
            char buf[160];
            snprintf(buf, sizeof buf, "non-physical state in case %d at cell (%lld, %lld) during timing",
                     slot, cell / b->cases[slot]->nj, cell % b->cases[slot]->nj);
            return fail(JST_ERR_NONPHYSICAL, buf);
        }
    return JST_OK;
}

jst_status jst_batch_time_steps(jst_batch *b, int nsteps, double *total_ms, double *stage_ms)
{
    jst_status st = jst_batch_time_begin(b, nsteps);
    if (st != JST_OK) return st;
    HIPCHK(hipEventRecord(b->ev[0], b->stream));
    if ((st = launch_steps(b, nsteps)) != JST_OK) return st;
    HIPCHK(hipEventRecord(b->ev[1], b->stream));
    HIPCHK(hipEventSynchronize(b->ev[1]));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    if (total_ms) *total_ms = ms;
    if (stage_ms) {
        std::vector<hipEvent_t> evs(10 * (size_t)nsteps);
        for (auto &e : evs) HIPCHK(hipEventCreate(&e));
        for (int s = 0; s < nsteps; ++s) {
            for (int k = 0; k < 5; ++k) {
                HIPCHK(hipEventRecord(evs[10 * s + 2 * k], b->stream));
                HIPCHK(jst::launch_stage_mixed(b->d_cases, b->d_tiles, b->ntiles, b->grid, k, RK_ALPHA[k],
                                               false, b->neg, b->stream));
                HIPCHK(hipEventRecord(evs[10 * s + 2 * k + 1], b->stream));
            }
            HIPCHK(jst::launch_end_mixed(b->d_cases, b->d_slots, b->nslots, b->stream));
        }
        HIPCHK(hipStreamSynchronize(b->stream));
        double sum = 0.0;
        for (int q = 0; q < 5 * nsteps; ++q) {
            float e = 0.f;
            HIPCHK(hipEventElapsedTime(&e, evs[2 * q], evs[2 * q + 1]));
            sum += e;
        }
        for (auto &e : evs) (void)hipEventDestroy(e);
        *stage_ms = sum / (5.0 * nsteps);
    }
    return jst_batch_time_end(b);
}

jst_status jst_batch_time_kernels(jst_batch *b, int nsteps, double *inner_ms, double *last_ms,
                                  double *end_ms)
{
    if (!b || nsteps <= 0) return fail(JST_ERR_ARG, "bad arguments");
    jst_status st = fetch_ctl(b);
    if (st != JST_OK) return st;
    std::vector<int> touched;
    for (auto &hc : b->cases)
        if (hc && b->h_ctl[hc->slot].status == JST_RUN_ACTIVE) touched.push_back(hc->slot);
    if ((st = set_ctl_fields(b, touched, &CaseCtl::timing, 1)) != JST_OK) return st;
    if (b->dirty && (st = rebuild_lists(b)) != JST_OK) return st;
    std::vector<hipEvent_t> evs(12 * (size_t)nsteps);
    for (auto &e : evs) HIPCHK(hipEventCreate(&e));
    for (int s = 0; s < nsteps; ++s) {
        for (int k = 0; k < 5; ++k) {
            HIPCHK(hipEventRecord(evs[12 * s + 2 * k], b->stream));
            HIPCHK(jst::launch_stage_mixed(b->d_cases, b->d_tiles, b->ntiles, b->grid, k, RK_ALPHA[k],
                                           false, b->neg, b->stream));
            HIPCHK(hipEventRecord(evs[12 * s + 2 * k + 1], b->stream));
        }
        HIPCHK(hipEventRecord(evs[12 * s + 10], b->stream));
        HIPCHK(jst::launch_end_mixed(b->d_cases, b->d_slots, b->nslots, b->stream));
        HIPCHK(hipEventRecord(evs[12 * s + 11], b->stream));
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    double acc[3] = {0.0, 0.0, 0.0};
    for (int s = 0; s < nsteps; ++s)
        for (int k = 0; k < 6; ++k) {
            float e = 0.f;
            HIPCHK(hipEventElapsedTime(&e, evs[12 * s + 2 * k], evs[12 * s + 2 * k + 1]));
            acc[k < 4 ? 0 : (k == 4 ? 1 : 2)] += e;
        }
    for (auto &e : evs) (void)hipEventDestroy(e);
    if (inner_ms) *inner_ms = acc[0] / (4.0 * nsteps);
    if (last_ms) *last_ms = acc[1] / nsteps;
    if (end_ms) *end_ms = acc[2] / nsteps;
    if ((st = set_ctl_fields(b, touched, &CaseCtl::timing, 0)) != JST_OK) return st;
    if ((st = fetch_ctl(b)) != JST_OK) return st;
    for (int slot : touched)
        if (b->h_ctl[slot].err_key != JST_ERR_NONE)
            return fail(JST_ERR_NONPHYSICAL, "non-physical state during timing");
    return JST_OK;
}

// Calibration probe for the HBM byte counters: one device-to-device copy of `bytes`.
jst_status jst_copy_probe(size_t bytes, double *ms)
{
    void *a = nullptr, *c = nullptr;
    hipEvent_t e0, e1;
    HIPCHK(hipMalloc(&a, bytes));
    HIPCHK(hipMalloc(&c, bytes));
    HIPCHK(hipMemset(a, 1, bytes));
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, nullptr));
    HIPCHK(hipMemcpyAsync(c, a, bytes, hipMemcpyDeviceToDevice, nullptr));
    HIPCHK(hipEventRecord(e1, nullptr));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(hipFree(a));
    HIPCHK(hipFree(c));
    return JST_OK;
}

static jst_status tune_impl(jst_batch *b, int trial_steps, bool variants, int *variant_out, int *grid_out)
{
    if (!b || trial_steps <= 0) return fail(JST_ERR_ARG, "bad arguments");
    int best_v = b->variant, best_g = b->grid;
    const int v0 = variants ? 0 : b->variant, v1 = variants ? MIX_VARIANT_FIT + 1 : b->variant + 1;
    // grid 0: one workgroup per four work items; otherwise a persistent grid of k workgroups
    // per CU.  Each candidate is timed in two rounds and keeps its faster time, so that a
    // one-off stall in one timing does not decide the choice.
    const int grids[3] = {0, 256 * 2, 256 * 4};
    std::vector<double> t((size_t)(v1 - v0) * 3, 1e300);
    for (int round = 0; round < 2; ++round)
        for (int v = v0; v < v1; ++v)
            for (int q = 0; q < 3; ++q) {
                jst_status st = jst_batch_set_tiling(b, v, grids[q]);
                if (st != JST_OK) return st;
                double ms = 0.0;
                if ((st = jst_batch_time_steps(b, trial_steps, &ms, nullptr)) != JST_OK) return st;
                double &m = t[(size_t)(v - v0) * 3 + q];
                m = std::min(m, ms);
            }
    double best = 1e300;
    for (int v = v0; v < v1; ++v)
        for (int q = 0; q < 3; ++q)
            if (t[(size_t)(v - v0) * 3 + q] < best) {
                best = t[(size_t)(v - v0) * 3 + q];
                best_v = v;
                best_g = grids[q];
            }
    jst_status st = jst_batch_set_tiling(b, best_v, best_g);
    if (st != JST_OK) return st;
    if (variant_out) *variant_out = best_v;
    if (grid_out) *grid_out = best_g;
    return JST_OK;
}

jst_status jst_batch_tune(jst_batch *b, int trial_steps, int *variant_out, int *grid_out)
{
    return tune_impl(b, trial_steps, true, variant_out, grid_out);
}

// ---------------------------------------------------------------------------------------
// Serving loop
// ---------------------------------------------------------------------------------------
jst_status jst_serve(const jst_request *reqs, int n, const jst_serve_opts *opts,
                     jst_request_result *results, jst_serve_summary *summary)
{
    if (!reqs || n < 0 || !opts || !results) return fail(JST_ERR_ARG, "bad arguments");
    jst_batch *b = nullptr;
    jst_status st = jst_batch_create(JST_MIXED, opts->hbm_bytes, opts->max_cases, &b);
    if (st != JST_OK) return st;
    std::unique_ptr<jst_batch> guard(b);
    // 0: every case its own segment length (from its size alone) while the GPU has waves to
    // spare, one length fitted to the whole batch while it is saturated; < 0: every case its
    // own length always (results bitwise independent of the load); else one fixed length
    const bool adaptive = opts->segment_rows == 0;
    {
        int v = opts->segment_rows <= 0 ? MIX_VARIANT_PER_CASE : -1;
        if (opts->segment_rows >= 1 && opts->segment_rows <= MIX_MAX_ROWS) v = MIX_VARIANT_ROWS + opts->segment_rows;
        if (v < 0) return fail(JST_ERR_ARG, "unsupported segment_rows");
        if ((st = jst_batch_set_tiling(b, v, 0)) != JST_OK) return st;
    }
    const int spr = opts->steps_per_round > 0 ? opts->steps_per_round : 50;
    std::vector<int> order(n);
    for (int q = 0; q < n; ++q) {
        order[q] = q;
        results[q] = jst_request_result{-1.0, -1.0, 0, -1, NAN, NAN, NAN};
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int c) { return reqs[a].arrival_s < reqs[c].arrival_s; });
    std::vector<int> slot_req(opts->max_cases, -1);
    std::vector<long> last_step(opts->max_cases, 0);
    double cell_stages = 0.0;
    int next = 0, done = 0, nonphys = 0, maxc = 0, active = 0, tuned_for = -1;
    double step_s = 0.0; // measured wall time per step of the current mix (round average)
    int switches = 0;    // adaptive tiling changes
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto now = [&]() { return std::chrono::duration<double>(clk::now() - t0).count(); };
    while (done < n) {
        // admission: in arrival order, while a slot and arena memory are free
        bool admitted = false;
        while (next < n && reqs[order[next]].arrival_s <= now()) {
            const jst_request &r = reqs[order[next]];
            int id = -1;
            st = jst_batch_add(b, r.config_json, r.mesh_text, r.max_steps, &id);
            // no space while other cases run: wait for a retirement.  No space in an empty
            // batch (whole arena free, every slot free): the request can never fit, so it is
            // rejected like an invalid one instead of being retried forever.
            if (st == JST_ERR_NOSPACE && active > 0) break;
            if (st != JST_OK) {
                results[order[next]].status = -1;
                ++done;
                ++next;
                continue;
            }
            results[order[next]].admit_s = now();
            slot_req[id] = order[next];
            last_step[id] = 0;
            ++next;
            ++active;
            admitted = true;
        }
        maxc = std::max(maxc, active);
        if (active == 0) {
            if (next < n) {
                const double wait = reqs[order[next]].arrival_s - now();
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.05)));
            }
            continue;
        }
        if (adaptive) {
            // Step time of each tiling, in row steps of one wave: a work item of L rows walks
            // L + 6 rows, and the launch lasts about the longer of one item and the total
            // over the 2048 waves two per SIMD keep resident.  Short own lengths win while the
            // GPU has waves to spare; one length fitted to the batch (fewer halo rows, every
            // item in one round where possible) once it is saturated.  The tiling changes only
            // when the other one is estimated 20 % faster.
            double lat_total = 0.0, lat_item = 0.0, t_fit = 0.0;
            for (auto &hc : b->cases)
                if (hc) {
                    const long groups = (hc->nj + MIX_COLS - 1) / MIX_COLS;
                    const int rows = own_segment_rows(hc->ni, hc->nj);
                    lat_total += (double)groups * (double)((hc->ni + rows - 1) / rows) * (rows + 6);
                    lat_item = std::max(lat_item, (double)std::min(rows, hc->ni) + 6.0);
                }
            (void)fit_segment_rows(b, &t_fit);
            const double t_lat = std::max(lat_item, lat_total / 2048.0);
            int want = b->variant;
            if (b->variant == MIX_VARIANT_PER_CASE && t_fit < 0.8 * t_lat) want = MIX_VARIANT_FIT;
            if (b->variant == MIX_VARIANT_FIT && t_lat < 0.8 * t_fit) want = MIX_VARIANT_PER_CASE;
            if (want != b->variant) {
                if ((st = jst_batch_set_tiling(b, want, b->grid)) != JST_OK) return st;
                ++switches;
            }
        }
        if (opts->tune && admitted && active != tuned_for && active >= 4) {
            int v, g;
            // grid only: the segment length stays fixed, so a request's result does not
            // depend on which other requests share the batch
            if ((st = tune_impl(b, 3, false, &v, &g)) != JST_OK) return st;
            tuned_for = active;
        }
        // Round length: at most steps_per_round steps, ending at the nearest step cap of a
        // running case or at the next arrival still to come (from the measured step time),
        // so that neither a retirement nor an admission waits for the rest of a long round.
        // A request that has arrived but is still queued waits for memory or a slot, which
        // only a retirement frees: it does not shorten the round.  Results do not depend on
        // the round length: every case advances by whole steps on its own data.
        int nsteps = spr;
        for (int id = 0; id < opts->max_cases; ++id)
            if (slot_req[id] >= 0) {
                const long left = get_case(b, id)->max_steps - last_step[id];
                if (left > 0 && left < nsteps) nsteps = (int)left;
            }
        if (next < n && step_s > 0.0) {
            const double wait = reqs[order[next]].arrival_s - now();
            if (wait > 0.0) {
                const double k = std::ceil(wait / step_s);
                if (k < (double)nsteps) nsteps = std::max(1, (int)k);
            }
        }
        const double t_round = now();
        if ((st = jst_batch_step(b, nsteps, 1)) != JST_OK) return st;
        const double per = (now() - t_round) / nsteps;
        step_s = step_s > 0.0 ? 0.5 * (step_s + per) : per;
        // account and retire
        bool progressed = false;
        for (int id = 0; id < opts->max_cases; ++id) {
            const int rq = slot_req[id];
            if (rq < 0) continue;
            jst_case_info ci;
            if ((st = jst_batch_info(b, id, &ci)) != JST_OK) return st;
            progressed = progressed || ci.step > last_step[id] || ci.status != JST_CASE_RUNNING;
            cell_stages += (double)ci.cells * 5.0 * (double)(ci.step - last_step[id]);
            last_step[id] = ci.step;
            if (ci.status == JST_CASE_RUNNING) continue;
            jst_request_result &res = results[rq];
            res.done_s = now();
            res.steps = ci.step;
            res.status = ci.status;
            res.residual = ci.residual;
            if (ci.status != JST_CASE_NONPHYSICAL) {
                if ((st = jst_batch_forces(b, id, &res.cl, &res.cd)) != JST_OK) return st;
            } else {
                ++nonphys;
            }
            if ((st = jst_batch_remove(b, id)) != JST_OK) return st;
            slot_req[id] = -1;
            --active;
            ++done;
        }
        if (!progressed) return fail(JST_ERR_HIP, "the device completed no step of an active case");
    }
    if (summary) {
        // latency = done - arrival; over every retired request, and over the converged ones
        std::vector<double> lat, lat_conv;
        for (int q = 0; q < n; ++q)
            if (results[q].done_s >= 0 && results[q].status >= 0) {
                lat.push_back(results[q].done_s - reqs[q].arrival_s);
                if (results[q].status == JST_CASE_CONVERGED) lat_conv.push_back(lat.back());
            }
        std::sort(lat.begin(), lat.end());
        std::sort(lat_conv.begin(), lat_conv.end());
        auto pct = [](const std::vector<double> &v, double p) -> double {
            if (v.empty()) return (double)NAN;
            const double idx = p * (v.size() - 1);
            const size_t lo = (size_t)std::floor(idx), hi = (size_t)std::ceil(idx);
            return v[lo] + (v[hi] - v[lo]) * (idx - lo);
        };
        summary->wall_s = now();
        summary->cell_stages = cell_stages;
        summary->p50_s = pct(lat, 0.50);
        summary->p99_s = pct(lat, 0.99);
        summary->completed = (int)lat.size();
        summary->nonphysical = nonphys;
        summary->max_concurrent = maxc;
        summary->tiling_switches = switches;
        summary->converged = (int)lat_conv.size();
        summary->p50_converged_s = pct(lat_conv, 0.50);
        summary->p99_converged_s = pct(lat_conv, 0.99);
        int rejected = 0;
        for (int q = 0; q < n; ++q) rejected += results[q].status < 0 ? 1 : 0;
        summary->rejected = rejected;
    }
    return JST_OK;
}

} // extern "C"
