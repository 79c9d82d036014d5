// jst_launch.h — host-side launchers of the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "jst_dev.h"

// FP64 stage tile: TI rows x TJ columns, one cell per thread.
#define FP64_TI 8
#define FP64_TJ 32
#define FP64_NT 256

namespace jst {

struct GeomPtrs {
    double *vol, *cx, *cy, *wd, *inv_vol, *inv_wd2;
    double *tnx, *tny, *tmx, *tmy; // (ni+1) x pitch
    double *nnx, *nny, *nmx, *nmy; // ni x pitch
};

hipError_t launch_geometry(const double *xy, int n_theta, int ni, int nj, int pitch,
                           const GeomPtrs &g, hipStream_t s);
hipError_t launch_stage_fp64(const CaseDev *cases, const TileDesc *tiles, int ntiles, int stage,
                             double alpha, bool terms, hipStream_t s);
hipError_t launch_post_fp64(const CaseDev *cases, const TileDesc *tiles, int ntiles, int mode,
                            hipStream_t s);
hipError_t launch_forces_fp64(const CaseDev *cases, int ci, double *scratch, double *out,
                              hipStream_t s);

// FP32-mixed row-march variants: rows per work item (one wave per 58 columns x that many
// rows), MIX_VARIANT_PER_CASE: each case its own length from its size alone, or
// MIX_VARIANT_FIT: one length fitted to the batch; the autotuner picks one per batch.
// Stage 5 leaves one residual sum and one dt minimum per (row segment, column group), from
// the case's segment length (CaseMix::seg_rows).  Results depend on the variant at FP32
// rounding level only (where a segment's edge rows run the general row body; the residual's
// FP64 summation order follows the segments), never on the launch grid.
#define MIX_NVARIANTS 10
#define MIX_VARIANT_PER_CASE MIX_NVARIANTS
// one length for every case, fitted to the whole batch (jst_batch.cpp fit_segment_rows)
#define MIX_VARIANT_FIT (MIX_NVARIANTS + 1)
// variant MIX_VARIANT_ROWS + L (1 <= L <= MIX_MAX_ROWS): every case L rows per work item
#define MIX_VARIANT_ROWS 100
#define MIX_MAX_ROWS 256
#define MIX_COLS 58
extern const int MIX_SEGMENT_ROWS[MIX_NVARIANTS];
// neg: launch the kernels compiled with the negative-nu~ SA branch (needed when any case of
// the item list has spalart_allmaras.negative; bitwise the same results for the others)
hipError_t launch_stage_mixed(const CaseMix *cases, const TileDesc *items, int nitems, int grid,
                              int stage, double alpha, bool terms, bool neg, hipStream_t s);
hipError_t launch_end_mixed(const CaseMix *cases, const int *slots, int nslots, hipStream_t s);
hipError_t launch_dt_mixed(const CaseMix *cases, const PostDesc *posts, int nposts, hipStream_t s);
hipError_t launch_to_f32(const double *a, float *b, size_t n, hipStream_t s);
// b[i * stride + lane] = (float)a[i], i < n: FP32 copies interleaved into a packed array
hipError_t launch_to_f32_strided(const double *a, float *b, size_t n, int stride, int lane, hipStream_t s);

// Allocation-free work-item lists: a fixed device array of JST_NOALLOC_MAX_ITEMS work items
// per device.  noalloc_scratch returns its address on `device` (looked up once per device
// with that device current); launch_make_items writes one case's items into it (every
// 58-column group x segment of `rows` rows) and fails with hipErrorInvalidValue when they
// exceed the array.
#define JST_NOALLOC_MAX_ITEMS 65536
hipError_t noalloc_scratch(int device, TileDesc **ptr);
hipError_t launch_make_items(TileDesc *items, int slot, int ni, int nj, int rows, int *nitems,
                             hipStream_t s);
hipError_t launch_shadow(const CaseMix *cases, int ci, size_t n5, hipStream_t s);

} // namespace jst
