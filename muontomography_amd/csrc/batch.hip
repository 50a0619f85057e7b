// Many independent scans in one FP32C iteration (one persistent launch per
// pass for all of them, one captured graph per iteration).
//
// The scans' grids are laid side by side along x in one grid (each offset
// to a multiple of the brick size, so no brick straddles two scans, and
// padded to the largest ny and nz); their track-major lists are
// concatenated with voxel indices moved into that grid.  The brick-major
// stream built from it holds the tasks of every scan in one size-sorted
// queue, so the persistent kernels balance large and small scans together.
// Tracks are ranked scan first, each scan's costs are summed on their own,
// and every scan keeps its own per-voxel update; voxels of the padding have
// no entries and stay 0.  Scans share the EM configuration (mean update,
// measurement errors); grid, voxel size and tracks differ.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "em_plan.hpp"

int muto_plan_alloc_common(muto_em_plan* pl);
int muto_read_counters(muto_em_plan* pl);

namespace muto {
int fp32c_build(muto_em_plan* pl);
int fp32c_run(muto_em_plan* pl, int iters, double* sums, int* done);
int fp32c_mirror(muto_em_plan* pl);
}  // namespace muto

using namespace muto;

struct muto_batch {
  muto_ctx* ctx = nullptr;
  muto_em_plan* grid = nullptr;  // the combined problem
  std::vector<muto_em_plan*> scans;
  std::vector<uint32_t> xoff;    // x offset of each scan in the combined grid
  std::vector<size_t> trk_off;   // first track of each scan
  double wall_ms = 0.0;
};

namespace {

// scan voxel j -> combined-grid voxel; out-of-grid indices stay invalid
__global__ void k_move_vox(uint64_t n, const uint32_t* __restrict__ in, uint32_t nx, uint32_t ny,
                           uint64_t nv, uint32_t xoff, uint32_t X, uint32_t Y,
                           uint32_t* __restrict__ out) {
  uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t j = in[k];
  if (j >= nv) {
    out[k] = 0xFFFFFFFFu;
    return;
  }
  const uint32_t x = j % nx, y = (j / nx) % ny, z = j / (nx * ny);
  out[k] = (x + xoff) + X * (y + Y * z);
}

__global__ void k_shift_rows(size_t n, const uint64_t* __restrict__ in, uint64_t add,
                             uint64_t* __restrict__ out) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = in[k] + add;
}

// scan image <-> its block of the combined grid
template <bool TO_GRID>
__global__ void k_image_block(uint64_t nv, uint32_t nx, uint32_t ny, uint32_t xoff, uint32_t X,
                              uint32_t Y, const double* __restrict__ src, double* __restrict__ dst,
                              const uint32_t* __restrict__ mj_src, uint32_t* __restrict__ mj_dst) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nv) return;
  const uint32_t x = (uint32_t)(j % nx), y = (uint32_t)((j / nx) % ny), z = (uint32_t)(j / ((uint64_t)nx * ny));
  const size_t g = (size_t)(x + xoff) + (size_t)X * (y + (size_t)Y * z);
  if (TO_GRID) {
    if (src) dst[g] = src[j];
    if (mj_src) mj_dst[g] = mj_src[j];
  } else {
    dst[j] = src[g];
  }
}

__global__ void k_fill_u16(size_t n, uint16_t v, uint16_t* __restrict__ out) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = v;
}

bool same_em(const muto_mlsd_cfg& a, const muto_mlsd_cfg& b) {
  return a.error_xy == b.error_xy && a.error_dz_outer == b.error_dz_outer &&
         a.error_dz_inner == b.error_dz_inner && a.update_with_mean == b.update_with_mean &&
         a.mean_percentile == b.mean_percentile;
}

}  // namespace

extern "C" int muto_batch_create(muto_ctx* ctx, size_t n, muto_em_plan* const* plans,
                                 muto_batch** out) {
  if (!ctx) return MUTO_ERR_ARG;
  if (!out || !plans || n == 0) { ctx->set_error("muto_batch_create: null or empty argument"); return MUTO_ERR_ARG; }
  *out = nullptr;
  if (n > 65535) { ctx->set_error("muto_batch_create: at most 65535 scans"); return MUTO_ERR_LIMIT; }
  for (size_t s = 0; s < n; ++s) {
    const muto_em_plan* p = plans[s];
    if (!p || p->ctx != ctx) { ctx->set_error("muto_batch_create: a plan is null or on another context"); return MUTO_ERR_ARG; }
    if (!p->has_lists) { ctx->set_error("muto_batch_create: a plan keeps no path lists"); return MUTO_ERR_STATE; }
    if (!(p->cfg.update_with_mean && p->cfg.mean_percentile == 100)) {
      ctx->set_error("muto_batch_create: batches run the FP32C mean update (update_with_mean, "
                     "mean_percentile 100)");
      return MUTO_ERR_ARG;
    }
    if (!same_em(p->cfg, plans[0]->cfg)) {
      ctx->set_error("muto_batch_create: scans must share the EM configuration (measurement errors)");
      return MUTO_ERR_ARG;
    }
  }
  hipStream_t st = ctx->stream;
  auto* b = new muto_batch();
  b->ctx = ctx;
  b->scans.assign(plans, plans + n);
  // combined grid: scans side by side along x, brick-aligned
  muto_em_plan* g = new muto_em_plan();
  g->ctx = ctx;
  g->cfg = plans[0]->cfg;
  g->f32.bs = plans[0]->f32.bs;  // the first plan's brick shape and block size
  const uint32_t bx = g->f32.bs.bx, by = g->f32.bs.by, bz = g->f32.bs.bz;
  uint64_t X = 0, Y = 0, Z = 0;
  double d[3] = {0, 0, 0};
  size_t nt = 0;
  uint64_t nnz = 0;
  for (size_t s = 0; s < n; ++s) {
    const muto_grid& sg = plans[s]->cfg.grid;
    b->xoff.push_back((uint32_t)X);
    b->trk_off.push_back(nt);
    X += (sg.nx + bx - 1) / bx * bx;
    Y = std::max<uint64_t>(Y, (sg.ny + by - 1) / by * by);
    Z = std::max<uint64_t>(Z, (sg.nz + bz - 1) / bz * bz);
    d[0] = std::max(d[0], sg.dx);
    d[1] = std::max(d[1], sg.dy);
    d[2] = std::max(d[2], sg.dz);
    nt += plans[s]->n_trk;
    nnz += plans[s]->nnz;
  }
  if (X * Y * Z >= 0xFFFFFFFFull || X > 0xFFFFFFFFull) {
    delete g;
    delete b;
    ctx->set_error("muto_batch_create: the scans' combined grid exceeds 2^32 voxels");
    return MUTO_ERR_LIMIT;
  }
  muto_grid cg{};
  cg.nx = (uint32_t)X;
  cg.ny = (uint32_t)Y;
  cg.nz = (uint32_t)Z;
  cg.dx = d[0];
  cg.dy = d[1];
  cg.dz = d[2];
  g->cfg.grid = cg;
  g->g = to_dev(cg);
  g->nv = n_vox(cg);
  g->n_trk = nt;
  g->nnz = nnz;
  g->n_scans = (uint32_t)n;
  auto fail = [&](int rc) {
    delete g;
    delete b;
    return rc;
  };
  if (int rc = muto_plan_alloc_common(g)) return fail(rc);
  if (n > 4) {
    g->cost_acc.release();
    if (g->cost_acc.alloc(n) != hipSuccess) { ctx->set_error("out of device memory"); return fail(MUTO_ERR_NOMEM); }
  }
  if (g->trk4.alloc(4 * nt) != hipSuccess || g->pr.alloc(nt) != hipSuccess ||
      g->row_ptr.alloc(nt + 1) != hipSuccess || g->vox.alloc(nnz) != hipSuccess ||
      g->len.alloc(nnz) != hipSuccess || g->tail.alloc(nnz) != hipSuccess ||
      g->mj.alloc(g->nv) != hipSuccess || g->trk_scan.alloc(nt) != hipSuccess) {
    ctx->set_error("muto_batch_create: out of device memory");
    return fail(MUTO_ERR_NOMEM);
  }
  // the combined grid's M_j is counted from its FP32C stream when that is built
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(g->mj.p, 0, g->nv * sizeof(uint32_t), st));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(g->dcount.p, 0, 8 * sizeof(unsigned long long), st));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(g->row_ptr.p, 0, sizeof(uint64_t), st));
  uint64_t eoff = 0;
  for (size_t s = 0; s < n; ++s) {
    const muto_em_plan* p = plans[s];
    const size_t to = b->trk_off[s], ni = p->n_trk;
    const muto_grid& sg = p->cfg.grid;
    if (ni) {
      MUTO_CHECK_HIP(ctx, hipMemcpyAsync(g->trk4.p + 4 * to, p->trk4.p, 4 * ni * sizeof(double), hipMemcpyDeviceToDevice, st));
      MUTO_CHECK_HIP(ctx, hipMemcpyAsync(g->pr.p + to, p->pr.p, ni * sizeof(double), hipMemcpyDeviceToDevice, st));
      k_shift_rows<<<grid_for(ni + 1, 256), 256, 0, st>>>(ni + 1, p->row_ptr.p, eoff, g->row_ptr.p + to);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
      k_fill_u16<<<grid_for(ni, 256), 256, 0, st>>>(ni, (uint16_t)s, g->trk_scan.p + to);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    }
    if (p->nnz) {
      MUTO_CHECK_HIP(ctx, hipMemcpyAsync(g->len.p + eoff, p->len.p, p->nnz * sizeof(double), hipMemcpyDeviceToDevice, st));
      MUTO_CHECK_HIP(ctx, hipMemcpyAsync(g->tail.p + eoff, p->tail.p, p->nnz * sizeof(double), hipMemcpyDeviceToDevice, st));
      k_move_vox<<<grid_for(p->nnz, 256), 256, 0, st>>>(p->nnz, p->vox.p, sg.nx, sg.ny, p->nv, b->xoff[s],
                                                        (uint32_t)X, (uint32_t)Y, g->vox.p + eoff);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    }
    eoff += p->nnz;
  }
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(st));
  g->counters.n_total = 0;
  for (size_t s = 0; s < n; ++s) g->counters.n_total += plans[s]->counters.n_total;
  g->counters.n_scat_in_grid = (int64_t)nt;
  g->counters.nnz = (int64_t)nnz;
  b->grid = g;
  *out = b;
  return MUTO_OK;
}

extern "C" int muto_batch_destroy(muto_batch* b) {
  if (!b) return MUTO_OK;
  if (b->grid) {
    (void)hipStreamSynchronize(b->ctx->stream);
    if (b->grid->f32.graph) (void)hipGraphExecDestroy(b->grid->f32.graph);
    for (int k = 0; k < 8; ++k)
      if (b->grid->ev[k]) (void)hipEventDestroy(b->grid->ev[k]);
    delete b->grid;
  }
  delete b;
  return MUTO_OK;
}

// the scans' images (their plans' current images) into the combined grid
static int batch_load_images(muto_batch* b) {
  if (!b) return MUTO_ERR_ARG;
  muto_ctx* ctx = b->ctx;
  muto_em_plan* g = b->grid;
  hipStream_t st = ctx->stream;
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(g->lam64.p, 0, g->nv * sizeof(double), st));
  for (size_t s = 0; s < b->scans.size(); ++s) {
    const muto_em_plan* p = b->scans[s];
    if (p->cur_prec < 0) { ctx->set_error("muto_batch_run: a scan has no image set"); return MUTO_ERR_STATE; }
    k_image_block<true><<<grid_for(p->nv, 256), 256, 0, st>>>(p->nv, p->cfg.grid.nx, p->cfg.grid.ny, b->xoff[s],
                                                              g->cfg.grid.nx, g->cfg.grid.ny, p->lam64.p,
                                                              g->lam64.p, nullptr, nullptr);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
  }
  g->cur_prec = MUTO_FP64;
  return MUTO_OK;
}

// `iters` FP32C mean-update iterations of every scan in lock step; the
// scans' images are read from their plans before and written back after.
// costs (may be NULL): iters x n_scans, scan fastest, each the scan's mean
// cost (MutoMLSD.cpp:202).
extern "C" int muto_batch_run(muto_batch* b, int iters, double* costs) {
  if (!b) return MUTO_ERR_ARG;
  muto_ctx* ctx = b->ctx;
  if (iters < 0) { ctx->set_error("muto_batch_run: iters < 0"); return MUTO_ERR_ARG; }
  muto_em_plan* g = b->grid;
  hipStream_t st = ctx->stream;
  if (int rc = batch_load_images(b)) return rc;
  if (int rc = fp32c_build(g)) return rc;
  if (!g->lam32.p) MUTO_CHECK_HIP(ctx, g->lam32.alloc(g->nv));
  g->is_map = false;
  g->cfg.early_stop = 0;
  g->cur_prec = MUTO_FP64;
  // the float mirror of the combined image
  {
    hipEvent_t e0 = g->ev[6], e1 = g->ev[7];
    if (int rc = fp32c_mirror(g)) return rc;
    MUTO_CHECK_HIP(ctx, hipEventRecord(e0, st));
    const size_t ns = b->scans.size();
    std::vector<double> sums((size_t)std::max(iters, 1) * ns, 0.0);
    int done = 0;
    if (int rc = fp32c_run(g, iters, sums.data(), &done)) return rc;
    MUTO_CHECK_HIP(ctx, hipEventRecord(e1, st));
    MUTO_CHECK_HIP(ctx, hipEventSynchronize(e1));
    float ms = 0.f;
    MUTO_CHECK_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
    b->wall_ms = ms;
    if (costs)
      for (int it = 0; it < done; ++it)
        for (size_t s = 0; s < ns; ++s)
          costs[it * ns + s] = sums[it * ns + s] / (double)b->scans[s]->n_trk;
  }
  // images back to the scans' plans
  for (size_t s = 0; s < b->scans.size(); ++s) {
    muto_em_plan* p = b->scans[s];
    k_image_block<false><<<grid_for(p->nv, 256), 256, 0, st>>>(p->nv, p->cfg.grid.nx, p->cfg.grid.ny, b->xoff[s],
                                                               g->cfg.grid.nx, g->cfg.grid.ny, g->lam64.p,
                                                               p->lam64.p, nullptr, nullptr);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
    p->cur_prec = MUTO_FP64;
  }
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(st));
  return MUTO_OK;
}

// wall time of the last muto_batch_run's iterations (ms, HIP events), and
// the combined problem's counters (entries, segments, tracks)
extern "C" int muto_batch_stats(const muto_batch* b, double* wall_ms, muto_counters* c) {
  if (!b) return MUTO_ERR_ARG;
  if (wall_ms) *wall_ms = b->wall_ms;
  if (c) {
    *c = b->grid->counters;
    c->n_segments = (int64_t)b->grid->f32.n_seg;
  }
  return MUTO_OK;
}
