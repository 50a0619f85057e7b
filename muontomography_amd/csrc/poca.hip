// PoCA on the GPU: per-ray closest-approach point, voxel assignment and the
// PoCA image (src/MutoPoCA.cpp:51-174).  One lane per ray; the Siddon legs
// add their weights into the image with FP64 atomics (HBM-resident image,
// coalesced per wavefront only by chance: the PoCA image is a set-up step,
// not an iteration).
#include <string.h>

#include "geom_dev.hpp"

namespace muto {
namespace {

struct PocaCounters {
  unsigned long long not_parallel, in_grid, out_of_grid;
};

// Counts of a 256-thread block added to dst[0..2] with one atomic each
// (every thread must call it).  Per-wavefront atomics on a single address
// serialise behind each other: 1e7 rays make 3e5 of them.
__device__ __forceinline__ void block_counts(unsigned long long* d0, bool c0,
                                             unsigned long long* d1, bool c1,
                                             unsigned long long* d2, uint32_t c2) {
  __shared__ uint32_t ws[4][3];
  const uint32_t w = threadIdx.x >> 6;
  const uint32_t a = (uint32_t)__popcll(__ballot(c0));
  const uint32_t b = (uint32_t)__popcll(__ballot(c1));
  uint32_t c = c2;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0) { ws[w][0] = a; ws[w][1] = b; ws[w][2] = c; }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t0 = 0, t1 = 0, t2 = 0;
    for (int k = 0; k < 4; ++k) { t0 += ws[k][0]; t1 += ws[k][1]; t2 += ws[k][2]; }
    if (d0 && t0) atomicAdd(d0, t0);
    if (d1 && t1) atomicAdd(d1, t1);
    if (d2 && t2) atomicAdd(d2, t2);
  }
}

__global__ void __launch_bounds__(256)
k_poca_points(GridD g, const muto_ray* __restrict__ rays, size_t n,
              double eps, int mode, double* __restrict__ pscat,
              uint32_t* __restrict__ vox, uint8_t* __restrict__ accept,
              PocaCounters* __restrict__ cnt) {
  __shared__ __align__(16) muto_ray sh[256];
  stage_rays<256>(rays, n, sh);
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool npar = false, acc = false;
  if (i < n) {
    const muto_ray r = sh[threadIdx.x];
    const double zmin_det = rays[0].pout[2], zmax_det = rays[0].pin[2];
    double P[3];
    npar = poca_point(r, eps, P);
    const uint32_t xs = x86_u32((P[0] - g.mn[0]) / g.d[0]);
    const uint32_t ys = x86_u32((P[1] - g.mn[1]) / g.d[1]);
    const uint32_t zs = x86_u32((P[2] - g.mn[2]) / g.d[2]);
    if (mode == 0)
      acc = !(g.n[0] <= xs || g.n[1] <= ys || g.n[2] <= zs ||
              P[2] < zmin_det || zmax_det < P[2]);
    else
      acc = mlsd_accept(g, P, eps, zmin_det, zmax_det);
    if (pscat) {
      pscat[3 * i] = P[0]; pscat[3 * i + 1] = P[1]; pscat[3 * i + 2] = P[2];
    }
    if (vox) { vox[3 * i] = xs; vox[3 * i + 1] = ys; vox[3 * i + 2] = zs; }
    if (accept) accept[i] = acc ? 1 : 0;
  }
  block_counts(&cnt->not_parallel, npar, &cnt->in_grid, acc, nullptr, 0u);
}

struct PocaCfgD {
  GridD g;
  int length_weighted, straight_line, has_energy;
  double angmag, eps;
};

template <bool CUT>
__global__ void __launch_bounds__(256)
k_poca_image(PocaCfgD c, const muto_ray* __restrict__ rays, size_t n,
             double* __restrict__ img, double* __restrict__ wimg,
             PocaCounters* __restrict__ cnt, uint32_t* __restrict__ cut_vox,
             double* __restrict__ cut_val) {
  __shared__ __align__(16) muto_ray sh[256];
  stage_rays<256>(rays, n, sh);
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool npar = false, acc = false;
  uint32_t nout = 0;
  const GridD& g = c.g;
  if (i < n) {
    const muto_ray r = sh[threadIdx.x];
    const double zmin_det = rays[0].pout[2], zmax_det = rays[0].pin[2];
    double P[3];
    npar = poca_point(r, c.eps, P);
    // getScatAngleSquare (MutoPoCA.cpp:165-174)
    const double axz = (atan2(r.din[0], r.din[2]) - atan2(r.dout[0], r.dout[2])) * c.angmag;
    const double ayz = (atan2(r.din[1], r.din[2]) - atan2(r.dout[1], r.dout[2])) * c.angmag;
    const double a2 = (axz * axz + ayz * ayz) / 2.0;
    const uint32_t xs = x86_u32((P[0] - g.mn[0]) / g.d[0]);
    const uint32_t ys = x86_u32((P[1] - g.mn[1]) / g.d[1]);
    const uint32_t zs = x86_u32((P[2] - g.mn[2]) / g.d[2]);
    acc = !(g.n[0] <= xs || g.n[1] <= ys || g.n[2] <= zs || P[2] < zmin_det ||
            zmax_det < P[2]);
    if (CUT) {  // PoCut: the value goes to the ray's record, not the image
      cut_vox[i] = acc ? (uint32_t)(xs + (size_t)g.n[0] * (ys + (size_t)g.n[1] * zs)) : 0xFFFFFFFFu;
      cut_val[i] = c.has_energy ? a2 * r.ein * r.ein : a2;
    }
    if (acc) {
      const size_t nx = g.n[0], ny = g.n[1];
      if (!CUT)
        atomicAdd(&img[xs + nx * (ys + ny * (size_t)zs)],
                  c.has_energy ? a2 * r.ein * r.ein : a2);
      auto add = [&](uint32_t x, uint32_t y, uint32_t z, double w) {
        if (x >= g.n[0] || y >= g.n[1] || z >= g.n[2]) { ++nout; return; }
        atomicAdd(&wimg[x + nx * (y + ny * (size_t)z)], w);
      };
      if (c.straight_line) {
        siddon_walk(g, r.pin, r.pout, [&](uint32_t x, uint32_t y, uint32_t z,
                                          double L) {
          add(x, y, z, c.length_weighted ? L : g.d[2]);
        });
      } else {
        siddon_walk(g, r.pin, P, [&](uint32_t x, uint32_t y, uint32_t z,
                                     double L) {
          add(x, y, z, c.length_weighted ? L : g.d[2]);
        });
        siddon_walk(g, P, r.pout, [&](uint32_t x, uint32_t y, uint32_t z,
                                      double L) {
          if (c.length_weighted)
            add(x, y, z, L);
          else if (x != xs || y != ys || z != zs)
            add(x, y, z, g.d[2]);
        });
      }
    }
  }
  block_counts(&cnt->not_parallel, npar, &cnt->in_grid, acc, &cnt->out_of_grid, nout);
}

// SLR (MutoSLR.cpp:47-103): theta^2 / 2 spread along the chord between
// the grid's top and bottom planes, one lane per ray.
__global__ void __launch_bounds__(256)
k_slr_image(PocaCfgD c, const muto_ray* __restrict__ rays, size_t n,
            double* __restrict__ img, double* __restrict__ wimg,
            PocaCounters* __restrict__ cnt) {
  __shared__ __align__(16) muto_ray sh[256];
  stage_rays<256>(rays, n, sh);
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t nout = 0;
  const GridD& g = c.g;
  if (i < n) {
    const muto_ray r = sh[threadIdx.x];
    const double axz = (atan2(r.din[0], r.din[2]) - atan2(r.dout[0], r.dout[2])) * c.angmag;
    const double ayz = (atan2(r.din[1], r.din[2]) - atan2(r.dout[1], r.dout[2])) * c.angmag;
    const double zlo = g.mn[2];
    const double zhi = g.mn[2] + g.d[2] * (double)g.n[2];
    const double tin = (zhi - r.pin[2]) / r.din[2];
    const double tout = (zlo - r.pout[2]) / r.dout[2];
    double pi[3], po[3], d[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      pi[k] = r.pin[k] + tin * r.din[k];
      po[k] = r.pout[k] + tout * r.dout[k];
      d[k] = pi[k] - po[k];
    }
    const double ltot = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    const double a2h = (axz * axz + ayz * ayz) * 0.5;
    const double e2 = r.ein * r.ein;
    // the path's voxel count (the weight of an unweighted entry is 1/k^2)
    uint64_t k = 0;
    if (!c.length_weighted)
      k = siddon_walk(g, pi, po, [](uint32_t, uint32_t, uint32_t, double) {});
    const size_t nx = g.n[0], ny = g.n[1];
    siddon_walk(g, pi, po, [&](uint32_t x, uint32_t y, uint32_t z, double L) {
      if (x >= g.n[0] || y >= g.n[1] || z >= g.n[2]) { ++nout; return; }
      const size_t t = x + nx * (y + ny * (size_t)z);
      double v;
      if (c.length_weighted) {
        atomicAdd(&wimg[t], L / g.d[2]);
        v = a2h * (L / ltot) * (L / ltot);
      } else {
        atomicAdd(&wimg[t], 1.0);
        v = a2h / (double)(k * k);
      }
      if (c.has_energy) v *= e2;
      atomicAdd(&img[t], v);
    });
  }
  block_counts(nullptr, false, nullptr, false, &cnt->out_of_grid, nout);
}

// PoCut per-voxel lists: counts, slots, then the cut sums.
__global__ void __launch_bounds__(256)
k_cut_count(size_t n, const uint32_t* __restrict__ vox, uint32_t* __restrict__ m) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && vox[i] != 0xFFFFFFFFu) atomicAdd(&m[vox[i]], 1u);
}

__global__ void __launch_bounds__(256)
k_cut_scatter(size_t n, const uint32_t* __restrict__ vox, const double* __restrict__ val,
              const uint64_t* __restrict__ off, uint32_t* __restrict__ fill,
              double* __restrict__ slot) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || vox[i] == 0xFFFFFFFFu) return;
  const uint32_t j = vox[i];
  slot[off[j] + atomicAdd(&fill[j], 1u)] = val[i];
}

// sum of the lowest thr of the sorted a[0..m) in ascending order, from 0
// (std::accumulate over the sorted list, MutoPoCut.cpp:112-116)
__device__ __forceinline__ double cut_sum(const double* a, uint32_t m, double pct) {
  uint32_t thr = x86_u32((double)m * pct);
  if (thr > m) thr = m;
  double sum = 0.0;
  for (uint32_t k = 0; k < thr; ++k) sum += a[k];
  return sum;
}

constexpr uint32_t CUT_SMALL = 16;   // lists sorted by one lane
constexpr uint32_t CUT_LDS = 4096;   // lists sorted by one block in LDS

// one lane per voxel: lists of up to CUT_SMALL values (insertion sort);
// longer lists are left to k_cut_big
__global__ void __launch_bounds__(256)
k_cut_small(size_t nv, const uint32_t* __restrict__ m, const uint64_t* __restrict__ off,
            const double* __restrict__ slot, const double* __restrict__ wimg, double pct,
            double eps, double* __restrict__ img) {
  size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nv) return;
  const uint32_t mj = m[j];
  if (mj > CUT_SMALL) return;
  double a[CUT_SMALL];
  const double* src = slot + off[j];
  for (uint32_t k = 0; k < mj; ++k) {
    const double v = src[k];
    uint32_t q = k;
    while (q > 0 && v < a[q - 1]) { a[q] = a[q - 1]; --q; }
    a[q] = v;
  }
  const double w = wimg[j];
  img[j] = w < eps ? 0 : cut_sum(a, mj, pct) / w;
}

// bitonic sort of a[0..np) (np a power of two) by one block
__device__ void cut_bitonic(double* a, uint32_t np) {
  for (uint32_t k = 2; k <= np; k <<= 1) {
    for (uint32_t jj = k >> 1; jj > 0; jj >>= 1) {
      for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) {
        const uint32_t l = i ^ jj;
        if (l > i) {
          const double x = a[i], y = a[l];
          const bool up = (i & k) == 0;
          if (up ? (y < x) : (x < y)) { a[i] = y; a[l] = x; }
        }
      }
      __syncthreads();
    }
  }
}

// one block per listed voxel (CUT_SMALL < M_j): sorted in LDS when M_j <=
// CUT_LDS, else in the global scratch (then the list holds one voxel per
// launch's block and the scratch has room for the longest list)
__global__ void __launch_bounds__(256)
k_cut_big(const uint32_t* __restrict__ list, const uint32_t* __restrict__ m,
          const uint64_t* __restrict__ off, const double* __restrict__ slot,
          const double* __restrict__ wimg, double pct, double eps,
          double* __restrict__ scratch, double* __restrict__ img) {
  __shared__ double a[CUT_LDS];
  const uint32_t j = list[blockIdx.x];
  const uint32_t mj = m[j];
  uint32_t np = 1;
  while (np < mj) np <<= 1;
  double* buf = mj > CUT_LDS ? scratch : a;
  const double* src = slot + off[j];
  for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) buf[i] = i < mj ? src[i] : INFINITY;
  __syncthreads();
  cut_bitonic(buf, np);
  if (threadIdx.x == 0) {
    const double w = wimg[j];
    img[j] = w < eps ? 0 : cut_sum(buf, mj, pct) / w;
  }
}

__global__ void k_poca_normalize(double* __restrict__ img,
                                 const double* __restrict__ wimg, size_t nv,
                                 double eps) {
  size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nv) return;
  const double w = wimg[j];
  img[j] = w < eps ? 0 : img[j] / w;  // MutoPoCA.cpp:130-132
}

}  // namespace
}  // namespace muto

using namespace muto;

static int check_grid(muto_ctx* ctx, const muto_grid* g, const char* fn) {
  if (!g || g->nx == 0 || g->ny == 0 || g->nz == 0) {
    ctx->set_error(std::string(fn) + ": empty or null grid");
    return MUTO_ERR_ARG;
  }
  return MUTO_OK;
}

extern "C" int muto_poca_points(muto_ctx* ctx, const muto_grid* grid,
                                size_t n, const muto_ray* rays, double eps,
                                int mode, double* p_scat, uint32_t* vox,
                                uint8_t* accept, muto_counters* counters) {
  if (!ctx) return MUTO_ERR_ARG;
  if (int rc = check_grid(ctx, grid, "muto_poca_points")) return rc;
  if (n && !rays) { ctx->set_error("muto_poca_points: null rays"); return MUTO_ERR_ARG; }
  if (mode != 0 && mode != 1) { ctx->set_error("muto_poca_points: mode must be 0 or 1"); return MUTO_ERR_ARG; }
  if (counters) memset(counters, 0, sizeof *counters);
  if (n == 0) return MUTO_OK;
  hipStream_t s = ctx->stream;
  DBuf<muto_ray> dr;
  DBuf<double> dp;
  DBuf<uint32_t> dv;
  DBuf<uint8_t> da;
  DBuf<PocaCounters> dc;
  MUTO_CHECK_HIP(ctx, dr.alloc(n));
  MUTO_CHECK_HIP(ctx, dc.alloc(1));
  if (p_scat) MUTO_CHECK_HIP(ctx, dp.alloc(3 * n));
  if (vox) MUTO_CHECK_HIP(ctx, dv.alloc(3 * n));
  if (accept) MUTO_CHECK_HIP(ctx, da.alloc(n));
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(dr.p, rays, n * sizeof(muto_ray), hipMemcpyHostToDevice, s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(dc.p, 0, sizeof(PocaCounters), s));
  ctx->mark_begin();
  k_poca_points<<<grid_for(n, 256), 256, 0, s>>>(to_dev(*grid), dr.p, n, eps, mode,
                                                 dp.p, dv.p, da.p, dc.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  ctx->mark_end();
  if (p_scat) MUTO_CHECK_HIP(ctx, hipMemcpyAsync(p_scat, dp.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost, s));
  if (vox) MUTO_CHECK_HIP(ctx, hipMemcpyAsync(vox, dv.p, 3 * n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  if (accept) MUTO_CHECK_HIP(ctx, hipMemcpyAsync(accept, da.p, n, hipMemcpyDeviceToHost, s));
  PocaCounters hc;
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(&hc, dc.p, sizeof hc, hipMemcpyDeviceToHost, s));
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  if (counters) {
    counters->n_total = (int64_t)n;
    counters->n_not_parallel = (int64_t)hc.not_parallel;
    counters->n_scat_in_grid = (int64_t)hc.in_grid;
  }
  return MUTO_OK;
}

extern "C" int muto_poca_reconstruct(muto_ctx* ctx, const muto_poca_cfg* cfg,
                                     size_t n, const muto_ray* rays,
                                     double* img, muto_counters* counters) {
  if (!ctx) return MUTO_ERR_ARG;
  if (!cfg || !img) { ctx->set_error("muto_poca_reconstruct: null argument"); return MUTO_ERR_ARG; }
  if (int rc = check_grid(ctx, &cfg->grid, "muto_poca_reconstruct")) return rc;
  if (n && !rays) { ctx->set_error("muto_poca_reconstruct: null rays"); return MUTO_ERR_ARG; }
  const size_t nv = n_vox(cfg->grid);
  if (counters) memset(counters, 0, sizeof *counters);
  if (n == 0) { memset(img, 0, nv * sizeof(double)); return MUTO_OK; }
  size_t ntot = cfg->total_num_rays > 0 ? (size_t)cfg->total_num_rays : n;
  if (ntot > n) ntot = n;  // the reference reads past the end (undefined)
  hipStream_t s = ctx->stream;
  DBuf<muto_ray> dr;
  DBuf<double> di, dw;
  DBuf<PocaCounters> dc;
  MUTO_CHECK_HIP(ctx, dr.alloc(ntot));
  MUTO_CHECK_HIP(ctx, di.alloc(nv));
  MUTO_CHECK_HIP(ctx, dw.alloc(nv));
  MUTO_CHECK_HIP(ctx, dc.alloc(1));
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(dr.p, rays, ntot * sizeof(muto_ray), hipMemcpyHostToDevice, s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(di.p, 0, nv * sizeof(double), s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(dw.p, 0, nv * sizeof(double), s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(dc.p, 0, sizeof(PocaCounters), s));
  PocaCfgD c;
  c.g = to_dev(cfg->grid);
  c.length_weighted = cfg->length_weighted;
  c.straight_line = cfg->straight_line;
  c.has_energy = cfg->has_energy;
  c.angmag = cfg->angle_unit_mag;
  c.eps = cfg->eps;
  ctx->mark_begin();
  k_poca_image<false><<<grid_for(ntot, 256), 256, 0, s>>>(c, dr.p, ntot, di.p, dw.p, dc.p,
                                                         nullptr, nullptr);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  k_poca_normalize<<<grid_for(nv, 256), 256, 0, s>>>(di.p, dw.p, nv, cfg->eps);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  ctx->mark_end();
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(img, di.p, nv * sizeof(double), hipMemcpyDeviceToHost, s));
  PocaCounters hc;
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(&hc, dc.p, sizeof hc, hipMemcpyDeviceToHost, s));
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  if (counters) {
    counters->n_total = (int64_t)ntot;
    counters->n_not_parallel = (int64_t)hc.not_parallel;
    counters->n_scat_in_grid = (int64_t)hc.in_grid;
    counters->n_out_of_grid = (int64_t)hc.out_of_grid;
  }
  if (cfg->log_results) {
    printf("==================================================\n");
    printf("|| PoCA reconstruction results                  ||\n");
    printf("==================================================\n");
    printf("number of rays processed: %zu\n", ntot);
    printf("number of unparalleled rays: %llu\n", hc.not_parallel);
    printf("number of interactions in voxels: %llu\n", hc.in_grid);
  }
  return MUTO_OK;
}

// the run-config copy of a PoCA-style call and the rays on the device
static int slr_like_setup(muto_ctx* ctx, const muto_poca_cfg* cfg, size_t n,
                          const muto_ray* rays, double* img, const char* fn,
                          size_t* ntot) {
  if (!cfg || !img) { ctx->set_error(std::string(fn) + ": null argument"); return MUTO_ERR_ARG; }
  if (int rc = check_grid(ctx, &cfg->grid, fn)) return rc;
  if (n && !rays) { ctx->set_error(std::string(fn) + ": null rays"); return MUTO_ERR_ARG; }
  size_t nt = cfg->total_num_rays > 0 ? (size_t)cfg->total_num_rays : n;
  if (nt > n) nt = n;  // the reference reads past the end (undefined)
  *ntot = nt;
  return MUTO_OK;
}

extern "C" int muto_pocut_reconstruct(muto_ctx* ctx, const muto_poca_cfg* cfg,
                                      double cut_percentile, size_t n,
                                      const muto_ray* rays, double* img,
                                      muto_counters* counters) {
  if (!ctx) return MUTO_ERR_ARG;
  size_t ntot = 0;
  if (int rc = slr_like_setup(ctx, cfg, n, rays, img, "muto_pocut_reconstruct", &ntot)) return rc;
  if (!(cut_percentile >= 0.0)) {
    ctx->set_error("muto_pocut_reconstruct: cut_percentile must be >= 0");
    return MUTO_ERR_ARG;
  }
  const size_t nv = n_vox(cfg->grid);
  if (counters) memset(counters, 0, sizeof *counters);
  if (ntot == 0) { memset(img, 0, nv * sizeof(double)); return MUTO_OK; }
  hipStream_t s = ctx->stream;
  DBuf<muto_ray> dr;
  DBuf<double> di, dw, val, slot, scratch;
  DBuf<uint32_t> vox, cnt, fill;
  DBuf<uint64_t> off;
  DBuf<PocaCounters> dc;
  MUTO_CHECK_HIP(ctx, dr.alloc(ntot));
  MUTO_CHECK_HIP(ctx, di.alloc(nv));
  MUTO_CHECK_HIP(ctx, dw.alloc(nv));
  MUTO_CHECK_HIP(ctx, val.alloc(ntot));
  MUTO_CHECK_HIP(ctx, vox.alloc(ntot));
  MUTO_CHECK_HIP(ctx, cnt.alloc(nv + 1));
  MUTO_CHECK_HIP(ctx, fill.alloc(nv));
  MUTO_CHECK_HIP(ctx, off.alloc(nv + 1));
  MUTO_CHECK_HIP(ctx, dc.alloc(1));
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(dr.p, rays, ntot * sizeof(muto_ray), hipMemcpyHostToDevice, s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(di.p, 0, nv * sizeof(double), s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(dw.p, 0, nv * sizeof(double), s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(cnt.p, 0, (nv + 1) * sizeof(uint32_t), s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(fill.p, 0, nv * sizeof(uint32_t), s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(dc.p, 0, sizeof(PocaCounters), s));
  PocaCfgD c;
  c.g = to_dev(cfg->grid);
  c.length_weighted = cfg->length_weighted;
  c.straight_line = 0;  // MutoPoCut forces the two-leg path (MutoPoCut.cpp:5,14)
  c.has_energy = cfg->has_energy;
  c.angmag = cfg->angle_unit_mag;
  c.eps = cfg->eps;
  ctx->mark_begin();
  k_poca_image<true><<<grid_for(ntot, 256), 256, 0, s>>>(c, dr.p, ntot, nullptr, dw.p, dc.p,
                                                        vox.p, val.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  k_cut_count<<<grid_for(ntot, 256), 256, 0, s>>>(ntot, vox.p, cnt.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  MUTO_CHECK_HIP(ctx, exclusive_scan_u32_to_u64(cnt.p, off.p, nv + 1, s));
  uint64_t nrec = 0;
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(&nrec, off.p + nv, sizeof nrec, hipMemcpyDeviceToHost, s));
  std::vector<uint32_t> hm(nv);
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(hm.data(), cnt.p, nv * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  uint32_t mmax = 0;
  for (uint32_t v : hm) mmax = v > mmax ? v : mmax;
  MUTO_CHECK_HIP(ctx, slot.alloc(nrec ? nrec : 1));
  if (mmax > CUT_LDS) {
    uint64_t np = 1;
    while (np < mmax) np <<= 1;
    MUTO_CHECK_HIP(ctx, scratch.alloc(np));
  }
  k_cut_scatter<<<grid_for(ntot, 256), 256, 0, s>>>(ntot, vox.p, val.p, off.p, fill.p, slot.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  k_cut_small<<<grid_for(nv, 256), 256, 0, s>>>(nv, cnt.p, off.p, slot.p, dw.p, cut_percentile,
                                                cfg->eps, di.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  if (mmax > CUT_SMALL) {
    // lists longer than one lane sorts: one block each; those above the LDS
    // capacity one launch at a time, sharing the scratch
    std::vector<uint32_t> big, huge;
    for (size_t j = 0; j < nv; ++j) {
      if (hm[j] > CUT_LDS) huge.push_back((uint32_t)j);
      else if (hm[j] > CUT_SMALL) big.push_back((uint32_t)j);
    }
    DBuf<uint32_t> dl;
    MUTO_CHECK_HIP(ctx, dl.alloc(big.size() + huge.size()));
    if (!big.empty())
      MUTO_CHECK_HIP(ctx, hipMemcpyAsync(dl.p, big.data(), big.size() * sizeof(uint32_t),
                                         hipMemcpyHostToDevice, s));
    if (!huge.empty())
      MUTO_CHECK_HIP(ctx, hipMemcpyAsync(dl.p + big.size(), huge.data(), huge.size() * sizeof(uint32_t),
                                         hipMemcpyHostToDevice, s));
    if (!big.empty()) {
      k_cut_big<<<(unsigned)big.size(), 256, 0, s>>>(dl.p, cnt.p, off.p, slot.p, dw.p, cut_percentile,
                                                     cfg->eps, nullptr, di.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    }
    for (size_t h = 0; h < huge.size(); ++h) {
      k_cut_big<<<1, 256, 0, s>>>(dl.p + big.size() + h, cnt.p, off.p, slot.p, dw.p, cut_percentile,
                                  cfg->eps, scratch.p, di.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    }
    MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));  // the host lists die here
  }
  ctx->mark_end();
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(img, di.p, nv * sizeof(double), hipMemcpyDeviceToHost, s));
  PocaCounters hc;
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(&hc, dc.p, sizeof hc, hipMemcpyDeviceToHost, s));
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  if (counters) {
    counters->n_total = (int64_t)ntot;
    counters->n_not_parallel = (int64_t)hc.not_parallel;
    counters->n_scat_in_grid = (int64_t)hc.in_grid;
    counters->n_out_of_grid = (int64_t)hc.out_of_grid;
  }
  if (cfg->log_results) {
    printf("==================================================\n");
    printf("|| PoCA with threshold for scattering angles    ||\n");
    printf("==================================================\n");
    printf("number of rays processed: %zu\n", ntot);
    printf("number of interactions in voxels: %llu\n", hc.in_grid);
  }
  return MUTO_OK;
}

extern "C" int muto_slr_reconstruct(muto_ctx* ctx, const muto_poca_cfg* cfg, size_t n,
                                    const muto_ray* rays, double* img,
                                    muto_counters* counters) {
  if (!ctx) return MUTO_ERR_ARG;
  size_t ntot = 0;
  if (int rc = slr_like_setup(ctx, cfg, n, rays, img, "muto_slr_reconstruct", &ntot)) return rc;
  const size_t nv = n_vox(cfg->grid);
  if (counters) memset(counters, 0, sizeof *counters);
  if (ntot == 0) { memset(img, 0, nv * sizeof(double)); return MUTO_OK; }
  hipStream_t s = ctx->stream;
  DBuf<muto_ray> dr;
  DBuf<double> di, dw;
  DBuf<PocaCounters> dc;
  MUTO_CHECK_HIP(ctx, dr.alloc(ntot));
  MUTO_CHECK_HIP(ctx, di.alloc(nv));
  MUTO_CHECK_HIP(ctx, dw.alloc(nv));
  MUTO_CHECK_HIP(ctx, dc.alloc(1));
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(dr.p, rays, ntot * sizeof(muto_ray), hipMemcpyHostToDevice, s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(di.p, 0, nv * sizeof(double), s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(dw.p, 0, nv * sizeof(double), s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(dc.p, 0, sizeof(PocaCounters), s));
  PocaCfgD c;
  c.g = to_dev(cfg->grid);
  c.length_weighted = cfg->length_weighted;
  c.straight_line = 1;
  c.has_energy = cfg->has_energy;
  c.angmag = cfg->angle_unit_mag;
  c.eps = cfg->eps;
  ctx->mark_begin();
  k_slr_image<<<grid_for(ntot, 256), 256, 0, s>>>(c, dr.p, ntot, di.p, dw.p, dc.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  k_poca_normalize<<<grid_for(nv, 256), 256, 0, s>>>(di.p, dw.p, nv, cfg->eps);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  ctx->mark_end();
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(img, di.p, nv * sizeof(double), hipMemcpyDeviceToHost, s));
  PocaCounters hc;
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(&hc, dc.p, sizeof hc, hipMemcpyDeviceToHost, s));
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  if (counters) {
    counters->n_total = (int64_t)ntot;
    counters->n_out_of_grid = (int64_t)hc.out_of_grid;
  }
  if (cfg->log_results) {
    printf("==================================================\n");
    printf("|| SLR reconstruction results                  ||\n");
    printf("==================================================\n");
    printf("number of rays processed: %zu\n", ntot);
  }
  return MUTO_OK;
}
