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

__device__ __forceinline__ void wave_count(unsigned long long* dst, bool v) {
  // one atomic per wavefront
  unsigned long long m = __ballot(v);
  if (m && (threadIdx.x & 63) == (unsigned)(__ffsll((long long)m) - 1))
    atomicAdd(dst, (unsigned long long)__popcll(m));
}

__global__ void __launch_bounds__(256)
k_poca_points(GridD g, const muto_ray* __restrict__ rays, size_t n,
              double eps, int mode, double* __restrict__ pscat,
              uint32_t* __restrict__ vox, uint8_t* __restrict__ accept,
              PocaCounters* __restrict__ cnt) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool npar = false, acc = false;
  if (i < n) {
    muto_ray r;
    load_ray(rays, i, r);
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
  wave_count(&cnt->not_parallel, npar);
  wave_count(&cnt->in_grid, acc);
}

struct PocaCfgD {
  GridD g;
  int length_weighted, straight_line, has_energy;
  double angmag, eps;
};

__global__ void __launch_bounds__(256)
k_poca_image(PocaCfgD c, const muto_ray* __restrict__ rays, size_t n,
             double* __restrict__ img, double* __restrict__ wimg,
             PocaCounters* __restrict__ cnt) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool npar = false, acc = false;
  uint32_t nout = 0;
  const GridD& g = c.g;
  if (i < n) {
    muto_ray r;
    load_ray(rays, i, r);
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
    if (acc) {
      const size_t nx = g.n[0], ny = g.n[1];
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
  wave_count(&cnt->not_parallel, npar);
  wave_count(&cnt->in_grid, acc);
  if (__any(nout != 0) && nout) atomicAdd(&cnt->out_of_grid, (unsigned long long)nout);
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
  k_poca_image<<<grid_for(ntot, 256), 256, 0, s>>>(c, dr.p, ntot, di.p, dw.p, dc.p);
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
