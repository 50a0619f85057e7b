// Muon records to rays on the GPU: NaN filter and per-side principal-axis
// line fits (src/MutoFile.cpp:19-82, 248-304), one lane per record, then an
// order-keeping compaction of the records that were kept.
#include <math.h>
#include <string.h>

#include "muto_internal.hpp"

namespace muto {
namespace {

// Centroid and principal axis of n points by cyclic Jacobi rotations of
// their 3x3 scatter matrix (the first right singular vector of the centred
// data, src/MutoFile.cpp:270-271), sign d.z > 0; the point is the line at
// height zref (:272-273).
__device__ void fit_line(const double* x, const double* y, const double* z, int n,
                         double zref, double* point, double* dir) {
  double m[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < n; ++i) { m[0] += x[i]; m[1] += y[i]; m[2] += z[i]; }
  for (int k = 0; k < 3; ++k) m[k] /= n;
  double a[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int i = 0; i < n; ++i) {
    const double d[3] = {x[i] - m[0], y[i] - m[1], z[i] - m[2]};
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) a[r][c] += d[r] * d[c];
  }
  double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int sweep = 0; sweep < 64; ++sweep) {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    const double diag = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
    if (off <= 1e-30 * diag || off == 0.0) break;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
#pragma unroll
      for (int q = p + 1; q < 3; ++q) {
        if (a[p][q] == 0.0) continue;
        const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double vkp = v[k][p], vkq = v[k][q];
          v[k][p] = c * vkp - s * vkq;
          v[k][q] = s * vkp + c * vkq;
        }
      }
    }
  }
  int best = 0;
  if (a[1][1] > a[best][best]) best = 1;
  if (a[2][2] > a[best][best]) best = 2;
  double d[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) d[k] = best == 0 ? v[k][0] : best == 1 ? v[k][1] : v[k][2];
  const double nrm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
#pragma unroll
  for (int k = 0; k < 3; ++k) d[k] /= nrm;
  if (d[2] < 0.0)
#pragma unroll
    for (int k = 0; k < 3; ++k) d[k] = -d[k];
  const double t = (zref - m[2]) / d[2];
  dir[0] = d[0]; dir[1] = d[1]; dir[2] = d[2];
  point[0] = t * d[0] + m[0];
  point[1] = t * d[1] + m[1];
  point[2] = zref;
}

struct Layers {
  double z[MUTO_MAX_LAYERS];
  int n, mid;
};

template <class T>
__global__ void __launch_bounds__(256)
k_fit_records(size_t n_rec, Layers ly, const T* __restrict__ rec, muto_ray* __restrict__ out,
              uint32_t* __restrict__ keep) {
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rec) return;
  const size_t each = 2 * (size_t)ly.n + 2;
  const T* p = rec + r * each;
  bool ok = true;
  for (size_t k = 0; k < each; ++k) ok = ok && !isnan((double)p[k]);
  keep[r] = ok ? 1u : 0u;
  if (!ok) return;
  double xs[MUTO_MAX_LAYERS], ys[MUTO_MAX_LAYERS];
  for (int i = 0; i < ly.n; ++i) {
    xs[i] = (double)p[1 + 2 * i];
    ys[i] = (double)p[2 + 2 * i];
  }
  muto_ray ray;
  ray.ein = (double)p[0];
  fit_line(xs, ys, ly.z, ly.mid, ly.z[ly.mid - 1], ray.pin, ray.din);
  fit_line(xs + ly.mid, ys + ly.mid, ly.z + ly.mid, ly.n - ly.mid, ly.z[ly.mid], ray.pout,
           ray.dout);
  ray.eout = (double)p[2 * ly.n + 1];
  out[r] = ray;
}

__global__ void __launch_bounds__(256)
k_keep_rays(size_t n, const uint32_t* __restrict__ keep, const uint64_t* __restrict__ pos,
            const muto_ray* __restrict__ in, muto_ray* __restrict__ out) {
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n && keep[r]) out[pos[r]] = in[r];
}

}  // namespace
}  // namespace muto

using namespace muto;

extern "C" int muto_rays_from_records(muto_ctx* ctx, size_t n_rec, int n_layers, int n_middle,
                                      const double* z, const void* records, int precision,
                                      muto_ray* rays, size_t* n_rays) {
  if (!ctx) return MUTO_ERR_ARG;
  if (!z || !n_rays || (n_rec && (!records || !rays))) {
    ctx->set_error("muto_rays_from_records: null argument");
    return MUTO_ERR_ARG;
  }
  if (precision != 4 && precision != 8) {
    ctx->set_error("muto_rays_from_records: precision must be 4 or 8");
    return MUTO_ERR_ARG;
  }
  if (n_middle < 2 || n_layers - n_middle < 2 || n_layers > MUTO_MAX_LAYERS) {
    ctx->set_error("muto_rays_from_records: at least two layers above and below, at most "
                   "MUTO_MAX_LAYERS in all");
    return MUTO_ERR_ARG;
  }
  *n_rays = 0;
  if (n_rec == 0) return MUTO_OK;
  Layers ly;
  ly.n = n_layers;
  ly.mid = n_middle;
  for (int i = 0; i < MUTO_MAX_LAYERS; ++i) ly.z[i] = i < n_layers ? z[i] : 0.0;
  hipStream_t s = ctx->stream;
  const size_t each = 2 * (size_t)n_layers + 2;
  DBuf<unsigned char> drec;
  DBuf<muto_ray> tmp, outd;
  DBuf<uint32_t> keep;
  DBuf<uint64_t> pos;
  MUTO_CHECK_HIP(ctx, drec.alloc(n_rec * each * (size_t)precision));
  MUTO_CHECK_HIP(ctx, tmp.alloc(n_rec));
  MUTO_CHECK_HIP(ctx, keep.alloc(n_rec + 1));
  MUTO_CHECK_HIP(ctx, pos.alloc(n_rec + 1));
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(drec.p, records, n_rec * each * (size_t)precision,
                                     hipMemcpyHostToDevice, s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(keep.p + n_rec, 0, sizeof(uint32_t), s));
  ctx->mark_begin();
  if (precision == 8)
    k_fit_records<double><<<grid_for(n_rec, 256), 256, 0, s>>>(
        n_rec, ly, reinterpret_cast<const double*>(drec.p), tmp.p, keep.p);
  else
    k_fit_records<float><<<grid_for(n_rec, 256), 256, 0, s>>>(
        n_rec, ly, reinterpret_cast<const float*>(drec.p), tmp.p, keep.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  MUTO_CHECK_HIP(ctx, exclusive_scan_u32_to_u64(keep.p, pos.p, n_rec + 1, s));
  uint64_t kept = 0;
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(&kept, pos.p + n_rec, sizeof kept, hipMemcpyDeviceToHost, s));
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  MUTO_CHECK_HIP(ctx, outd.alloc(kept ? kept : 1));
  k_keep_rays<<<grid_for(n_rec, 256), 256, 0, s>>>(n_rec, keep.p, pos.p, tmp.p, outd.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  ctx->mark_end();
  if (kept)
    MUTO_CHECK_HIP(ctx, hipMemcpyAsync(rays, outd.p, kept * sizeof(muto_ray), hipMemcpyDeviceToHost, s));
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  *n_rays = (size_t)kept;
  return MUTO_OK;
}
