// Device geometry: Siddon walk, PoCA point, selection tests, angles.
//
// Bit-exactness contract with the reference's FP64 CPU arithmetic: every
// translation unit that includes this file is compiled with
// -ffp-contract=off (the reference has no FMA: x86-64 baseline, no -march),
// FP64 '/' and sqrt are IEEE correctly rounded on gfx950, and the
// association of every expression follows the reference source.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "muto_internal.hpp"

namespace muto {

// static_cast<unsigned>(double) as GCC emits it on x86-64 (cvttsd2si to
// 64 bits, keep the low 32).  gfx950's v_cvt_u32_f64 saturates instead, so
// the conversion goes through int64.
__device__ __forceinline__ uint32_t x86_u32(double v) {
  if (!(v >= -9223372036854775808.0 && v < 9223372036854775808.0)) return 0u;
  return (uint32_t)(uint64_t)(int64_t)v;
}

// std::max / std::min semantics (signed zeros, ties)
__device__ __forceinline__ double smax(double a, double b) {
  return (a < b) ? b : a;
}
__device__ __forceinline__ double smin(double a, double b) {
  return (b < a) ? b : a;
}

// Siddon voxel walk of segment p1 -> p2 (src/MutoSiddon.cpp:9-141) as a
// resumable cursor: begin() sets the walk up (returns false for a segment
// that misses the grid), run(m, emit) takes steps until m voxels have been
// recorded or the walk ends, calling emit(x, y, z, length) for every
// recorded voxel in order.  The walk is capped at nx+ny+nz+8 steps, which a
// finite ray never reaches (the reference has no cap).
struct SiddonCursor {
  double length, amax, alpha, alpha_last;
  double ax[3], da[3];
  int di[3];
  uint32_t x, y, z, it, cap;
  bool live;

  __device__ __forceinline__ bool begin(const GridD& g, const double* p1, const double* p2) {
    const double eps = 1.0e-6;
    live = false;
    double d21[3], a0[3], an[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      d21[a] = p2[a] - p1[a];
      a0[a] = -0.1;
      an[a] = 1.1;
    }
    length = sqrt((d21[0] * d21[0] + d21[1] * d21[1]) + d21[2] * d21[2]);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double hi = g.mn[a] + g.d[a] * (double)g.n[a];
      if (eps < fabs(d21[a])) {
        a0[a] = (g.mn[a] - p1[a]) / d21[a];
        an[a] = (hi - p1[a]) / d21[a];
      } else if (p1[a] < g.mn[a] || hi <= p1[a]) {
        return false;
      }
    }
    double lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const bool sw = an[a] < a0[a];
      lo[a] = sw ? an[a] : a0[a];
      hi[a] = sw ? a0[a] : an[a];
    }
    const double amin = smax(smax(smax(lo[0], lo[1]), lo[2]), 0.0);
    amax = smin(smin(smin(hi[0], hi[1]), hi[2]), 1.0);
    if (amax <= amin) return false;
    alpha = amin;
    uint32_t idx[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      idx[a] = x86_u32((alpha * d21[a] + p1[a] - g.mn[a]) / g.d[a] +
                       (d21[a] <= 0 ? -eps : eps));
      di[a] = 0;
      da[a] = 0.0;
      if (eps < fabs(d21[a])) {
        da[a] = g.d[a] / (fabs(d21[a]));
        di[a] = d21[a] < 0 ? -1 : 1;
      }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      ax[a] = 1.1;
      if (fabs(amin - lo[a]) < eps) {
        ax[a] = amin + da[a];
      } else if (di[a] != 0) {
        ax[a] = ((double)idx[a] * g.d[a] + g.mn[a] - p1[a]) / d21[a] +
                (di[a] == 1 ? da[a] : 0.0);
      }
    }
    alpha_last = alpha;
    it = 0;
    cap = g.n[0] + g.n[1] + g.n[2] + 8;
    x = idx[0];
    y = idx[1];
    z = idx[2];
    live = true;
    return true;
  }

  template <class Emit>
  __device__ __forceinline__ uint32_t run(uint32_t max_emit, Emit&& emit) {
    const double eps = 1.0e-6;
    uint32_t cnt = 0;
    while (live && cnt < max_emit) {
      if (!(alpha < amax && it < cap)) {
        live = false;
        break;
      }
      ++it;
      const uint32_t cx = x, cy = y, cz = z;
      if (ax[0] < ax[1] && ax[0] < ax[2]) {
        x += (uint32_t)di[0]; alpha = ax[0]; ax[0] += da[0];
      } else if (ax[1] < ax[2]) {
        y += (uint32_t)di[1]; alpha = ax[1]; ax[1] += da[1];
      } else {
        z += (uint32_t)di[2]; alpha = ax[2]; ax[2] += da[2];
      }
      const double dalpha = smin(alpha, amax) - alpha_last;
      if (eps < dalpha) {
        emit(cx, cy, cz, dalpha * length);
        ++cnt;
        alpha_last = alpha;
      }
    }
    return cnt;
  }

  // The same steps as run(), until emit(x, y, z, length) returns false
  // (that voxel is then not taken and the cursor is not used again) or the
  // walk ends.
  template <class Emit>
  __device__ __forceinline__ void run_while(Emit&& emit) {
    const double eps = 1.0e-6;
    while (live) {
      if (!(alpha < amax && it < cap)) {
        live = false;
        break;
      }
      ++it;
      const uint32_t cx = x, cy = y, cz = z;
      if (ax[0] < ax[1] && ax[0] < ax[2]) {
        x += (uint32_t)di[0]; alpha = ax[0]; ax[0] += da[0];
      } else if (ax[1] < ax[2]) {
        y += (uint32_t)di[1]; alpha = ax[1]; ax[1] += da[1];
      } else {
        z += (uint32_t)di[2]; alpha = ax[2]; ax[2] += da[2];
      }
      const double dalpha = smin(alpha, amax) - alpha_last;
      if (eps < dalpha) {
        if (!emit(cx, cy, cz, dalpha * length)) return;
        alpha_last = alpha;
      }
    }
  }

  // Inside emit(cx, cy, cz, ...) of run(): the plane crossings as they were
  // before the step that recorded (cx, cy, cz).  Only the stepped axis's
  // index moved, and its crossing was taken as alpha before the increment
  // (a step along an axis the segment does not cross moves nothing and
  // adds nothing).  alpha_last is still the value before that step.
  __device__ __forceinline__ void snapshot(uint32_t cx, uint32_t cy, uint32_t cz,
                                           double* ax_pre) const {
    ax_pre[0] = x != cx ? alpha : ax[0];
    ax_pre[1] = y != cy ? alpha : ax[1];
    ax_pre[2] = z != cz ? alpha : ax[2];
  }

  // Right after begin() on the same segment: continue from a snapshot, so
  // that the next recorded voxel is (cx, cy, cz) with the same arithmetic
  // as the walk that took the snapshot.  The step count follows from the
  // index moves (every step before the last moves one index by one), and
  // the loop test of the first step passes as it did there (alpha_last is
  // at most the alpha of the step before, which was below amax).
  __device__ __forceinline__ void resume(uint32_t cx, uint32_t cy, uint32_t cz,
                                         const double* ax_pre, double al) {
    it = (cx - x) * (uint32_t)di[0] + (cy - y) * (uint32_t)di[1] + (cz - z) * (uint32_t)di[2];
    x = cx;
    y = cy;
    z = cz;
    ax[0] = ax_pre[0];
    ax[1] = ax_pre[1];
    ax[2] = ax_pre[2];
    alpha = al;
    alpha_last = al;
  }
};

// The whole walk in one call; returns the number of recorded voxels.
template <class Emit>
__device__ __forceinline__ uint32_t siddon_walk(const GridD& g,
                                                const double* p1,
                                                const double* p2,
                                                Emit&& emit) {
  SiddonCursor c;
  if (!c.begin(g, p1, p2)) return 0;
  return c.run(0xFFFFFFFFu, emit);
}

// Eigen 3-vector dot: (x0*y0 + x1*y1) + x2*y2
__device__ __forceinline__ double dot3(const double* a, const double* b) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

// PoCA point (src/MutoPoCA.cpp:142-163); returns true if not parallel.
__device__ __forceinline__ bool poca_point(const muto_ray& r, double eps,
                                           double* out) {
  double w0[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) w0[k] = r.pin[k] - r.pout[k];
  const double a = dot3(r.din, r.din);
  const double b = dot3(r.din, r.dout);
  const double c = dot3(r.dout, r.dout);
  const double d = dot3(r.din, w0);
  const double e = dot3(r.dout, w0);
  const double den = a * c - b * b;
  if (fabs(den) < eps) {
#pragma unroll
    for (int k = 0; k < 3; ++k) out[k] = 0.5 * (r.pin[k] + r.pout[k]);
    return false;
  }
  const double cin = (b * e - c * d) / den;
  const double cout = (a * e - b * d) / den;
#pragma unroll
  for (int k = 0; k < 3; ++k)
    out[k] = 0.5 * (((r.pin[k] + cin * r.din[k]) + r.pout[k]) +
                    cout * r.dout[k]);
  return true;
}

// MLSD selection (src/MutoMLSD.cpp:298-301)
__device__ __forceinline__ bool mlsd_accept(const GridD& g, const double* P,
                                            double eps, double zmin_det,
                                            double zmax_det) {
  return !(P[0] < g.mn[0] + eps || g.mn[0] + g.d[0] * g.n[0] - eps < P[0] ||
           P[1] < g.mn[1] + eps || g.mn[1] + g.d[1] * g.n[1] - eps < P[1] ||
           P[2] < g.mn[2] + eps || g.mn[2] + g.d[2] * g.n[2] - eps < P[2] ||
           P[2] < zmin_det + eps || zmax_det - eps <= P[2]);
}

__device__ __forceinline__ void load_ray(const muto_ray* __restrict__ rays,
                                         size_t i, muto_ray& r) {
  const double* s = reinterpret_cast<const double*>(rays + i);
  double* d = reinterpret_cast<double*>(&r);
#pragma unroll
  for (int k = 0; k < 14; ++k) d[k] = s[k];
}

// The block's rays [blockIdx.x * BS, +BS) (those below n) copied to LDS with
// 16-byte loads that consecutive threads issue at consecutive addresses (a
// ray is 112 bytes, seven such words); every thread of the block must call
// it.  Each thread then reads its own ray from LDS instead of making 14
// strided 8-byte loads.
static_assert(sizeof(muto_ray) == 112, "a ray is seven 16-byte words");

template <unsigned BS>
__device__ __forceinline__ void stage_rays(const muto_ray* __restrict__ rays, size_t n,
                                           muto_ray* sh) {
  const size_t base = (size_t)blockIdx.x * BS;
  const size_t cnt = base < n ? (n - base < BS ? n - base : BS) : 0;
  const uint4* src = reinterpret_cast<const uint4*>(rays + base);
  uint4* dst = reinterpret_cast<uint4*>(sh);
  for (size_t u = threadIdx.x; u < cnt * 7; u += BS) dst[u] = src[u];
  __syncthreads();
}

}  // namespace muto
