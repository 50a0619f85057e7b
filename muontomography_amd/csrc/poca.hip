// PoCA on the GPU: per-ray closest-approach point, voxel assignment and the
// PoCA image (src/MutoPoCA.cpp:51-174).  One lane per ray; the Siddon legs
// add their weights into the image with FP64 atomics (HBM-resident image,
// coalesced per wavefront only by chance: the PoCA image is a set-up step,
// not an iteration).
#include <string.h>

#include <algorithm>

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

// j appended to list[] where flag holds (one counter atomic per wavefront)
__device__ __forceinline__ void wave_append(bool flag, uint32_t* ctr, uint32_t* list, uint32_t j) {
  const unsigned long long b = __ballot(flag);
  if (!b) return;
  const int lane = (int)(threadIdx.x & 63u);
  const int lead = __ffsll((long long)b) - 1;
  uint32_t base = 0;
  if (lane == lead) base = atomicAdd(ctr, (uint32_t)__popcll(b));
  base = (uint32_t)__shfl((int)base, lead, 64);
  if (flag) list[base + (uint32_t)__popcll(b & ((1ull << lane) - 1ull))] = j;
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

// the voxels with lists above CUT_SMALL (ctr[1] of them in big) and above
// CUT_LDS (ctr[2] in huge), and the longest of the latter (ctr[0]: the
// scratch size; a max over every voxel would put one atomic per wavefront
// on one address)
__global__ void __launch_bounds__(256)
k_cut_lists(size_t nv, const uint32_t* __restrict__ m, uint32_t* __restrict__ ctr,
            uint32_t* __restrict__ big, uint32_t* __restrict__ huge) {
  const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t mj = j < nv ? m[j] : 0u;
  if (mj > CUT_LDS) atomicMax(&ctr[0], mj);
  wave_append(mj > CUT_SMALL && mj <= CUT_LDS, &ctr[1], big, (uint32_t)j);
  wave_append(mj > CUT_LDS, &ctr[2], huge, (uint32_t)j);
}

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

// ---------------- weights privatised per tile -----------------------------
// The PoCA, PoCut and SLR images add one value per Siddon entry into a
// voxel image.  Scattered FP64 global atomics run at about 2.4e10/s on this
// GPU (a microbenchmark: the same with per-XCD copies of the image, with
// workgroup scope, or with two values per line), against some 1e12/s for
// LDS atomics.  So the sums are made per tile of TILE_E^3 voxels in LDS:
//   1. the rays in a spatial order (direction bin of din, then the Morton
//      code of pin), so that the rays crossing one tile come in runs;
//   2. every leg walked once (k_*_split); where a recorded voxel lies in a
//      tile other than the previous recorded voxel's, the walk's state
//      before that step is kept as a record (SiddonCursor::snapshot), at a
//      slot of the ray's range (a bound per leg, tile_visit_bound);
//   3. the records sorted by tile, stably (in ray order inside a tile);
//   4. per task of up to TILE_TASK records of one tile, one block resumes
//      each record's walk (begin() on the same leg, then
//      SiddonCursor::resume) and adds the entries that stay in the tile in
//      LDS, then adds the tile into the image (stores it when one task
//      covers the tile).
// The voxels and lengths are the walk's own, bit for bit; only the order of
// the FP64 sums differs from one atomic per entry.
constexpr uint32_t TILE_E = 16;
constexpr uint32_t TILE_V = TILE_E * TILE_E * TILE_E;
constexpr uint32_t TILE_BS = 256;
constexpr uint32_t TILE_TASK = 16384;

struct TileGeo {
  uint32_t nt[3];
  uint32_t n_tiles;
};

__device__ __forceinline__ uint32_t tile_of(const TileGeo& T, uint32_t x, uint32_t y, uint32_t z) {
  return x / TILE_E + T.nt[0] * (y / TILE_E + T.nt[1] * (z / TILE_E));
}

// the walk's state before the step that recorded voxel `vox` of leg `leg`
struct TileRec {
  double ax[3];
  double alpha_last;
  uint32_t leg;  // 2 ray + leg, the ray in the spatial order
  uint32_t vox;  // linear index of the voxel
};
static_assert(sizeof(TileRec) == 40, "tile record");

struct TileTask {
  uint32_t tile, q0, q1, ntask;  // records [q0, q1) of the sorted list; tasks of the tile
};

// Upper bound on the tile visits of the walk of p1 -> p2.  Its recorded
// voxels lie inside the segment's bounding box widened by two voxels per
// side, and each index moves one way only, so the tile changes at most once
// per tile boundary per axis inside that range.  (A violation is caught by
// the split kernel, which then leaves the call to the untiled kernel.)
__device__ __forceinline__ uint32_t tile_visit_bound(const GridD& g, const double* p1, const double* p2) {
  uint32_t b = 1;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double hi_v = (double)(g.n[a] - 1);
    double lo = floor((fmin(p1[a], p2[a]) - g.mn[a]) / g.d[a]) - 2.0;
    double hi = floor((fmax(p1[a], p2[a]) - g.mn[a]) / g.d[a]) + 2.0;
    lo = fmin(fmax(lo, 0.0), hi_v);
    hi = fmin(fmax(hi, 0.0), hi_v);
    b += (uint32_t)hi / TILE_E - (uint32_t)lo / TILE_E;
  }
  return b;
}

__device__ __forceinline__ uint32_t interleave12(uint32_t v) {
  v &= 0xFFFu;
  v = (v | (v << 8)) & 0x00FF00FFu;
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

// Spatial key of every ray: direction bin (1/8 in din.x/din.z and
// din.y/din.z over [-2, 2]) above the Morton code of pin in voxel units.
// KIND 0 (PoCA, PoCut): the rays whose PoCA point is accepted come first
// (the same test as k_poca_prep), so the walks run over a dense prefix.
template <int KIND>
__global__ void __launch_bounds__(256)
k_ray_keys(PocaCfgD c, const muto_ray* __restrict__ rays, size_t n, uint64_t* __restrict__ key,
           uint32_t* __restrict__ val) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const GridD& g = c.g;
  muto_ray r;
  load_ray(rays, i, r);
  const double dz = fabs(r.din[2]) > 1e-12 ? r.din[2] : 1e-12;
  const double sx = fmin(fmax(r.din[0] / dz, -2.0), 2.0), sy = fmin(fmax(r.din[1] / dz, -2.0), 2.0);
  const uint32_t dbin = (uint32_t)((sx + 2.0) * 8.0 + 0.5) * 33u + (uint32_t)((sy + 2.0) * 8.0 + 0.5);
  const double xt = fmin(fmax((r.pin[0] - g.mn[0]) / g.d[0] + 1024.0, 0.0), 4095.0);
  const double yt = fmin(fmax((r.pin[1] - g.mn[1]) / g.d[1] + 1024.0, 0.0), 4095.0);
  uint64_t k = ((uint64_t)dbin << 24) | (interleave12((uint32_t)xt) | (interleave12((uint32_t)yt) << 1));
  if (KIND == 0) {
    const double zmin_det = rays[0].pout[2], zmax_det = rays[0].pin[2];
    double P[3];
    (void)poca_point(r, c.eps, P);
    const uint32_t xs = x86_u32((P[0] - g.mn[0]) / g.d[0]);
    const uint32_t ys = x86_u32((P[1] - g.mn[1]) / g.d[1]);
    const uint32_t zs = x86_u32((P[2] - g.mn[2]) / g.d[2]);
    const bool acc = !(g.n[0] <= xs || g.n[1] <= ys || g.n[2] <= zs || P[2] < zmin_det ||
                       zmax_det < P[2]);
    if (!acc) k |= 1ull << 35;  // above the 24 Morton and 11 direction bits
  }
  key[i] = k;
  val[i] = (uint32_t)i;
}

__global__ void k_fill_u32(size_t n, uint32_t v, uint32_t* __restrict__ a) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}
__global__ void k_iota_u32(size_t n, uint32_t* __restrict__ a) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = (uint32_t)i;
}

// PoCA / PoCut per ray in the spatial order (MutoPoCA.cpp:84-121): the
// PoCA point and acceptance, the image's value at the PoCA voxel (PoCut:
// the ray's record instead), the legs' end points and the record bound.
template <bool CUT>
__global__ void __launch_bounds__(256)
k_poca_prep(PocaCfgD c, const muto_ray* __restrict__ rays, size_t n,
            const uint32_t* __restrict__ order, double* __restrict__ img,
            PocaCounters* __restrict__ cnt, uint32_t* __restrict__ cut_vox,
            double* __restrict__ cut_val, double* __restrict__ legs,
            uint32_t* __restrict__ pvox, uint32_t* __restrict__ bound) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool npar = false, acc = false;
  const GridD& g = c.g;
  if (i < n) {
    const uint32_t ri = order[i];
    muto_ray r;
    load_ray(rays, ri, r);
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
    const uint32_t lin = acc ? (uint32_t)(xs + (size_t)g.n[0] * (ys + (size_t)g.n[1] * zs)) : 0xFFFFFFFFu;
    if (CUT) {
      cut_vox[ri] = lin;
      cut_val[ri] = c.has_energy ? a2 * r.ein * r.ein : a2;
    }
    uint32_t b = 0;
    if (acc) {
      if (!CUT) atomicAdd(&img[lin], c.has_energy ? a2 * r.ein * r.ein : a2);
      b = c.straight_line ? tile_visit_bound(g, r.pin, r.pout)
                          : tile_visit_bound(g, r.pin, P) + tile_visit_bound(g, P, r.pout);
    }
    double* L = legs + 9 * i;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      L[k] = r.pin[k];
      L[3 + k] = P[k];
      L[6 + k] = r.pout[k];
    }
    pvox[i] = lin;
    bound[i] = b;
  }
  block_counts(&cnt->not_parallel, npar, &cnt->in_grid, acc, nullptr, 0u);
}

// SLR per ray in the spatial order (MutoSLR.cpp:47-103): the chord between
// the grid's top and bottom planes, its length, theta^2 / 2 and the energy
// factor, and the record bound.  chord: pi, po, ltot, a2h, e2.
__global__ void __launch_bounds__(256)
k_slr_prep(PocaCfgD c, const muto_ray* __restrict__ rays, size_t n,
           const uint32_t* __restrict__ order, double* __restrict__ chord,
           uint32_t* __restrict__ bound) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const GridD& g = c.g;
  muto_ray r;
  load_ray(rays, order[i], r);
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
  double* C = chord + 9 * i;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    C[k] = pi[k];
    C[3 + k] = po[k];
  }
  C[6] = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
  C[7] = (axz * axz + ayz * ayz) * 0.5;
  C[8] = r.ein * r.ein;
  bound[i] = tile_visit_bound(g, pi, po);
}

// The leg l of sorted ray i: PoCA legs pin -> P, P -> pout (or pin -> pout
// along a straight line) from {pin, P, pout}; the SLR chord pi -> po.
template <int KIND>
__device__ __forceinline__ void leg_ends(const double* T, int straight, uint32_t l,
                                         const double*& p1, const double*& p2) {
  if (KIND == 1) {
    p1 = T;
    p2 = T + 3;
  } else {
    p1 = l == 0 ? T : T + 3;
    p2 = (straight || l == 1) ? T + 6 : T + 3;
  }
}

// Step 2: each ray's legs walked once.  Records go to the ray's slots
// [roff[i], roff[i + 1]); SLR also keeps the chord's voxel count (every
// recorded voxel, the 1/k^2 weight's k).
// (KIND 0 runs over the accepted rays, the first n in the order.)
template <int KIND>
__global__ void __launch_bounds__(256)
k_tile_split(GridD g, TileGeo T, int straight, size_t n, const double* __restrict__ tab,
             const uint64_t* __restrict__ roff, TileRec* __restrict__ rec,
             uint32_t* __restrict__ rkey, uint64_t* __restrict__ kcount,
             PocaCounters* __restrict__ cnt, unsigned int* __restrict__ overflow) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t nout = 0;
  bool ovf = false;
  if (i < n) {
    const uint64_t b = roff[i], cap = roff[i + 1] - b;
    uint64_t k = 0, ne = 0;
    const double* tr = tab + 9 * i;
    const int nl = (KIND == 1 || straight) ? 1 : 2;
    for (int l = 0; l < nl; ++l) {
      const double *p1, *p2;
      leg_ends<KIND>(tr, straight, (uint32_t)l, p1, p2);
      SiddonCursor c;
      if (!c.begin(g, p1, p2)) continue;
      uint32_t cur = 0xFFFFFFFFu;
      c.run(0xFFFFFFFFu, [&](uint32_t x, uint32_t y, uint32_t z, double) {
        ++ne;
        if (x >= g.n[0] || y >= g.n[1] || z >= g.n[2]) {
          ++nout;
          return;
        }
        const uint32_t t = tile_of(T, x, y, z);
        if (t == cur) return;
        cur = t;
        if (k < cap) {
          TileRec R;
          c.snapshot(x, y, z, R.ax);
          R.alpha_last = c.alpha_last;
          R.leg = (uint32_t)(2 * i + (size_t)l);
          R.vox = x + g.n[0] * (y + g.n[1] * z);
          rec[b + k] = R;
          rkey[b + k] = t;
        } else {
          ovf = true;
        }
        ++k;
      });
    }
    if (KIND == 1) kcount[i] = ne;
  }
  block_counts(nullptr, false, nullptr, false, &cnt->out_of_grid, nout);
  if (ovf) atomicOr(overflow, 1u);
}

// first and end record of every tile in the sorted keys
__global__ void k_tile_bounds(size_t n, const uint32_t* __restrict__ key, uint32_t n_tiles,
                              uint32_t* __restrict__ ts, uint32_t* __restrict__ te) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const uint32_t t = key[q];
  if (t >= n_tiles) return;
  if (q == 0 || key[q - 1] != t) ts[t] = (uint32_t)q;
  if (q + 1 == n || key[q + 1] != t) te[t] = (uint32_t)(q + 1);
}

__global__ void k_tile_ntask(uint32_t n_tiles, const uint32_t* __restrict__ ts,
                             const uint32_t* __restrict__ te, uint32_t* __restrict__ nt) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n_tiles) nt[t] = (te[t] - ts[t] + TILE_TASK - 1) / TILE_TASK;
}

// a tile's records cut into its tasks of nearly equal size
__global__ void k_tile_tasks(uint32_t n_tiles, const uint32_t* __restrict__ ts,
                             const uint32_t* __restrict__ te, const uint64_t* __restrict__ toff,
                             TileTask* __restrict__ tasks) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tiles) return;
  const uint32_t k = (uint32_t)(toff[t + 1] - toff[t]);
  if (k == 0) return;
  const uint32_t m = te[t] - ts[t], per = (m + k - 1) / k;
  for (uint32_t j = 0; j < k; ++j) {
    TileTask tk;
    tk.tile = t;
    tk.q0 = ts[t] + j * per;
    tk.q1 = min(ts[t] + (j + 1) * per, te[t]);
    tk.ntask = k;
    tasks[toff[t] + j] = tk;
  }
}

// Step 4, persistent blocks over the tasks.  KIND 0 (PoCA, PoCut): the
// weight image (length, or the voxel height, per entry; without length
// weights the out-leg skips the PoCA voxel).  KIND 1 (SLR): the weight
// image and the value image.
template <int KIND>
__global__ void __launch_bounds__(TILE_BS)
k_tile_accum(GridD g, TileGeo T, int straight, int length_weighted, int has_energy,
             const TileTask* __restrict__ tasks, const uint64_t* __restrict__ n_tasks_p,
             unsigned int* __restrict__ work, const uint32_t* __restrict__ sval,
             const TileRec* __restrict__ rec, const double* __restrict__ tab,
             const uint32_t* __restrict__ pvox, const uint64_t* __restrict__ kcount,
             double* __restrict__ wimg, double* __restrict__ img) {
  __shared__ double w_s[TILE_V];
  __shared__ double v_s[KIND == 1 ? TILE_V : 1];
  __shared__ uint32_t sh;
  const uint32_t t = threadIdx.x;
  const uint32_t n_tasks = (uint32_t)n_tasks_p[0];
  for (;;) {
    if (t == 0) sh = atomicAdd(work, 1u);
    for (uint32_t v = t; v < TILE_V; v += TILE_BS) {
      w_s[v] = 0.0;
      if (KIND == 1) v_s[v] = 0.0;
    }
    __syncthreads();
    const uint32_t ti = sh;
    if (ti >= n_tasks) break;
    const TileTask tk = tasks[ti];
    const uint32_t tyz = tk.tile / T.nt[0];
    const uint32_t o0 = (tk.tile % T.nt[0]) * TILE_E, o1 = (tyz % T.nt[1]) * TILE_E,
                   o2 = (tyz / T.nt[1]) * TILE_E;
    for (uint32_t q = tk.q0 + t; q < tk.q1; q += TILE_BS) {
      const TileRec R = rec[sval[q]];
      const uint32_t i = R.leg >> 1, l = R.leg & 1u;
      const double* tr = tab + 9 * (size_t)i;
      const double *p1, *p2;
      leg_ends<KIND>(tr, straight, l, p1, p2);
      SiddonCursor c;
      c.begin(g, p1, p2);  // succeeded when the record was taken
      const uint32_t vr = R.vox / g.n[0];
      c.resume(R.vox - vr * g.n[0], vr % g.n[1], vr / g.n[1], R.ax, R.alpha_last);
      const uint32_t skip = (KIND == 0 && !length_weighted && l == 1) ? pvox[i] : 0xFFFFFFFFu;
      // SLR: the ray's factor of an entry's value, theta^2 / 2 (ltot^-2 or
      // k^-2, times e^2 with energies): the value is f L^2 (length
      // weights) or f, and the weight sums L / dz as (sum L) / dz
      double f = 0.0;
      if (KIND == 1) {
        const double ltot = tr[6], a2h = tr[7], e2 = tr[8];
        const uint64_t kk = kcount[i];
        f = length_weighted ? a2h / (ltot * ltot) : a2h / (double)(kk * kk);
        if (has_energy) f *= e2;
      }
      c.run_while([&](uint32_t x, uint32_t y, uint32_t z, double L) -> bool {
        const uint32_t lx = x - o0, ly = y - o1, lz = z - o2;
        if (lx >= TILE_E || ly >= TILE_E || lz >= TILE_E || x >= g.n[0] || y >= g.n[1] || z >= g.n[2])
          return false;  // the next record of the leg, or the leg's end
        const uint32_t lv = lx + TILE_E * (ly + TILE_E * lz);
        if (KIND == 0) {
          if (x + g.n[0] * (y + g.n[1] * z) == skip) return true;
          atomicAdd(&w_s[lv], length_weighted ? L : g.d[2]);
        } else {
          atomicAdd(&w_s[lv], length_weighted ? L : 1.0);
          atomicAdd(&v_s[lv], length_weighted ? f * (L * L) : f);
        }
        return true;
      });
    }
    __syncthreads();
    for (uint32_t v = t; v < TILE_V; v += TILE_BS) {
      const uint32_t x = o0 + v % TILE_E, y = o1 + (v / TILE_E) % TILE_E, z = o2 + v / (TILE_E * TILE_E);
      if (x >= g.n[0] || y >= g.n[1] || z >= g.n[2]) continue;
      const size_t j = x + (size_t)g.n[0] * (y + (size_t)g.n[1] * z);
      const double wv = (KIND == 1 && length_weighted) ? w_s[v] / g.d[2] : w_s[v];
      if (tk.ntask == 1) {
        wimg[j] = wv;
        if (KIND == 1) img[j] = v_s[v];
      } else {
        if (wv != 0.0) atomicAdd(&wimg[j], wv);
        if (KIND == 1 && v_s[v] != 0.0) atomicAdd(&img[j], v_s[v]);
      }
    }
    __syncthreads();
  }
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

static int bits_of(uint64_t v) {
  int b = 0;
  while ((1ull << b) <= v && b < 63) ++b;
  return b;
}

// Bump allocation in a workspace (a first pass with base == nullptr only
// measures).
struct Arena {
  char* base = nullptr;
  size_t off = 0;
  template <class T>
  T* take(size_t n) {
    off = (off + 255) & ~(size_t)255;
    T* p = reinterpret_cast<T*>(base + off);
    off += (n ? n : 1) * sizeof(T);
    return p;
  }
};

// The weight image (KIND 0: PoCA, PoCut) or the weight and value images
// (KIND 1: SLR) of the rays dr[0..n) by tiles (the kernels above), into dw
// and di.  PoCA also adds the PoCA voxels' values to di, PoCut writes the
// rays' records.  The buffers come from the context's workspaces, so a call
// allocates only when it needs more than an earlier one.  Sets *fallback
// when the records would not fit the sort or a record bound was exceeded:
// the caller then clears its images and counters and runs the kernel with
// one atomic per entry.
template <int KIND>
static int tiled_images(muto_ctx* ctx, const PocaCfgD& c, const muto_ray* dr, size_t n,
                        double* di, double* dw, PocaCounters* dc, bool cut, uint32_t* cut_vox,
                        double* cut_val, bool* fallback) {
  *fallback = false;
  hipStream_t s = ctx->stream;
  const GridD& g = c.g;
  TileGeo T;
  for (int a = 0; a < 3; ++a) T.nt[a] = (g.n[a] + TILE_E - 1) / TILE_E;
  const uint64_t ntl = (uint64_t)T.nt[0] * T.nt[1] * T.nt[2];
  if (ntl >= 0x7FFFFFFFull || n >= 0x7FFFFFFFull) { *fallback = true; return MUTO_OK; }
  T.n_tiles = (uint32_t)ntl;
  // per ray: keys and order (and their sort buffers), legs / chord, record
  // bounds and offsets, SLR voxel counts
  const int key_bits = 24 + bits_of(33 * 33) + (KIND == 0 ? 1 : 0);  // 35, and PoCA's accept bit
  static_assert(33 * 33 < (1 << 11), "direction bins fit 11 bits");
  size_t t_sort = 0, t_scan = 0;
  MUTO_CHECK_HIP(ctx, sort_pairs_u64_u32_ws(nullptr, nullptr, nullptr, nullptr, n, 0, key_bits,
                                           nullptr, &t_sort, nullptr, s));
  MUTO_CHECK_HIP(ctx, exclusive_scan_u32_to_u64_ws(nullptr, nullptr, n + 1, nullptr, &t_scan, s));
  uint64_t *key, *key2, *roff, *kcount;
  uint32_t *ord, *ord2, *bound, *live;
  double* tab;
  char *tsort, *tscan;
  auto carve1 = [&](Arena& A) {
    key = A.take<uint64_t>(n);
    key2 = A.take<uint64_t>(n);
    ord = A.take<uint32_t>(n);
    ord2 = A.take<uint32_t>(n);
    tab = A.take<double>(9 * n);
    bound = A.take<uint32_t>(n + 1);
    roff = A.take<uint64_t>(n + 1);
    live = A.take<uint32_t>(n);
    kcount = A.take<uint64_t>(KIND == 1 ? n : 1);
    tsort = A.take<char>(t_sort);
    tscan = A.take<char>(t_scan);
  };
  Arena A1;
  carve1(A1);
  void* w1 = nullptr;
  MUTO_CHECK_HIP(ctx, ctx->workspace(0, A1.off, &w1));
  A1 = Arena{static_cast<char*>(w1), 0};
  carve1(A1);
  // 1. the rays' spatial order
  if (KIND == 0)
    k_ray_keys<0><<<grid_for(n, 256), 256, 0, s>>>(c, dr, n, key, ord);
  else
    k_ray_keys<1><<<grid_for(n, 256), 256, 0, s>>>(c, dr, n, key, ord);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  bool alt = false;
  MUTO_CHECK_HIP(ctx, sort_pairs_u64_u32_ws(key, ord, key2, ord2, n, 0, key_bits, tsort, &t_sort,
                                           &alt, s));
  const uint32_t* order = alt ? ord2 : ord;
  // per ray in that order
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(bound + n, 0, sizeof(uint32_t), s));
  if (KIND == 0) {
    if (cut)
      k_poca_prep<true><<<grid_for(n, 256), 256, 0, s>>>(c, dr, n, order, di, dc, cut_vox, cut_val,
                                                          tab, live, bound);
    else
      k_poca_prep<false><<<grid_for(n, 256), 256, 0, s>>>(c, dr, n, order, di, dc, nullptr, nullptr,
                                                           tab, live, bound);
  } else {
    k_slr_prep<<<grid_for(n, 256), 256, 0, s>>>(c, dr, n, order, tab, bound);
  }
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  MUTO_CHECK_HIP(ctx, exclusive_scan_u32_to_u64_ws(bound, roff, n + 1, tscan, &t_scan, s));
  uint64_t nrec = 0;
  PocaCounters hc{};
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(&nrec, roff + n, sizeof nrec, hipMemcpyDeviceToHost, s));
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(&hc, dc, sizeof hc, hipMemcpyDeviceToHost, s));
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  if (nrec >= 0x7FFFFFFFull) { *fallback = true; return MUTO_OK; }
  const size_t n_walk = KIND == 0 ? (size_t)hc.in_grid : n;  // accepted rays come first
  // per record (and per tile)
  const size_t n_task_max = ntl + nrec / TILE_TASK + 1;
  size_t t_sort2 = 0, t_scan2 = 0;
  MUTO_CHECK_HIP(ctx, sort_pairs_u32_u32_ws(nullptr, nullptr, nullptr, nullptr, nrec, 0,
                                           bits_of(T.n_tiles), nullptr, &t_sort2, nullptr, s));
  MUTO_CHECK_HIP(ctx, exclusive_scan_u32_to_u64_ws(nullptr, nullptr, ntl + 1, nullptr, &t_scan2, s));
  TileRec* rec;
  uint32_t *rkey, *rval, *rkey2, *rval2, *ts, *te, *ntk;
  uint64_t* toff;
  TileTask* tasks;
  unsigned int *ctr;  // [0] overflow, [1] task counter
  char *tsort2, *tscan2;
  auto carve2 = [&](Arena& A) {
    rec = A.take<TileRec>(nrec);
    rkey = A.take<uint32_t>(nrec);
    rval = A.take<uint32_t>(nrec);
    rkey2 = A.take<uint32_t>(nrec);
    rval2 = A.take<uint32_t>(nrec);
    ts = A.take<uint32_t>(ntl);
    te = A.take<uint32_t>(ntl);
    ntk = A.take<uint32_t>(ntl + 1);
    toff = A.take<uint64_t>(ntl + 1);
    tasks = A.take<TileTask>(n_task_max);
    ctr = A.take<unsigned int>(2);
    tsort2 = A.take<char>(t_sort2);
    tscan2 = A.take<char>(t_scan2);
  };
  Arena A2;
  carve2(A2);
  void* w2 = nullptr;
  MUTO_CHECK_HIP(ctx, ctx->workspace(1, A2.off, &w2));
  A2 = Arena{static_cast<char*>(w2), 0};
  carve2(A2);
  // 2. the walks, records
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(ctr, 0, 2 * sizeof(unsigned int), s));
  if (nrec) {
    k_fill_u32<<<grid_for(nrec, 256), 256, 0, s>>>(nrec, T.n_tiles, rkey);  // unused slots sort last
    k_iota_u32<<<grid_for(nrec, 256), 256, 0, s>>>(nrec, rval);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
  }
  if (n_walk) {
    k_tile_split<KIND><<<grid_for(n_walk, 256), 256, 0, s>>>(g, T, c.straight_line, n_walk, tab, roff,
                                                             rec, rkey, kcount, dc, ctr);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
  }
  // 3. records by tile, tile ranges, tasks
  MUTO_CHECK_HIP(ctx, sort_pairs_u32_u32_ws(rkey, rval, rkey2, rval2, nrec, 0, bits_of(T.n_tiles),
                                           tsort2, &t_sort2, &alt, s));
  const uint32_t* skey = alt ? rkey2 : rkey;
  const uint32_t* sval = alt ? rval2 : rval;
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(ts, 0, ntl * sizeof(uint32_t), s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(te, 0, ntl * sizeof(uint32_t), s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(ntk + ntl, 0, sizeof(uint32_t), s));
  if (nrec) {
    k_tile_bounds<<<grid_for(nrec, 256), 256, 0, s>>>(nrec, skey, T.n_tiles, ts, te);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
  }
  k_tile_ntask<<<grid_for(ntl, 256), 256, 0, s>>>(T.n_tiles, ts, te, ntk);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  MUTO_CHECK_HIP(ctx, exclusive_scan_u32_to_u64_ws(ntk, toff, ntl + 1, tscan2, &t_scan2, s));
  k_tile_tasks<<<grid_for(ntl, 256), 256, 0, s>>>(T.n_tiles, ts, te, toff, tasks);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  // 4. the tiles' sums
  int occ = 1;
  MUTO_CHECK_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_tile_accum<KIND>, TILE_BS, 0));
  const unsigned grid = (unsigned)(std::max(1, occ) * ctx->n_cu);
  k_tile_accum<KIND><<<grid, TILE_BS, 0, s>>>(g, T, c.straight_line, c.length_weighted, c.has_energy,
                                               tasks, toff + ntl, ctr + 1, sval, rec, tab, live,
                                               kcount, dw, di);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  unsigned int hov = 0;
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(&hov, ctr, sizeof hov, hipMemcpyDeviceToHost, s));
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  if (hov) *fallback = true;  // a bound was exceeded: the images are redone
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
  bool fallback = false;
  if (int rc = tiled_images<0>(ctx, c, dr.p, ntot, di.p, dw.p, dc.p, false, nullptr, nullptr, &fallback))
    return rc;
  if (fallback) {
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(di.p, 0, nv * sizeof(double), s));
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(dw.p, 0, nv * sizeof(double), s));
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(dc.p, 0, sizeof(PocaCounters), s));
    k_poca_image<false><<<grid_for(ntot, 256), 256, 0, s>>>(c, dr.p, ntot, di.p, dw.p, dc.p,
                                                           nullptr, nullptr);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
  }
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
  bool fallback = false;
  if (int rc = tiled_images<0>(ctx, c, dr.p, ntot, nullptr, dw.p, dc.p, true, vox.p, val.p, &fallback))
    return rc;
  if (fallback) {
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(dw.p, 0, nv * sizeof(double), s));
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(dc.p, 0, sizeof(PocaCounters), s));
    k_poca_image<true><<<grid_for(ntot, 256), 256, 0, s>>>(c, dr.p, ntot, nullptr, dw.p, dc.p,
                                                          vox.p, val.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
  }
  k_cut_count<<<grid_for(ntot, 256), 256, 0, s>>>(ntot, vox.p, cnt.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  MUTO_CHECK_HIP(ctx, exclusive_scan_u32_to_u64(cnt.p, off.p, nv + 1, s));
  // the longest list and the voxels whose lists one lane does not sort,
  // listed on the device
  DBuf<uint32_t> lctr, lbig, lhuge;
  MUTO_CHECK_HIP(ctx, lctr.alloc(3));
  MUTO_CHECK_HIP(ctx, lbig.alloc(nv));
  MUTO_CHECK_HIP(ctx, lhuge.alloc(nv));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(lctr.p, 0, 3 * sizeof(uint32_t), s));
  k_cut_lists<<<grid_for(nv, 256), 256, 0, s>>>(nv, cnt.p, lctr.p, lbig.p, lhuge.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  uint64_t nrec = 0;
  uint32_t hl[3] = {0, 0, 0};  // longest list, big lists, huge lists
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(&nrec, off.p + nv, sizeof nrec, hipMemcpyDeviceToHost, s));
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(hl, lctr.p, sizeof hl, hipMemcpyDeviceToHost, s));
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  const uint32_t mmax = hl[0];  // 0 when no list exceeds CUT_LDS
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
  // lists longer than one lane sorts: one block each; those above the LDS
  // capacity one launch at a time, sharing the scratch
  if (hl[1]) {
    k_cut_big<<<hl[1], 256, 0, s>>>(lbig.p, cnt.p, off.p, slot.p, dw.p, cut_percentile, cfg->eps,
                                    nullptr, di.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
  }
  for (uint32_t h = 0; h < hl[2]; ++h) {
    k_cut_big<<<1, 256, 0, s>>>(lhuge.p + h, cnt.p, off.p, slot.p, dw.p, cut_percentile, cfg->eps,
                                scratch.p, di.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
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
  bool fallback = false;
  if (int rc = tiled_images<1>(ctx, c, dr.p, ntot, di.p, dw.p, dc.p, false, nullptr, nullptr, &fallback))
    return rc;
  if (fallback) {
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(di.p, 0, nv * sizeof(double), s));
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(dw.p, 0, nv * sizeof(double), s));
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(dc.p, 0, sizeof(PocaCounters), s));
    k_slr_image<<<grid_for(ntot, 256), 256, 0, s>>>(c, dr.p, ntot, di.p, dw.p, dc.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
  }
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
