// Device helpers shared by the FP64 EM/MAP kernels (em.hip) and the MAP
// objective operator (map_op.hip).
#pragma once
#include <math.h>

#include "em_plan.hpp"
#include "geom_dev.hpp"

namespace muto {

// ------------------------------------------------------------------ //
// shared device helpers                                              //
// ------------------------------------------------------------------ //
struct Mat2 {
  double m00, m01, m10, m11;
};

// Eigen computeInverseAndDetWithCheck for 2x2: det = m00*m11 - m10*m01,
// invertible iff |det| > 1e-12, inverse = adj * (1/det), untouched (zero)
// otherwise.
__device__ __forceinline__ bool inv_det(const Mat2& m, Mat2& inv,
                                        double& det) {
  det = m.m00 * m.m11 - m.m10 * m.m01;
  inv.m00 = inv.m01 = inv.m10 = inv.m11 = 0.0;
  if (!(fabs(det) > 1e-12)) return false;
  const double invdet = 1.0 / det;
  inv.m00 = m.m11 * invdet;
  inv.m10 = -m.m10 * invdet;
  inv.m01 = -m.m01 * invdet;
  inv.m11 = m.m00 * invdet;
  return true;
}

__device__ __forceinline__ void weight(double L, double T, Mat2& W) {
  const double LT = L * T;  // generateWeightMatrix, MutoMLSD.cpp:417-430
  W.m00 = L;
  W.m01 = W.m10 = LT + L * L * 0.5;
  W.m11 = L * L * L / 3.0 + (L + T) * LT;
}

struct ErrM {
  double e00, e01, e11;
};

ErrM error_matrix(const muto_mlsd_cfg& c);  // MutoMLSD.cpp:168-170

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block-level sum of one double per thread into *dst (one atomic per wave).
__device__ __forceinline__ void atomic_block_sum(double* dst, double v) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0 && v != 0.0) atomicAdd(dst, v);
}

// Deterministic block sum: the block's total of v (waves combined in a
// fixed order) goes to part[blockIdx.x]; every thread of the block must
// call it.  k_sum_parts then adds the parts in a fixed order, so the total
// is the same in every run.
__device__ __forceinline__ void block_sum_to(double* part, double v) {
  __shared__ double ws[16];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (unsigned k = 0; k < (blockDim.x + 63) / 64; ++k) t += ws[k];
    part[blockIdx.x] = t;
  }
}

namespace {
// out[0] = sum of part[0..n) in a fixed order (launch with one block of 256)
__global__ void __launch_bounds__(256)
k_sum_parts_f64(const double* __restrict__ part, size_t n, double* __restrict__ out) {
  double v = 0.0;
  for (size_t k = threadIdx.x; k < n; k += blockDim.x) v += part[k];
  block_sum_to(out, v);
}
}  // namespace

__device__ __forceinline__ void wave_count(unsigned long long* dst, bool v) {
  unsigned long long m = __ballot(v);
  if (m && (threadIdx.x & 63) == (unsigned)(__ffsll((long long)m) - 1))
    atomicAdd(dst, (unsigned long long)__popcll(m));
}

// device counter slots
enum { C_SINGULAR = 0, C_CUBIC = 1, C_NEWTON = 2, C_OUT = 3, C_EMPTY = 4, C_MISS = 5 };

struct MapCfgD {
  GridD g;
  double beta, p, delta, eps;
  int obj_prior, all_nb;
};

#define MAP_SCALE 1.0e7

// The clamp-to-edge neighbours of voxel j (index2xyz, MutoMAP.cpp:283-290;
// getSixNeighbours / getAllNeighbours, :292-350, in their order) with their
// weights (1, 1/sqrt2, 1/sqrt3, :210-212): one call f(k, index, weight)
// each, NN (6 or 26) fixed at compile time, so that the offsets fold into
// the unrolled loop and nothing is kept in per-lane arrays (which the
// compiler would place in scratch memory).
template <int NN, class F>
__device__ __forceinline__ void for_each_neighbour(const MapCfgD& c, uint32_t j, F&& f) {
  const uint32_t nx = c.g.n[0], ny = c.g.n[1], nz = c.g.n[2];
  const uint32_t x = j % (nx * ny) % nx;
  const uint32_t y = j % (nx * ny) / nx;
  const uint32_t z = j / (nx * ny);
  const uint32_t X[3] = {0 < x ? x - 1 : 0, x, x < nx - 1 ? x + 1 : nx - 1};
  const uint32_t Y[3] = {0 < y ? y - 1 : 0, y, y < ny - 1 ? y + 1 : ny - 1};
  const uint32_t Z[3] = {0 < z ? z - 1 : 0, z, z < nz - 1 ? z + 1 : nz - 1};
  // the offsets (x, y, z) as indices 0 (-1), 1 (0), 2 (+1)
  constexpr int8_t O[26][3] = {
      {2, 1, 1}, {0, 1, 1}, {1, 2, 1}, {1, 0, 1}, {1, 1, 2}, {1, 1, 0}, {2, 2, 1},
      {0, 2, 1}, {2, 0, 1}, {0, 0, 1}, {1, 2, 2}, {1, 0, 2}, {1, 2, 0}, {1, 0, 0},
      {2, 1, 2}, {0, 1, 2}, {2, 1, 0}, {0, 1, 0}, {2, 2, 2}, {0, 2, 2}, {2, 0, 2},
      {0, 0, 2}, {2, 2, 0}, {0, 2, 0}, {2, 0, 0}, {0, 0, 0}};
  const double w2 = 1.0 / sqrt(2.0), w3 = 1.0 / sqrt(3.0);
#pragma unroll
  for (int k = 0; k < NN; ++k) {
    const size_t idx = (size_t)X[O[k][0]] + (size_t)nx * ((size_t)Y[O[k][1]] + (size_t)ny * Z[O[k][2]]);
    f(k, idx, k < 6 ? 1.0 : (k < 18 ? w2 : w3));
  }
}


MapCfgD map_cfg(const muto_em_plan* pl);

// Walks a selected track's merged path again (getMuonPathAndTail,
// MutoMLSD.cpp:382-415): Siddon(pin, P) then Siddon(P, pout), the out-leg's
// first length added to the in-leg's last entry (and dropped when the
// in-leg is empty); straight: Siddon(pin, pout).  f(j, L, s, oob) is called
// per entry in path order with s the path length before the entry and oob
// set when the voxel coordinates lie outside the grid; returns the total
// length.  geo = pin, pout, P (9 doubles).  Same voxel indices (xyz2index
// in u32 arithmetic) and lengths as the stored lists.
template <class F>
__device__ __forceinline__ double walk_path(const GridD& g, const double* geo, int straight, F&& f) {
  const double* pin = geo;
  const double* pout = geo + 3;
  const double* P = geo + 6;
  const uint32_t nx = g.n[0], nxy = g.n[0] * g.n[1];
  auto oob = [&](uint32_t x, uint32_t y, uint32_t z) {
    return x >= g.n[0] || y >= g.n[1] || z >= g.n[2];
  };
  double s = 0.0;
  if (straight) {
    siddon_walk(g, pin, pout, [&](uint32_t x, uint32_t y, uint32_t z, double L) {
      f(x + y * nx + z * nxy, L, s, oob(x, y, z));
      s += L;
    });
    return s;
  }
  bool have = false, pout_of = false;
  uint32_t pj = 0;
  double pL = 0.0;
  siddon_walk(g, pin, P, [&](uint32_t x, uint32_t y, uint32_t z, double L) {
    if (have) {
      f(pj, pL, s, pout_of);
      s += pL;
    }
    pj = x + y * nx + z * nxy;
    pL = L;
    pout_of = oob(x, y, z);
    have = true;
  });
  bool first = true;
  siddon_walk(g, P, pout, [&](uint32_t x, uint32_t y, uint32_t z, double L) {
    if (first) {
      first = false;
      if (have) {
        pL += L;
        f(pj, pL, s, pout_of);
        s += pL;
        have = false;
      }
      return;
    }
    f(x + y * nx + z * nxy, L, s, oob(x, y, z));
    s += L;
  });
  if (have) {
    f(pj, pL, s, pout_of);
    s += pL;
  }
  return s;
}

}  // namespace muto
