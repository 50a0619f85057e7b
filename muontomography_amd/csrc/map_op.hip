// MAP objective as a differentiable operator: value and gradient with
// respect to the voxel scattering densities, behind a forward/backward pair
// with a saved context.
//
//   Phi(lambda) = 1/2 sum_i cost_i(lambda) + beta R(lambda)
//
// cost_i is the EM iteration's per-track cost (MutoMLSD.cpp:456-468: log det
// Sigma_i + 1/2 <D_i, Sigma_i^-1>, 0 for a singular or non-positive Sigma_i)
// and R the MAP prior (MutoMAP.cpp:207-237): over every voxel j and its 6 or
// 26 clamp-to-edge neighbours k, w_jk psi(l_j - l_k), plus psi(l_j) with the
// object prior, psi(t) = (t^2 + delta^2)^(p/2), l = 1e7 lambda.  The
// reference never evaluates Phi or its gradient; its MAP step minimises a
// surrogate that is tangent to Phi in mean mode.
//
// Gradient: d cost_i / d lambda_j = -<W_ij, G_i> with
// G_i = p_i^2 (1/2 Sigma_i^-1 D_i Sigma_i^-1 - Sigma_i^-1), the transpose of
// the Siddon-weighted forward map through each track's covariance.  The
// forward pass stores G_i per track; the backward pass scatters
// -1/2 <W_ij, G_i> over each track's path (FP64 atomics) and adds the prior
// gradient.  FP64 reads the stored track-major lists or, when the plan keeps
// none (or the caller asks), walks every path again by Siddon; FP32C runs
// the brick-stream passes of the EM iteration.
#include <math.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "em_dev.hpp"
#include "em_plan.hpp"

namespace muto {
int fp32c_build(muto_em_plan* pl);
int fp32c_map_forward(muto_em_plan* pl, const double* lam, float* g_out,
                      unsigned long long* cmax_out, double* cost_sum);
int fp32c_map_backward(muto_em_plan* pl, const float* g, const unsigned long long* cmax,
                       double* acc);
int fp64s_map_forward(muto_em_plan* pl, const double* lam, double* g_out, double* cost_sum);
int fp64s_map_backward(muto_em_plan* pl, const double* g, double* acc);
}  // namespace muto

using namespace muto;

struct muto_map_saved {
  muto_em_plan* pl = nullptr;
  int precision = MUTO_FP64;
  int recompute = 0;
  double phi = 0.0;
  DBuf<double> lam;                // [nv] the lambda of the forward call
  DBuf<double> g64;                // FP64: [4 n_trk] G00, G01, G11, path length
  DBuf<float> g32;                 // FP32C: [4 n_trk] the track kernel's G'
  DBuf<unsigned long long> cmax;   // FP32C: the |c_ij| bound of that pass
  DBuf<double> grad;               // [nv] device gradient
  DBuf<double> sum;                // [2] cost sum, prior sum
  DBuf<double> part;               // per-block partial sums
  bool dev = false;                // MUTO_MAP_DEVICE_BUFFERS: caller buffers in device memory
  DBuf<float> acc32;               // FP32C with the paths walked again: [nv] float sums
  bool stream64 = false;           // FP64 on the FP64 brick stream (g64: p^2 G in rank order)
};

namespace {

// What the per-track passes read: the stored track-major lists (RC = 0) or
// the track geometry whose path is walked again (RC = 1).
struct MapTrackArgs {
  size_t n_trk, nv;
  const uint64_t* row_ptr;
  const uint32_t* vox;
  const double* len;
  const double* tail;
  GridD g;
  const double* geo;
  int straight;
  const double* trk4;
  const double* pr;
  ErrM E;
};

MapTrackArgs track_args(const muto_em_plan* pl) {
  MapTrackArgs a;
  a.n_trk = pl->n_trk;
  a.nv = pl->nv;
  a.row_ptr = pl->row_ptr.p;
  a.vox = pl->vox.p;
  a.len = pl->len.p;
  a.tail = pl->tail.p;
  a.g = pl->g;
  a.geo = pl->geo.p;
  a.straight = pl->cfg.straight_line;
  a.trk4 = pl->trk4.p;
  a.pr = pl->pr.p;
  a.E = error_matrix(pl->cfg);
  return a;
}

__device__ __forceinline__ void track_finish(const MapTrackArgs& a, size_t i, Mat2 cov, double& cost,
                                             bool& sing, double& G00, double& G01, double& G11);

// Track i: Sigma_i from lambda (stored lists in the reference's order, or
// the path walked again with moments about its start moved to the tail
// basis), Eigen's 2x2 det/inverse, cost_i, and G_i (zero where cost_i is
// 0); P = the path length (RC = 1).
template <int RC>
__device__ __forceinline__ void map_track(const MapTrackArgs& a, size_t i, const double* __restrict__ lam,
                                          double& cost, bool& sing, double& G00, double& G01,
                                          double& G11, double& P) {
  Mat2 cov = {0.0, 0.0, 0.0, 0.0};
  P = 0.0;
  if (RC == 0) {
    for (uint64_t k = a.row_ptr[i]; k < a.row_ptr[i + 1]; ++k) {
      Mat2 W;
      weight(a.len[k], a.tail[k], W);
      const uint32_t j = a.vox[k];
      const double lj = j < a.nv ? lam[j] : 0.0;
      cov.m00 = cov.m00 + lj * W.m00;
      cov.m01 = cov.m01 + lj * W.m01;
      cov.m10 = cov.m10 + lj * W.m10;
      cov.m11 = cov.m11 + lj * W.m11;
    }
  } else {
    // moments in tau = path length from the start: int (1, tau)(1, tau)^T
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    P = walk_path(a.g, a.geo + 9 * i, a.straight, [&](uint32_t j, double L, double s, bool) {
      const double lj = j < a.nv ? lam[j] : 0.0;
      a0 += lj * L;
      a1 += lj * (L * s + 0.5 * L * L);
      a2 += lj * (L * s * s + L * L * s + L * L * L / 3.0);
    });
    // tail t = P - tau: u(t) = A u(tau), A = [[1, 0], [P, -1]]
    cov.m00 = a0;
    cov.m01 = cov.m10 = P * a0 - a1;
    cov.m11 = P * (P * a0 - 2.0 * a1) + a2;
  }
  track_finish(a, i, cov, cost, sing, G00, G01, G11);
}

// Sigma_i = p_i^2 cov + E (cov: the path's weighted moments in the tail
// basis), Eigen's 2x2 det/inverse, cost_i, and G_i (zero where cost_i is 0)
__device__ __forceinline__ void track_finish(const MapTrackArgs& a, size_t i, Mat2 cov, double& cost,
                                             bool& sing, double& G00, double& G01, double& G11) {
  const double p = a.pr[i], pp = p * p;
  cov.m00 = pp * cov.m00 + a.E.e00;
  cov.m01 = pp * cov.m01 + a.E.e01;
  cov.m10 = pp * cov.m10 + a.E.e01;
  cov.m11 = pp * cov.m11 + a.E.e11;
  const double Sx = a.trk4[4 * i], Sy = a.trk4[4 * i + 1];
  const double Dx = a.trk4[4 * i + 2], Dy = a.trk4[4 * i + 3];
  const double d00 = Sx * Sx + Sy * Sy, d11 = Dx * Dx + Dy * Dy, d01 = Sx * Dx + Sy * Dy;
  Mat2 inv;
  double det;
  const bool ok = inv_det(cov, inv, det);
  G00 = G01 = G11 = 0.0;
  if (!ok || det <= 0) {
    sing = true;
  } else {
    cost = log(det) + (d00 * inv.m00 + d11 * inv.m11 + d01 * 2.0 * inv.m01) * 0.5;
    // G = p^2 (1/2 inv D inv - inv)
    const double A00 = inv.m00 * d00 + inv.m01 * d01, A01 = inv.m00 * d01 + inv.m01 * d11;
    const double A10 = inv.m10 * d00 + inv.m11 * d01, A11 = inv.m10 * d01 + inv.m11 * d11;
    G00 = pp * (0.5 * (A00 * inv.m00 + A01 * inv.m10) - inv.m00);
    G01 = pp * (0.5 * (A00 * inv.m01 + A01 * inv.m11) - inv.m01);
    G11 = pp * (0.5 * (A10 * inv.m01 + A11 * inv.m11) - inv.m11);
  }
}

// grad_j += scale <W_ij, G_i> over track i's path (FP64 atomics)
template <int RC>
__device__ __forceinline__ void map_scatter(const MapTrackArgs& a, size_t i, double G00, double G01,
                                            double G11, double P, double scale,
                                            double* __restrict__ grad) {
  if (G00 == 0.0 && G01 == 0.0 && G11 == 0.0) return;
  auto add = [&](uint32_t j, double L, double T) {
    Mat2 W;
    weight(L, T, W);
    atomicAdd(&grad[j], scale * (W.m00 * G00 + 2.0 * W.m01 * G01 + W.m11 * G11));
  };
  if (RC == 0) {
    for (uint64_t k = a.row_ptr[i]; k < a.row_ptr[i + 1]; ++k) {
      const uint32_t j = a.vox[k];
      if (j < a.nv) add(j, a.len[k], a.tail[k]);
    }
  } else {
    walk_path(a.g, a.geo + 9 * i, a.straight, [&](uint32_t j, double L, double s, bool) {
      if (j < a.nv) add(j, L, P - s - L);
    });
  }
}

// Forward: cost_i per block, G_i (and the path length) per track.
template <int RC>
__global__ void __launch_bounds__(256)
k_map_fwd64(MapTrackArgs a, const double* __restrict__ lam, double* __restrict__ gsave,
            double* __restrict__ cost_part, unsigned long long* __restrict__ cnt) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double cost = 0.0;
  bool sing = false;
  if (i < a.n_trk) {
    double G00, G01, G11, P;
    map_track<RC>(a, i, lam, cost, sing, G00, G01, G11, P);
    gsave[4 * i] = G00;
    gsave[4 * i + 1] = G01;
    gsave[4 * i + 2] = G11;
    gsave[4 * i + 3] = P;
  }
  block_sum_to(cost_part, cost);
  wave_count(&cnt[C_SINGULAR], sing);
}

// Backward: grad_j += scale <W_ij, G_i> (scale = -1/2 grad_phi).
template <int RC>
__global__ void __launch_bounds__(256)
k_map_bwd64(MapTrackArgs a, const double* __restrict__ gsave, double scale,
            double* __restrict__ grad) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_trk) return;
  map_scatter<RC>(a, i, gsave[4 * i], gsave[4 * i + 1], gsave[4 * i + 2], gsave[4 * i + 3], scale,
                  grad);
}

// FP32C backward with the paths walked again: the FP64 walk (the voxel
// indices and lengths of the stored lists), each entry's term
// scale <W_ij, G_i> rounded to FP32 and added with FP32 atomics.
__global__ void __launch_bounds__(256)
k_map_bwd32rc(MapTrackArgs a, const double* __restrict__ gsave, double scale,
              float* __restrict__ acc) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_trk) return;
  const double G00 = gsave[4 * i], G01 = gsave[4 * i + 1], G11 = gsave[4 * i + 2];
  if (G00 == 0.0 && G01 == 0.0 && G11 == 0.0) return;
  const double P = gsave[4 * i + 3];
  const float g00 = (float)(scale * G00), g01 = (float)(2.0 * scale * G01), g11 = (float)(scale * G11);
  walk_path(a.g, a.geo + 9 * i, a.straight, [&](uint32_t j, double L, double s, bool) {
    if (j >= a.nv) return;
    // W from the FP64 length and tail, its products with G in FP32
    const double T = P - s - L;
    const double LT = L * T;
    const float w00 = (float)L, w01 = (float)(LT + L * L * 0.5);
    const float w11 = (float)(L * L * L / 3.0 + (L + T) * LT);
    atomicAdd(&acc[j], fmaf(w11, g11, fmaf(w01, g01, w00 * g00)));
  });
}

// ---------------- the recomputing backward, per tile -------------------
// One global atomic per entry runs at about 2.4e10 entries/s on this GPU
// (FP64 or FP32, scattered voxels), so the backward that walks the paths
// again makes its sums per tile of MT_E^3 voxels in LDS instead:
//   1. once per plan, every track's merged path is walked and each run of
//      its entries inside one tile gets a record (MapRec: the Siddon walk's
//      state before the step of the run's first entry, and the path length
//      before that entry); the records are sorted by tile (and inside a
//      tile by leg and run length, see mt_key), stably, and cut into tasks
//      of at most MT_TASK records;
//   2. per backward call, one workgroup per task resumes each record's walk,
//      adds the terms of the run's entries into the tile in LDS, and adds
//      the tile into the gradient (stores it when one task covers the tile).
// The entries (voxel, length, and path length s before them, so the tail
// T = P - s - L) are walk_path's own, bit for bit; only the order of the
// sums differs from one atomic per entry.  A record takes 56 bytes per run
// of entries (about a dozen at 16^3 tiles), against 20 bytes per entry for
// the stored lists.
constexpr uint32_t MT_E = 16;
constexpr uint32_t MT_V = MT_E * MT_E * MT_E;
constexpr uint32_t MT_BS = 256;
constexpr uint32_t MT_TASK = 1024;

struct MtGeo {
  uint32_t nt[3];  // tiles per axis
};

// Tile of entry voxel j (j < nv) and the voxel's index in it, from the
// walk's coordinates or, where those lie outside the grid, from j (which
// then names another voxel, as the reference's xyz2index does)
__device__ __forceinline__ uint32_t mt_tile(const GridD& g, const MtGeo& T, uint32_t x, uint32_t y,
                                            uint32_t z, uint32_t j, uint32_t& lv) {
  if (!(x < g.n[0] && y < g.n[1] && z < g.n[2])) {
    const uint32_t r = j / g.n[0];
    x = j - r * g.n[0];
    y = r % g.n[1];
    z = r / g.n[1];
  }
  lv = (x % MT_E) + MT_E * ((y % MT_E) + MT_E * (z % MT_E));
  return x / MT_E + T.nt[0] * (y / MT_E + T.nt[1] * (z / MT_E));
}

// Every entry of a track's merged path in walk_path's order and arithmetic,
// f(j, s, leg, x, y, z, ax_pre, alpha_pre) with s the path length before
// the entry and the state of the step that recorded the entry's voxel on
// leg `leg` (0: the in-leg, or the whole straight path; 1: the out-leg).
// The in-leg's last entry, whose length takes the out-leg's first length on
// top, carries its own step's state.  Returns the path length (walk_path's).
template <class F>
__device__ __forceinline__ double walk_entries(const GridD& g, const double* geo, int straight, F&& f) {
  const double* pin = geo;
  const double* pout = geo + 3;
  const double* P = geo + 6;
  const uint32_t nx = g.n[0], nxy = g.n[0] * g.n[1];
  double s = 0.0;
  if (straight) {
    SiddonCursor c;
    if (!c.begin(g, pin, pout)) return s;
    c.run(0xFFFFFFFFu, [&](uint32_t x, uint32_t y, uint32_t z, double L) {
      double ap[3];
      c.snapshot(x, y, z, ap);
      f(x + y * nx + z * nxy, s, 0u, x, y, z, ap, c.alpha_last);
      s += L;
    });
    return s;
  }
  bool have = false;
  uint32_t px = 0, py = 0, pz = 0;
  double pL = 0.0, pal = 0.0, pap[3] = {0.0, 0.0, 0.0};
  {
    SiddonCursor c;
    if (c.begin(g, pin, P))
      c.run(0xFFFFFFFFu, [&](uint32_t x, uint32_t y, uint32_t z, double L) {
        if (have) {
          f(px + py * nx + pz * nxy, s, 0u, px, py, pz, pap, pal);
          s += pL;
        }
        px = x;
        py = y;
        pz = z;
        pL = L;
        c.snapshot(x, y, z, pap);
        pal = c.alpha_last;
        have = true;
      });
  }
  SiddonCursor c2;
  bool first = true;
  if (c2.begin(g, P, pout))
    c2.run(0xFFFFFFFFu, [&](uint32_t x, uint32_t y, uint32_t z, double L) {
      if (first) {
        first = false;
        if (have) {
          pL += L;
          f(px + py * nx + pz * nxy, s, 0u, px, py, pz, pap, pal);
          s += pL;
          have = false;
        }
        return;
      }
      double ap[3];
      c2.snapshot(x, y, z, ap);
      f(x + y * nx + z * nxy, s, 1u, x, y, z, ap, c2.alpha_last);
      s += L;
    });
  if (have) {
    f(px + py * nx + pz * nxy, s, 0u, px, py, pz, pap, pal);
    s += pL;
  }
  return s;
}

// The entries of track geo's merged path from record R's entry on, in path
// order, f(j, x, y, z, L, s) for each until it returns false or the path
// ends; the same voxels, lengths and s as walk_path.
template <class F>
__device__ __forceinline__ void resume_entries(const GridD& g, const double* geo, int straight,
                                               const MapRec& R, F&& f) {
  const double* pin = geo;
  const double* pout = geo + 3;
  const double* P = geo + 6;
  const uint32_t nx = g.n[0], nxy = g.n[0] * g.n[1];
  double s = R.s;
  bool go = true;
  if (straight || (R.tl & 1u)) {
    SiddonCursor c;
    c.begin(g, straight ? pin : P, pout);  // succeeded when the record was taken
    c.resume(R.x, R.y, R.z, R.ax, R.alpha_last);
    c.run_while([&](uint32_t x, uint32_t y, uint32_t z, double L) -> bool {
      if (!f(x + y * nx + z * nxy, x, y, z, L, s)) return false;
      s += L;
      return true;
    });
    return;
  }
  // the in-leg: each entry taken when the next step is found, the last one
  // with the out-leg's first length added
  bool have = false;
  uint32_t px = 0, py = 0, pz = 0;
  double pL = 0.0;
  {
    SiddonCursor c;
    c.begin(g, pin, P);
    c.resume(R.x, R.y, R.z, R.ax, R.alpha_last);
    c.run_while([&](uint32_t x, uint32_t y, uint32_t z, double L) -> bool {
      if (have) {
        if (!f(px + py * nx + pz * nxy, px, py, pz, pL, s)) {
          go = false;
          return false;
        }
        s += pL;
      }
      px = x;
      py = y;
      pz = z;
      pL = L;
      have = true;
      return true;
    });
  }
  if (!go) return;
  SiddonCursor c2;
  bool first = true;
  if (c2.begin(g, P, pout))
    c2.run_while([&](uint32_t x, uint32_t y, uint32_t z, double L) -> bool {
      if (first) {
        first = false;
        if (have) {
          pL += L;
          if (!f(px + py * nx + pz * nxy, px, py, pz, pL, s)) {
            go = false;
            return false;
          }
          s += pL;
          have = false;
        }
        return true;
      }
      if (!f(x + y * nx + z * nxy, x, y, z, L, s)) {
        go = false;
        return false;
      }
      s += L;
      return true;
    });
  if (go && have) f(px + py * nx + pz * nxy, px, py, pz, pL, s);
}

// Sort key of a record: its tile, then its leg and its run's length
// (longest first, 63 entries and more alike), so that the lanes of a
// wavefront, which take consecutive records, walk runs of one kind and of
// similar lengths.  Tiles below 2^25.
constexpr int MT_SUB_BITS = 7;
__device__ __forceinline__ uint32_t mt_key(uint32_t tile, uint32_t leg, uint32_t run) {
  return tile << MT_SUB_BITS | leg << 6 | (63u - min(run, 63u));
}

// Step 1 per track: the runs of its entries (voxel below nv) in one tile.
// COUNT: their number into cnt[i] and the path length into plen[i]; else
// the records at roff[i], each with its sort key and its own index as the
// value.
template <bool COUNT>
__global__ void __launch_bounds__(256)
k_mt_split(MapTrackArgs a, MtGeo T, const uint64_t* __restrict__ roff, uint32_t* __restrict__ cnt,
           double* __restrict__ plen, MapRec* __restrict__ rec, uint32_t* __restrict__ key,
           uint32_t* __restrict__ idx) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_trk) return;
  const uint64_t b = COUNT ? 0 : roff[i];
  uint32_t cur = 0xFFFFFFFFu, k = 0, run = 0, run_leg = 0;
  const double P = walk_entries(a.g, a.geo + 9 * i, a.straight,
               [&](uint32_t j, double s, uint32_t leg, uint32_t x, uint32_t y, uint32_t z,
                   const double* ap, double al) {
                 if (j >= a.nv) return;
                 uint32_t lv;
                 const uint32_t t = mt_tile(a.g, T, x, y, z, j, lv);
                 if (t == cur) {
                   ++run;
                   return;
                 }
                 if (!COUNT && k > 0) key[b + k - 1] = mt_key(cur, run_leg, run);
                 cur = t;
                 run = 1;
                 run_leg = leg;
                 if (!COUNT) {
                   MapRec R;
                   R.ax[0] = ap[0];
                   R.ax[1] = ap[1];
                   R.ax[2] = ap[2];
                   R.alpha_last = al;
                   R.s = s;
                   R.x = x;
                   R.y = y;
                   R.z = z;
                   R.tl = (uint32_t)(i << 1) | leg;
                   rec[b + k] = R;
                   idx[b + k] = (uint32_t)(b + k);
                 }
                 ++k;
               });
  if (COUNT) {
    cnt[i] = k;
    plen[i] = P;
  } else if (k > 0) {
    key[b + k - 1] = mt_key(cur, run_leg, run);
  }
}

__global__ void k_mt_gather(size_t n, const uint32_t* __restrict__ idx, const MapRec* __restrict__ src,
                            MapRec* __restrict__ dst) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) dst[q] = src[idx[q]];
}

// first and end record of every tile in the sorted keys
__global__ void k_mt_bounds(size_t n, const uint32_t* __restrict__ key, uint32_t* __restrict__ ts,
                            uint32_t* __restrict__ te) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const uint32_t t = key[q] >> MT_SUB_BITS;
  if (q == 0 || (key[q - 1] >> MT_SUB_BITS) != t) ts[t] = (uint32_t)q;
  if (q + 1 == n || (key[q + 1] >> MT_SUB_BITS) != t) te[t] = (uint32_t)(q + 1);
}

__global__ void k_mt_reset(unsigned int* __restrict__ work) {
  if (threadIdx.x == 0 && blockIdx.x == 0) work[0] = 0u;
}

// Step 2: grad_j += sum over the tile's runs of scale <W_ij, G_i>, with
// the terms of k_map_bwd64<1> (F32 false) or of k_map_bwd32rc (F32 true:
// W and its product with G in FP32), summed in FP64 in LDS; persistent
// workgroups over the tasks.
template <bool F32>
__global__ void __launch_bounds__(MT_BS)
k_mt_bwd(MapTrackArgs a, MtGeo T, const MapTask* __restrict__ tasks, uint32_t n_tasks,
         unsigned int* __restrict__ work, const MapRec* __restrict__ rec,
         const double* __restrict__ gsave, double scale, double* __restrict__ grad) {
  __shared__ double acc[MT_V];
  __shared__ uint32_t sh;
  const uint32_t t = threadIdx.x;
  for (;;) {
    if (t == 0) sh = atomicAdd(work, 1u);
    for (uint32_t v = t; v < MT_V; v += MT_BS) acc[v] = 0.0;
    __syncthreads();
    const uint32_t ti = sh;
    if (ti >= n_tasks) break;
    const MapTask tk = tasks[ti];
    for (uint32_t q = tk.q0 + t; q < tk.q1; q += MT_BS) {
      const MapRec R = rec[q];
      const size_t i = R.tl >> 1;
      const double G00 = gsave[4 * i], G01 = gsave[4 * i + 1], G11 = gsave[4 * i + 2];
      if (G00 == 0.0 && G01 == 0.0 && G11 == 0.0) continue;
      const double P = gsave[4 * i + 3];
      const float g00 = (float)(scale * G00), g01 = (float)(2.0 * scale * G01),
                  g11 = (float)(scale * G11);
      resume_entries(a.g, a.geo + 9 * i, a.straight, R,
                     [&](uint32_t j, uint32_t x, uint32_t y, uint32_t z, double L, double s) -> bool {
                       if (j >= a.nv) return true;
                       uint32_t lv;
                       if (mt_tile(a.g, T, x, y, z, j, lv) != tk.tile) return false;
                       const double Tl = P - s - L;
                       double term;
                       if (F32) {
                         const double LT = L * Tl;
                         const float w00 = (float)L, w01 = (float)(LT + L * L * 0.5);
                         const float w11 = (float)(L * L * L / 3.0 + (L + Tl) * LT);
                         term = (double)fmaf(w11, g11, fmaf(w01, g01, w00 * g00));
                       } else {
                         Mat2 W;
                         weight(L, Tl, W);
                         term = scale * (W.m00 * G00 + 2.0 * W.m01 * G01 + W.m11 * G11);
                       }
                       atomicAdd(&acc[lv], term);
                       return true;
                     });
    }
    __syncthreads();
    const uint32_t tyz = tk.tile / T.nt[0];
    const uint32_t o0 = (tk.tile - tyz * T.nt[0]) * MT_E, o1 = (tyz % T.nt[1]) * MT_E,
                   o2 = (tyz / T.nt[1]) * MT_E;
    for (uint32_t v = t; v < MT_V; v += MT_BS) {
      const double val = acc[v];
      if (val == 0.0) continue;  // (a nonzero sum lies inside the grid)
      const uint32_t x = o0 + v % MT_E, y = o1 + (v / MT_E) % MT_E, z = o2 + v / (MT_E * MT_E);
      const size_t j = (size_t)x + (size_t)a.g.n[0] * ((size_t)y + (size_t)a.g.n[1] * z);
      if (tk.ntask == 1)
        grad[j] = val;
      else
        atomicAdd(&grad[j], val);
    }
    __syncthreads();
  }
}

// Forward per tile: each run's moments about its path's start,
// sum lambda_j (L, L s + L^2/2, L s^2 + L^2 s + L^3/3) with the entry
// arithmetic of map_track<1> and lambda from the tile in LDS, into
// part[3 orig[q]] (the run's place in track order).
__global__ void __launch_bounds__(MT_BS)
k_mt_fwd(MapTrackArgs a, MtGeo T, const MapTask* __restrict__ tasks, uint32_t n_tasks,
         unsigned int* __restrict__ work, const MapRec* __restrict__ rec,
         const uint32_t* __restrict__ orig, const double* __restrict__ lam, double* __restrict__ part) {
  __shared__ double lam_s[MT_V];
  __shared__ uint32_t sh;
  const uint32_t t = threadIdx.x;
  for (;;) {
    if (t == 0) sh = atomicAdd(work, 1u);
    __syncthreads();
    const uint32_t ti = sh;
    if (ti >= n_tasks) break;
    const MapTask tk = tasks[ti];
    const uint32_t tyz = tk.tile / T.nt[0];
    const uint32_t o0 = (tk.tile - tyz * T.nt[0]) * MT_E, o1 = (tyz % T.nt[1]) * MT_E,
                   o2 = (tyz / T.nt[1]) * MT_E;
    for (uint32_t v = t; v < MT_V; v += MT_BS) {
      const uint32_t x = o0 + v % MT_E, y = o1 + (v / MT_E) % MT_E, z = o2 + v / (MT_E * MT_E);
      lam_s[v] = (x < a.g.n[0] && y < a.g.n[1] && z < a.g.n[2])
                     ? lam[(size_t)x + (size_t)a.g.n[0] * ((size_t)y + (size_t)a.g.n[1] * z)]
                     : 0.0;
    }
    __syncthreads();
    for (uint32_t q = tk.q0 + t; q < tk.q1; q += MT_BS) {
      const MapRec R = rec[q];
      const size_t i = R.tl >> 1;
      double a0 = 0.0, a1 = 0.0, a2 = 0.0;
      resume_entries(a.g, a.geo + 9 * i, a.straight, R,
                     [&](uint32_t j, uint32_t x, uint32_t y, uint32_t z, double L, double s) -> bool {
                       if (j >= a.nv) return true;
                       uint32_t lv;
                       if (mt_tile(a.g, T, x, y, z, j, lv) != tk.tile) return false;
                       const double lj = lam_s[lv];
                       a0 += lj * L;
                       a1 += lj * (L * s + 0.5 * L * L);
                       a2 += lj * (L * s * s + L * L * s + L * L * L / 3.0);
                       return true;
                     });
      double* o = part + 3 * (size_t)orig[q];
      o[0] = a0;
      o[1] = a1;
      o[2] = a2;
    }
    __syncthreads();  // the tile is loaded again for the next task
  }
}

// Per track: its runs' moments added in path order, moved to the tail
// basis with the path length P as map_track<1> does, then cost_i and G_i.
__global__ void __launch_bounds__(256)
k_mt_track(MapTrackArgs a, const uint64_t* __restrict__ roff, const double* __restrict__ part,
           const double* __restrict__ plen, double* __restrict__ gsave, double* __restrict__ cost_part,
           unsigned long long* __restrict__ cnt) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double cost = 0.0;
  bool sing = false;
  if (i < a.n_trk) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (uint64_t r = roff[i]; r < roff[i + 1]; ++r) {
      a0 += part[3 * r];
      a1 += part[3 * r + 1];
      a2 += part[3 * r + 2];
    }
    const double P = plen[i];
    Mat2 cov;
    cov.m00 = a0;
    cov.m01 = cov.m10 = P * a0 - a1;
    cov.m11 = P * (P * a0 - 2.0 * a1) + a2;
    double G00, G01, G11;
    track_finish(a, i, cov, cost, sing, G00, G01, G11);
    gsave[4 * i] = G00;
    gsave[4 * i + 1] = G01;
    gsave[4 * i + 2] = G11;
    gsave[4 * i + 3] = P;
  }
  block_sum_to(cost_part, cost);
  wave_count(&cnt[C_SINGULAR], sing);
}

__global__ void k_zero_f32(size_t n, float* __restrict__ x) {
  size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) x[j] = 0.0f;
}

__global__ void k_widen_f32(size_t n, const float* __restrict__ a, double* __restrict__ b) {
  size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) b[j] = (double)a[j];
}

// Forward and backward of one track in one pass, nothing stored per track
// (muto_map_value_grad_noalloc): the same cost_i, G_i and scatter as the
// two kernels above.
template <int RC>
__global__ void __launch_bounds__(256)
k_map_fb64(MapTrackArgs a, const double* __restrict__ lam, double scale, double* __restrict__ grad,
           double* __restrict__ cost_part, unsigned long long* __restrict__ cnt) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double cost = 0.0;
  bool sing = false;
  if (i < a.n_trk) {
    double G00, G01, G11, P;
    map_track<RC>(a, i, lam, cost, sing, G00, G01, G11, P);
    map_scatter<RC>(a, i, G00, G01, G11, P, scale, grad);
  }
  block_sum_to(cost_part, cost);
  wave_count(&cnt[C_SINGULAR], sing);
}

__device__ __forceinline__ double psi(double t, double p, double dd) {
  return pow(t * t + dd, p / 2.0);
}
__device__ __forceinline__ double dpsi(double t, double p, double dd) {
  return p * t * pow(t * t + dd, (p - 2.0) / 2.0);
}

// prior R(lambda), summed over all voxels (NN: 6 or 26 neighbours)
template <int NN>
__global__ void __launch_bounds__(256)
k_prior_phi(MapCfgD c, size_t nv, const double* __restrict__ lam, double* __restrict__ part) {
  size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (j < nv) {
    const double lj = lam[j] * MAP_SCALE, dd = c.delta * c.delta;
    for_each_neighbour<NN>(c, (uint32_t)j, [&](int, size_t idx, double w) {
      v += w * psi(lj - lam[idx] * MAP_SCALE, c.p, dd);
    });
    if (c.obj_prior) v += psi(lj, c.p, dd);
  }
  block_sum_to(part, v);
}

// out[0] = sum of part[0..n) in a fixed order (one block of 256)
__global__ void __launch_bounds__(256)
k_sum_parts(const double* __restrict__ part, size_t n, double* __restrict__ out) {
  double v = 0.0;
  for (size_t k = threadIdx.x; k < n; k += blockDim.x) v += part[k];
  block_sum_to(out, v);
}

// grad += scale dR/dlambda (scale = grad_phi beta): each pair term
// w psi(l_j - l_k) adds to j and subtracts from k (FP64 atomics)
template <int NN>
__global__ void __launch_bounds__(256)
k_prior_grad(MapCfgD c, size_t nv, const double* __restrict__ lam, double scale,
             double* __restrict__ grad) {
  size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nv) return;
  const double lj = lam[j] * MAP_SCALE, dd = c.delta * c.delta;
  const double s = scale * MAP_SCALE;
  double own = 0.0;
  for_each_neighbour<NN>(c, (uint32_t)j, [&](int, size_t idx, double w) {
    const double dv = w * dpsi(lj - lam[idx] * MAP_SCALE, c.p, dd);
    own += s * dv;
    atomicAdd(&grad[idx], -s * dv);
  });
  if (c.obj_prior) own += s * dpsi(lj, c.p, dd);
  atomicAdd(&grad[j], own);
}

// the prior kernels for the plan's neighbourhood
static void prior_phi(const MapCfgD& c, size_t nv, const double* lam, double* part, hipStream_t s) {
  if (c.all_nb)
    k_prior_phi<26><<<grid_for(nv, 256), 256, 0, s>>>(c, nv, lam, part);
  else
    k_prior_phi<6><<<grid_for(nv, 256), 256, 0, s>>>(c, nv, lam, part);
}
static void prior_grad(const MapCfgD& c, size_t nv, const double* lam, double scale, double* grad,
                       hipStream_t s) {
  if (c.all_nb)
    k_prior_grad<26><<<grid_for(nv, 256), 256, 0, s>>>(c, nv, lam, scale, grad);
  else
    k_prior_grad<6><<<grid_for(nv, 256), 256, 0, s>>>(c, nv, lam, scale, grad);
}

__global__ void k_scale(size_t n, double a, double* __restrict__ x) {
  size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) x[j] *= a;
}

// x[0..n) = 0 by a kernel (not a memset node: callers may capture the call)
__global__ void k_zero_f64(size_t n, double* __restrict__ x) {
  size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) x[j] = 0.0;
}

// phi = 1/2 cost sum + beta prior sum, on the device
__global__ void k_phi_combine(const double* __restrict__ sums, double beta, double* __restrict__ phi) {
  if (threadIdx.x == 0 && blockIdx.x == 0) phi[0] = 0.5 * sums[0] + beta * sums[1];
}

MtGeo mt_geo(const muto_em_plan* pl) {
  MtGeo T;
  for (int a = 0; a < 3; ++a) T.nt[a] = (pl->g.n[a] + MT_E - 1) / MT_E;
  return T;
}

// The tile records and tasks of the passes that walk the paths again, made
// at the first such MAP call on the plan, with each record's place in track
// order, each track's records there, its path length and the forward's
// moments per record (84 bytes per record, 16 per track).  Returns false
// when they are not kept: more than 2^31 records or tracks, more than
// 2^25 tiles, a plan that keeps no lists whose budget they exceed, or no
// device memory; the passes then walk each track in one lane, the backward
// with one atomic per entry.  Kept or not, the decision holds for the
// plan's life.
bool mt_build(muto_em_plan* pl) {
  if (pl->mt_state) return pl->mt_state > 0;
  pl->mt_state = -1;
  muto_ctx* ctx = pl->ctx;
  hipStream_t s = ctx->stream;
  const size_t nt = pl->n_trk;
  if (!nt || !pl->geo.p || nt >= (1ull << 31)) return false;
  const MtGeo T = mt_geo(pl);
  const uint64_t n_tiles = (uint64_t)T.nt[0] * T.nt[1] * T.nt[2];
  if (n_tiles >= (1ull << (32 - MT_SUB_BITS))) return false;
  const MapTrackArgs a = track_args(pl);
  auto fail = [&]() {
    (void)hipStreamSynchronize(s);
    (void)hipGetLastError();  // an allocation that failed leaves its status
    pl->mt_rec.release();
    pl->mt_task.release();
    pl->mt_work.release();
    pl->mt_orig.release();
    pl->mt_roff.release();
    pl->mt_plen.release();
    pl->mt_part.release();
    return false;
  };
  DBuf<uint32_t> cnt;
  if (cnt.alloc(nt + 1) != hipSuccess || pl->mt_roff.alloc(nt + 1) != hipSuccess ||
      pl->mt_plen.alloc(nt) != hipSuccess)
    return fail();
  if (hipMemsetAsync(cnt.p + nt, 0, sizeof(uint32_t), s) != hipSuccess) return fail();
  k_mt_split<true><<<grid_for(nt, 256), 256, 0, s>>>(a, T, nullptr, cnt.p, pl->mt_plen.p, nullptr, nullptr,
                                                     nullptr);
  if (hipGetLastError() != hipSuccess) return fail();
  if (exclusive_scan_u32_to_u64(cnt.p, pl->mt_roff.p, nt + 1, s) != hipSuccess) return fail();
  uint64_t n_rec = 0;
  if (copy_sync(ctx, &n_rec, pl->mt_roff.p + nt, sizeof n_rec, hipMemcpyDeviceToHost) != hipSuccess)
    return fail();
  cnt.release();
  if (n_rec == 0 || n_rec >= (1ull << 31)) return fail();
  constexpr double REC_BYTES = sizeof(MapRec) + sizeof(uint32_t) + 3 * sizeof(double);
  if (!pl->has_lists && (double)n_rec * REC_BYTES + 16.0 * (double)nt > (double)pl->list_budget)
    return fail();
  {
    DBuf<MapRec> rec0;
    DBuf<uint32_t> key;
    if (rec0.alloc(n_rec) != hipSuccess || key.alloc(n_rec) != hipSuccess ||
        pl->mt_orig.alloc(n_rec) != hipSuccess)
      return fail();
    k_mt_split<false><<<grid_for(nt, 256), 256, 0, s>>>(a, T, pl->mt_roff.p, nullptr, nullptr, rec0.p, key.p,
                                                        pl->mt_orig.p);
    if (hipGetLastError() != hipSuccess) return fail();
    int bits = 1;
    while ((1ull << bits) < n_tiles) ++bits;
    if (sort_pairs_u32_u32(key.p, pl->mt_orig.p, n_rec, 0, bits + MT_SUB_BITS, s) != hipSuccess) return fail();
    if (pl->mt_rec.alloc(n_rec) != hipSuccess) return fail();
    k_mt_gather<<<grid_for(n_rec, 256), 256, 0, s>>>(n_rec, pl->mt_orig.p, rec0.p, pl->mt_rec.p);
    if (hipGetLastError() != hipSuccess) return fail();
    rec0.release();
    if (pl->mt_part.alloc(3 * n_rec) != hipSuccess) return fail();
    DBuf<uint32_t> ts, te;
    if (ts.alloc(n_tiles) != hipSuccess || te.alloc(n_tiles) != hipSuccess) return fail();
    if (hipMemsetAsync(ts.p, 0, n_tiles * sizeof(uint32_t), s) != hipSuccess ||
        hipMemsetAsync(te.p, 0, n_tiles * sizeof(uint32_t), s) != hipSuccess)
      return fail();
    k_mt_bounds<<<grid_for(n_rec, 256), 256, 0, s>>>(n_rec, key.p, ts.p, te.p);
    if (hipGetLastError() != hipSuccess) return fail();
    std::vector<uint32_t> hs(n_tiles), he(n_tiles);
    if (copy_sync(ctx, hs.data(), ts.p, n_tiles * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess ||
        copy_sync(ctx, he.data(), te.p, n_tiles * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)
      return fail();
    // each tile's records cut into tasks of nearly equal size
    std::vector<MapTask> tasks;
    for (uint64_t t = 0; t < n_tiles; ++t) {
      const uint32_t m = he[t] - hs[t];
      if (m == 0) continue;
      const uint32_t k = (m + MT_TASK - 1) / MT_TASK, per = (m + k - 1) / k;
      for (uint32_t q = 0; q < k; ++q)
        tasks.push_back(MapTask{(uint32_t)t, hs[t] + q * per, std::min(hs[t] + (q + 1) * per, he[t]), k});
    }
    if (pl->mt_task.alloc(tasks.size()) != hipSuccess || pl->mt_work.alloc(1) != hipSuccess) return fail();
    if (copy_sync(ctx, pl->mt_task.p, tasks.data(), tasks.size() * sizeof(MapTask), hipMemcpyHostToDevice) !=
        hipSuccess)
      return fail();
    pl->mt_n_task = (uint32_t)tasks.size();
  }
  int ob = 0, of = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&ob, k_mt_bwd<false>, MT_BS, 0) != hipSuccess || ob < 1 ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&of, k_mt_fwd, MT_BS, 0) != hipSuccess || of < 1)
    return fail();
  pl->mt_grid = (int)std::min<uint64_t>((uint64_t)ob * ctx->n_cu, pl->mt_n_task);
  pl->mt_grid_fwd = (int)std::min<uint64_t>((uint64_t)of * ctx->n_cu, pl->mt_n_task);
  pl->mt_n_rec = n_rec;
  pl->mt_state = 1;
  return true;
}

}  // namespace

extern "C" int muto_map_forward(muto_em_plan* pl, const double* lambda, int precision, int flags,
                                double* phi, muto_map_saved** saved) {
  if (!pl) return MUTO_ERR_ARG;
  muto_ctx* ctx = pl->ctx;
  if (!lambda || !phi || !saved) { ctx->set_error("muto_map_forward: null argument"); return MUTO_ERR_ARG; }
  *saved = nullptr;
  if (precision != MUTO_FP64 && precision != MUTO_FP32C && precision != MUTO_FP64_STREAM) {
    ctx->set_error("muto_map_forward: precision must be MUTO_FP64, MUTO_FP64_STREAM or MUTO_FP32C");
    return MUTO_ERR_ARG;
  }
  if (flags & ~(MUTO_MAP_RECOMPUTE_SIDDON | MUTO_MAP_DEVICE_BUFFERS)) {
    ctx->set_error("muto_map_forward: unknown flag");
    return MUTO_ERR_ARG;
  }
  const bool recompute = (flags & MUTO_MAP_RECOMPUTE_SIDDON) || !pl->has_lists;
  const bool dev = (flags & MUTO_MAP_DEVICE_BUFFERS) != 0;
  if (recompute && !pl->geo.p && pl->n_trk) {
    ctx->set_error("muto_map_forward: the plan has no track geometry to walk (prepared plan)");
    return MUTO_ERR_STATE;
  }
  hipStream_t s = ctx->stream;
  auto* sv = new muto_map_saved();
  sv->pl = pl;
  sv->precision = precision;
  sv->recompute = recompute;
  sv->dev = dev;
  const size_t nv = pl->nv, nt = pl->n_trk;
  auto fail = [&](int rc) { delete sv; return rc; };
  const size_t n_part = std::max<size_t>(grid_for(nt, 256), grid_for(nv, 256));
  if (sv->lam.alloc(nv) != hipSuccess || sv->sum.alloc(2) != hipSuccess ||
      (!dev && sv->grad.alloc(nv) != hipSuccess) || sv->part.alloc(n_part) != hipSuccess) {
    ctx->set_error("muto_map_forward: out of device memory");
    return fail(MUTO_ERR_NOMEM);
  }
  // the saved context keeps its own copy of lambda (the caller may change
  // its buffer before the backward call)
  if (hipMemcpyAsync(sv->lam.p, lambda, nv * sizeof(double),
                     dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemsetAsync(sv->sum.p, 0, 2 * sizeof(double), s) != hipSuccess) {
    ctx->set_error("muto_map_forward: copy of lambda failed");
    return fail(MUTO_ERR_HIP);
  }
  double cost_sum = 0.0;
  if (nt) {
    // FP64, and FP32C with the paths walked again: the FP64 walk (index
    // exactness) with FP64 moment sums, G_i saved in FP64; FP32C's
    // recomputing backward then adds FP32 entry terms
    if (precision == MUTO_FP64_STREAM && recompute) {
      ctx->set_error("muto_map_forward: MUTO_FP64_STREAM reads the plan's lists; it does not walk the paths again");
      return fail(MUTO_ERR_ARG);
    }
    if (precision == MUTO_FP64_STREAM || (precision == MUTO_FP64 && !recompute && pl->fp64_kernel == 2)) {
      // FP64 on the FP64 brick stream (MUTO_FP64_STREAM, or MUTO_OPT_FP64_KERNEL 2)
      if (sv->g64.alloc(4 * nt) != hipSuccess) { ctx->set_error("muto_map_forward: out of device memory"); return fail(MUTO_ERR_NOMEM); }
      if (int rc = fp64s_map_forward(pl, sv->lam.p, sv->g64.p, &cost_sum)) return fail(rc);
      sv->stream64 = true;
    } else if (precision == MUTO_FP64 || recompute) {
      if (sv->g64.alloc(4 * nt) != hipSuccess) { ctx->set_error("muto_map_forward: out of device memory"); return fail(MUTO_ERR_NOMEM); }
      const MapTrackArgs a = track_args(pl);
      if (recompute && mt_build(pl)) {
        // paths walked again per tile: each run's moments, then per track
        k_mt_reset<<<1, 64, 0, s>>>(pl->mt_work.p);
        MUTO_CHECK_HIP(ctx, hipGetLastError());
        k_mt_fwd<<<pl->mt_grid_fwd, MT_BS, 0, s>>>(a, mt_geo(pl), pl->mt_task.p, pl->mt_n_task, pl->mt_work.p,
                                                   pl->mt_rec.p, pl->mt_orig.p, sv->lam.p, pl->mt_part.p);
        MUTO_CHECK_HIP(ctx, hipGetLastError());
        k_mt_track<<<grid_for(nt, 256), 256, 0, s>>>(a, pl->mt_roff.p, pl->mt_part.p, pl->mt_plen.p, sv->g64.p,
                                                     sv->part.p, pl->dcount.p);
      } else if (recompute) {
        k_map_fwd64<1><<<grid_for(nt, 256), 256, 0, s>>>(a, sv->lam.p, sv->g64.p, sv->part.p, pl->dcount.p);
      } else {
        k_map_fwd64<0><<<grid_for(nt, 256), 256, 0, s>>>(a, sv->lam.p, sv->g64.p, sv->part.p, pl->dcount.p);
      }
      MUTO_CHECK_HIP(ctx, hipGetLastError());
      k_sum_parts<<<1, 256, 0, s>>>(sv->part.p, grid_for(nt, 256), sv->sum.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
      MUTO_CHECK_HIP(ctx, hipMemcpyAsync(&cost_sum, sv->sum.p, sizeof(double), hipMemcpyDeviceToHost, s));
      MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
    } else {
      if (int rc = fp32c_build(pl)) return fail(rc);
      if (sv->g32.alloc(4 * nt) != hipSuccess || sv->cmax.alloc(1) != hipSuccess) {
        ctx->set_error("muto_map_forward: out of device memory");
        return fail(MUTO_ERR_NOMEM);
      }
      if (int rc = fp32c_map_forward(pl, sv->lam.p, sv->g32.p, sv->cmax.p, &cost_sum)) return fail(rc);
    }
  }
  double prior = 0.0;
  prior_phi(map_cfg(pl), nv, sv->lam.p, sv->part.p, s);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  k_sum_parts<<<1, 256, 0, s>>>(sv->part.p, grid_for(nv, 256), sv->sum.p + 1);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(&prior, sv->sum.p + 1, sizeof(double), hipMemcpyDeviceToHost, s));
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  sv->phi = 0.5 * cost_sum + pl->cfg.beta * prior;
  *phi = sv->phi;
  *saved = sv;
  return MUTO_OK;
}

extern "C" int muto_map_backward(muto_map_saved* sv, double grad_phi, double* grad_lambda) {
  if (!sv || !sv->pl) return MUTO_ERR_ARG;
  muto_em_plan* pl = sv->pl;
  muto_ctx* ctx = pl->ctx;
  if (!grad_lambda) { ctx->set_error("muto_map_backward: null argument"); return MUTO_ERR_ARG; }
  hipStream_t s = ctx->stream;
  const size_t nv = pl->nv, nt = pl->n_trk;
  // device buffers: the gradient accumulates in the caller's array
  double* grad = sv->dev ? grad_lambda : sv->grad.p;
  k_zero_f64<<<grid_for(nv, 256), 256, 0, s>>>(nv, grad);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  if (nt) {
    if (sv->stream64) {
      if (int rc = fp64s_map_backward(pl, sv->g64.p, grad)) return rc;
      k_scale<<<grid_for(nv, 256), 256, 0, s>>>(nv, -0.5 * grad_phi, grad);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    } else if (sv->recompute && mt_build(pl)) {
      // paths walked again, summed per tile in LDS (FP64 or FP32 terms)
      k_mt_reset<<<1, 64, 0, s>>>(pl->mt_work.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
      const MapTrackArgs a = track_args(pl);
      if (sv->precision == MUTO_FP64)
        k_mt_bwd<false><<<pl->mt_grid, MT_BS, 0, s>>>(a, mt_geo(pl), pl->mt_task.p, pl->mt_n_task,
                                                      pl->mt_work.p, pl->mt_rec.p, sv->g64.p,
                                                      -0.5 * grad_phi, grad);
      else
        k_mt_bwd<true><<<pl->mt_grid, MT_BS, 0, s>>>(a, mt_geo(pl), pl->mt_task.p, pl->mt_n_task,
                                                     pl->mt_work.p, pl->mt_rec.p, sv->g64.p,
                                                     -0.5 * grad_phi, grad);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    } else if (sv->precision == MUTO_FP64) {
      const MapTrackArgs a = track_args(pl);
      if (sv->recompute)
        k_map_bwd64<1><<<grid_for(nt, 256), 256, 0, s>>>(a, sv->g64.p, -0.5 * grad_phi, grad);
      else
        k_map_bwd64<0><<<grid_for(nt, 256), 256, 0, s>>>(a, sv->g64.p, -0.5 * grad_phi, grad);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    } else if (sv->recompute) {
      // FP32C, paths walked again: FP32 entry terms and FP32 atomics into a
      // float accumulator, then widened into the gradient
      if (!sv->acc32.p && sv->acc32.alloc(nv) != hipSuccess) {
        ctx->set_error("muto_map_backward: out of device memory");
        return MUTO_ERR_NOMEM;
      }
      k_zero_f32<<<grid_for(nv, 256), 256, 0, s>>>(nv, sv->acc32.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
      k_map_bwd32rc<<<grid_for(nt, 256), 256, 0, s>>>(track_args(pl), sv->g64.p, -0.5 * grad_phi,
                                                      sv->acc32.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
      k_widen_f32<<<grid_for(nv, 256), 256, 0, s>>>(nv, sv->acc32.p, grad);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    } else {
      if (int rc = fp32c_map_backward(pl, sv->g32.p, sv->cmax.p, grad)) return rc;
      k_scale<<<grid_for(nv, 256), 256, 0, s>>>(nv, -0.5 * grad_phi, grad);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    }
  }
  prior_grad(map_cfg(pl), nv, sv->lam.p, grad_phi * pl->cfg.beta, grad, s);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  if (!sv->dev)
    MUTO_CHECK_HIP(ctx, hipMemcpyAsync(grad_lambda, grad, nv * sizeof(double), hipMemcpyDeviceToHost, s));
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  return MUTO_OK;
}

extern "C" int muto_map_saved_destroy(muto_map_saved* sv) {
  if (!sv) return MUTO_OK;
  if (sv->pl && sv->pl->ctx) (void)hipStreamSynchronize(sv->pl->ctx->stream);
  delete sv;
  return MUTO_OK;
}

// ------------------------------------------------------------------ //
// Phi and its gradient without allocating (muto_map_value_grad_noalloc) //
// ------------------------------------------------------------------ //
// The entry point's only memory beyond its caller's buffers and the plan:
// the per-block partial sums of the track costs and of the prior, and their
// two totals.  One copy per GPU (each device holds its own, at its own
// address), kept until the process ends; every call on a GPU reuses it, so
// calls that overlap on one GPU race.
constexpr size_t NOALLOC_PARTS = 65536;  // at most 2^24 tracks and 2^24 voxels
constexpr size_t NOALLOC_DOUBLES = 2 * NOALLOC_PARTS + 32;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
hipError_t g_scratch_err[NOALLOC_MAX_DEVICES] = {};
}  // namespace

namespace muto {
// Called by muto_ctx_create with `device` current: the device address of
// this GPU's copy of `scratch` (host code must never pass the symbol itself
// to a kernel or a copy).
hipError_t noalloc_scratch_init(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  g_scratch_err[device] = e;
  if (e != hipSuccess) return e;
  g_scratch_ptr[device] = static_cast<double*>(ptr);
  return hipSuccess;
}
}  // namespace muto

extern "C" int muto_map_value_grad_noalloc(muto_em_plan* pl, const double* lambda, int flags,
                                           double* phi, double* grad_lambda) {
  if (!pl) return MUTO_ERR_ARG;
  muto_ctx* ctx = pl->ctx;
  if (!lambda || !phi || !grad_lambda) {
    ctx->set_error("muto_map_value_grad_noalloc: null argument");
    return MUTO_ERR_ARG;
  }
  if (flags & ~MUTO_MAP_RECOMPUTE_SIDDON) {
    ctx->set_error("muto_map_value_grad_noalloc: unknown flag (the buffers are always device memory)");
    return MUTO_ERR_ARG;
  }
  const size_t nv = pl->nv, nt = pl->n_trk;
  if (grid_for(nt, 256) > NOALLOC_PARTS || grid_for(nv, 256) > NOALLOC_PARTS) {
    ctx->set_error("muto_map_value_grad_noalloc: more than 2^24 tracks or 2^24 voxels (the size of "
                   "its static scratch); use muto_map_forward / muto_map_backward");
    return MUTO_ERR_LIMIT;
  }
  const bool recompute = (flags & MUTO_MAP_RECOMPUTE_SIDDON) || !pl->has_lists;
  if (recompute && !pl->geo.p && nt) {
    ctx->set_error("muto_map_value_grad_noalloc: the plan has no track geometry to walk (prepared plan)");
    return MUTO_ERR_STATE;
  }
  double* scr = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (ctx->device >= 0 && ctx->device < NOALLOC_MAX_DEVICES) scr = g_scratch_ptr[ctx->device];
  }
  if (!scr) {
    const hipError_t e = ctx->device >= 0 && ctx->device < NOALLOC_MAX_DEVICES
                             ? g_scratch_err[ctx->device] : hipErrorInvalidDevice;
    ctx->set_error(std::string("muto_map_value_grad_noalloc: no scratch address on this device: ") +
                   hipGetErrorString(e));
    return MUTO_ERR_HIP;
  }
  double* part_t = scr;                  // [NOALLOC_PARTS] track cost partials
  double* part_p = scr + NOALLOC_PARTS;  // [NOALLOC_PARTS] prior partials
  double* sums = scr + 2 * NOALLOC_PARTS;  // [2] cost sum, prior sum
  hipStream_t s = ctx->stream;
  k_zero_f64<<<grid_for(nv, 256), 256, 0, s>>>(nv, grad_lambda);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  k_zero_f64<<<1, 256, 0, s>>>(2, sums);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  if (nt) {
    const MapTrackArgs a = track_args(pl);
    if (recompute)
      k_map_fb64<1><<<grid_for(nt, 256), 256, 0, s>>>(a, lambda, -0.5, grad_lambda, part_t, pl->dcount.p);
    else
      k_map_fb64<0><<<grid_for(nt, 256), 256, 0, s>>>(a, lambda, -0.5, grad_lambda, part_t, pl->dcount.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
    k_sum_parts<<<1, 256, 0, s>>>(part_t, grid_for(nt, 256), sums);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
  }
  const MapCfgD mc = map_cfg(pl);
  prior_phi(mc, nv, lambda, part_p, s);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  k_sum_parts<<<1, 256, 0, s>>>(part_p, grid_for(nv, 256), sums + 1);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  prior_grad(mc, nv, lambda, pl->cfg.beta, grad_lambda, s);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  k_phi_combine<<<1, 64, 0, s>>>(sums, pl->cfg.beta, phi);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  return MUTO_OK;
}
