// MLSD/EM and MAP on the GPU.
//
// Plan set-up (once): MLSD selection and per-track angles/displacements
// (src/MutoMLSD.cpp:268-376), two-leg Siddon paths merged at the PoCA voxel
// and their tails (src/MutoMLSD.cpp:382-415), per-voxel entry counts.
//
// FP64 iteration (reference order, one lane per track): covariance
// Sigma_i = p^2 sum_j lambda_j W_ij + E accumulated in path order, Eigen's
// 2x2 determinant/inverse, S_ij = 2 lambda_j + (xy/2 - tr) p^2 lambda_j^2
// (src/MutoMLSD.cpp:442-491), then the per-voxel mean / median / trimmed
// mean (src/MutoMLSD.cpp:493-517).  Same operations in the same order as
// the CPU reference, so per-track values are bit-identical; only the
// order of the per-voxel mean sum (FP64 atomics) differs.
//
// FP32C iteration: see em32.hip.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "em_dev.hpp"
#include "em_plan.hpp"
#include "geom_dev.hpp"

int muto_plan_alloc_common(muto_em_plan* pl);  // shared with batch.hip
int muto_read_counters(muto_em_plan* pl);

namespace muto {

// ------------------------------------------------------------------ //
// host helpers shared with the other kernels' files                  //
// ------------------------------------------------------------------ //
ErrM error_matrix(const muto_mlsd_cfg& c) {  // MutoMLSD.cpp:168-170
  const double e = c.error_xy, o = c.error_dz_outer, i = c.error_dz_inner;
  ErrM E;
  E.e00 = 4.0 * e * e / (o * o);
  E.e11 = 2.0 * e * e * (1.0 + (i / o + 1.0) * i / o);
  E.e01 = 2.0 * e * e * i / (o * o);
  return E;
}

namespace {

// ------------------------------------------------------------------ //
// plan set-up kernels                                                //
// ------------------------------------------------------------------ //
__global__ void __launch_bounds__(256)
k_select(GridD g, const muto_ray* __restrict__ rays, size_t n, double eps,
         uint32_t* __restrict__ flag, double* __restrict__ pscat) {
  __shared__ __align__(16) muto_ray sh[256];
  stage_rays<256>(rays, n, sh);
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const muto_ray r = sh[threadIdx.x];
  double P[3];
  poca_point(r, eps, P);
  const bool a = mlsd_accept(g, P, eps, rays[0].pout[2], rays[0].pin[2]);
  flag[i] = a ? 1u : 0u;
  pscat[3 * i] = P[0];
  pscat[3 * i + 1] = P[1];
  pscat[3 * i + 2] = P[2];
}

__global__ void __launch_bounds__(256)
k_compact(GridD g, const muto_ray* __restrict__ rays, size_t n,
          const uint32_t* __restrict__ flag,
          const uint64_t* __restrict__ pos, const double* __restrict__ pscat,
          int has_energy, uint32_t* __restrict__ sel,
          double* __restrict__ sel_p, double* __restrict__ trk4,
          double* __restrict__ pr, double* __restrict__ geo) {
  __shared__ __align__(16) muto_ray sh[256];
  stage_rays<256>(rays, n, sh);
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i]) return;
  const size_t s = pos[i];
  sel[s] = (uint32_t)i;
  sel_p[3 * s] = pscat[3 * i];
  sel_p[3 * s + 1] = pscat[3 * i + 1];
  sel_p[3 * s + 2] = pscat[3 * i + 2];
  const muto_ray r = sh[threadIdx.x];
  for (int k = 0; k < 3; ++k) {  // path endpoints for recomputing the walk
    geo[9 * s + k] = r.pin[k];
    geo[9 * s + 3 + k] = r.pout[k];
    geo[9 * s + 6 + k] = pscat[3 * i + k];
  }
  // getScatteringAngle (MutoMLSD.cpp:350-354)
  trk4[4 * s] = (atan2(r.din[0], r.din[2]) - atan2(r.dout[0], r.dout[2]));
  trk4[4 * s + 1] = (atan2(r.din[1], r.din[2]) - atan2(r.dout[1], r.dout[2]));
  // getDisplacement (MutoMLSD.cpp:356-376)
  const double axin = atan2(r.din[0], r.din[2]);
  const double axout = atan2(r.dout[0], r.dout[2]);
  const double ayin = atan2(r.din[1], r.din[2]);
  const double ayout = atan2(r.dout[1], r.dout[2]);
  const double sp = (g.mn[2] - r.pin[2]) / r.din[2];
  const double st = (g.mn[2] - r.pout[2]) / r.dout[2];
  const double xp = r.pin[0] + sp * r.din[0];
  const double yp = r.pin[1] + sp * r.din[1];
  const double xt = r.pout[0] + st * r.dout[0];
  const double yt = r.pout[1] + st * r.dout[1];
  const double tx = tan(axin), ty = tan(ayin);
  const double lxy = sqrt(1 + tx * tx + ty * ty);
  trk4[4 * s + 2] = (xt - xp) * cos(axin) * lxy * cos(axout) / cos(axout - axin);
  trk4[4 * s + 3] = (yt - yp) * cos(ayin) * lxy * cos(ayout) / cos(ayout - ayin);
  pr[s] = has_energy ? 3.0 / r.ein : 3.0 / 3.0;
}

// Merged two-leg path of each selected track (MutoMLSD.cpp:382-415).
// k_paths_count: entry counts only.  k_paths_fill: voxel indices and
// lengths.  Each lane walks its own track, one leg after the other on one
// cursor, up to PATH_STAGE entries at a time into an LDS stage; then the
// wavefront writes every lane's staged run together, PATH_STAGE lanes per
// source lane.  As in the Siddon CSR fill, a window ends where the lane's
// output position reaches a multiple of PATH_STAGE and its entries sit at
// their position modulo PATH_STAGE, so that later windows write whole
// aligned runs.
constexpr int PATH_STAGE = 16;
constexpr uint32_t PATH_BS = 64;

__global__ void __launch_bounds__(256)
k_paths_count(GridD g, const muto_ray* __restrict__ rays,
              const uint32_t* __restrict__ sel, const double* __restrict__ sel_p,
              size_t n_trk, int straight, uint64_t* __restrict__ counts) {
  const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_trk) return;
  muto_ray r;
  load_ray(rays, sel[s], r);
  const double P[3] = {sel_p[3 * s], sel_p[3 * s + 1], sel_p[3 * s + 2]};
  uint64_t c;
  if (straight) {
    c = siddon_walk(g, r.pin, r.pout, [](uint32_t, uint32_t, uint32_t, double) {});
  } else {
    uint64_t ki = siddon_walk(g, r.pin, P, [](uint32_t, uint32_t, uint32_t, double) {});
    uint64_t ko = siddon_walk(g, P, r.pout, [](uint32_t, uint32_t, uint32_t, double) {});
    c = (ki ? ki : 0) + (ko ? ko - 1 : 0);
  }
  counts[s] = c;
}

__global__ void __launch_bounds__(PATH_BS)
k_paths_fill(GridD g, const muto_ray* __restrict__ rays,
             const uint32_t* __restrict__ sel, const double* __restrict__ sel_p,
             size_t n_trk, int straight, const uint64_t* __restrict__ row_ptr,
             uint32_t* __restrict__ vox, double* __restrict__ len,
             unsigned long long* __restrict__ cnt_out) {
  // slot k of thread t at k * PATH_BS + t
  __shared__ uint32_t st_j[PATH_STAGE * PATH_BS];
  __shared__ double st_l[PATH_STAGE * PATH_BS];
  const uint32_t tid = threadIdx.x;
  const int lane = (int)(tid & 63u);
  const uint32_t wbase = tid & ~63u;
  const size_t s = (size_t)blockIdx.x * PATH_BS + tid;
  const uint32_t nx = g.n[0], nxy = g.n[0] * g.n[1];
  uint32_t nout = 0;
  SiddonCursor c;
  bool live = false;
  bool in_leg = false;      // the cursor walks the in-leg (the out-leg follows)
  bool merge_next = false;  // the next recorded voxel is the out-leg's first
  uint32_t ki = 0;          // entries the in-leg recorded
  uint64_t o = 0, e = 0;
  double P[3] = {0.0, 0.0, 0.0}, pout[3] = {0.0, 0.0, 0.0};
  if (s < n_trk) {
    muto_ray r;
    load_ray(rays, sel[s], r);
    for (int k = 0; k < 3; ++k) {
      P[k] = sel_p[3 * s + k];
      pout[k] = r.pout[k];
    }
    o = row_ptr[s];
    e = row_ptr[s + 1];
    if (straight) {
      live = c.begin(g, r.pin, r.pout);
    } else {
      in_leg = c.begin(g, r.pin, P);
      live = in_leg;
      if (!live) {  // an empty in-leg: the out-leg's first entry is dropped
        merge_next = true;
        live = c.begin(g, P, pout);
      }
    }
  }
  while (__any(live)) {
    const int sh = (int)(o % (uint64_t)PATH_STAGE);
    int cnt = 0;
    auto put = [&](uint32_t x, uint32_t y, uint32_t z, double L) {
      if (merge_next) {
        // in.back().length += out[0].length; out[0] is dropped.  The
        // in-leg's last entry is staged in this window, or was written by
        // an earlier flush of this wavefront (then wait for that store)
        merge_next = false;
        if (ki) {
          if (cnt > 0) {
            st_l[(uint32_t)(sh + cnt - 1) * PATH_BS + tid] += L;
          } else {
            __threadfence_block();
            len[o - 1] += L;
          }
        }
        return;
      }
      if (x >= g.n[0] || y >= g.n[1] || z >= g.n[2]) ++nout;
      const uint32_t sl = (uint32_t)(sh + cnt) * PATH_BS + tid;
      st_j[sl] = x + y * nx + z * nxy;  // xyz2index, u32 arithmetic
      st_l[sl] = L;
      ++cnt;
      if (in_leg) ++ki;
    };
    while (live && cnt < PATH_STAGE - sh) {
      c.run((uint32_t)(PATH_STAGE - sh - cnt), put);
      if (!c.live) {
        if (in_leg) {
          in_leg = false;
          merge_next = true;
          live = c.begin(g, P, pout);
        } else {
          live = false;
        }
      }
    }
    const uint64_t rem = e > o ? e - o : 0;
    const int hi = sh + ((uint64_t)cnt < rem ? cnt : (int)rem);
    const uint64_t ob = o - (uint64_t)sh;
#pragma unroll
    for (int rd = 0; rd < PATH_STAGE; ++rd) {
      const int src = rd * (64 / PATH_STAGE) + lane / PATH_STAGE, k = lane % PATH_STAGE;
      const int kse = __shfl(sh | (hi << 8), src, 64), ks = kse & 255, ke = kse >> 8;
      const uint64_t os = (uint64_t)(uint32_t)__shfl((int)(uint32_t)ob, src, 64) |
                          ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(ob >> 32), src, 64) << 32);
      if (k >= ks && k < ke) {
        const uint32_t sl = (uint32_t)k * PATH_BS + wbase + (uint32_t)src;
        vox[os + k] = st_j[sl];
        len[os + k] = st_l[sl];
      }
    }
    o += (uint64_t)cnt;
  }
  if (nout) atomicAdd(&cnt_out[C_OUT], (unsigned long long)nout);
}

// Tails T_k = sum of the lengths after entry k (MutoMLSD.cpp:403-414, the
// reference's backward running sum, same order), one lane per track.  Each
// lane takes its list from the end in windows that stop at 16-entry
// boundaries of the list array; the wavefront reads every lane's window
// into LDS together (16 lanes per window, whole aligned runs), each lane
// turns its own window's lengths into tails in place, and the wavefront
// writes them back the same way.
constexpr int TAIL_STAGE = 16;
constexpr uint32_t TAIL_BS = 256;
__global__ void __launch_bounds__(TAIL_BS)
k_tails(size_t n_trk, const uint64_t* __restrict__ row_ptr,
        const double* __restrict__ len, double* __restrict__ tail,
        unsigned long long* __restrict__ cnt) {
  __shared__ double st[TAIL_STAGE * TAIL_BS];  // slot k of thread t at k * TAIL_BS + t
  const uint32_t tid = threadIdx.x;
  const int lane = (int)(tid & 63u);
  const uint32_t wbase = tid & ~63u;
  const size_t s = (size_t)blockIdx.x * blockDim.x + tid;
  uint64_t b = 0, p = 0;  // entries [b, p) still to do
  bool empty = false;
  if (s < n_trk) {
    b = row_ptr[s];
    p = row_ptr[s + 1];
    empty = (p == b);
  }
  double run = 0.0;  // the tail of entry p - 1
  while (__any(p > b)) {
    // the window [lo, p): within the aligned run holding entry p - 1
    const bool act = p > b;
    const uint64_t base = act ? (p - 1) / TAIL_STAGE * TAIL_STAGE : 0;
    const uint64_t lo = act ? (base > b ? base : b) : 0;
    const int ks = act ? (int)(lo - base) : 0, ke = act ? (int)(p - base) : 0;
#pragma unroll
    for (int r = 0; r < TAIL_STAGE; ++r) {
      const int src = r * (64 / TAIL_STAGE) + lane / TAIL_STAGE, k = lane % TAIL_STAGE;
      const int kse = __shfl(ks | (ke << 8), src, 64), k0 = kse & 255, k1 = kse >> 8;
      const uint64_t bs = (uint64_t)(uint32_t)__shfl((int)(uint32_t)base, src, 64) |
                          ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(base >> 32), src, 64) << 32);
      if (k >= k0 && k < k1) st[(uint32_t)k * TAIL_BS + wbase + (uint32_t)src] = len[bs + k];
    }
    if (act) {
      for (int k = ke - 1; k >= ks; --k) {  // tail[k - 1] = tail[k] + len[k]
        const double l = st[(uint32_t)k * TAIL_BS + tid];
        st[(uint32_t)k * TAIL_BS + tid] = run;
        run = run + l;
      }
    }
#pragma unroll
    for (int r = 0; r < TAIL_STAGE; ++r) {
      const int src = r * (64 / TAIL_STAGE) + lane / TAIL_STAGE, k = lane % TAIL_STAGE;
      const int kse = __shfl(ks | (ke << 8), src, 64), k0 = kse & 255, k1 = kse >> 8;
      const uint64_t bs = (uint64_t)(uint32_t)__shfl((int)(uint32_t)base, src, 64) |
                          ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(base >> 32), src, 64) << 32);
      if (k >= k0 && k < k1) tail[bs + k] = st[(uint32_t)k * TAIL_BS + wbase + (uint32_t)src];
    }
    p = lo;
  }
  wave_count(&cnt[C_EMPTY], empty);
}

// Entries per voxel, one thread per entry (entries outside the grid are
// not counted).
__global__ void __launch_bounds__(256)
k_mj_count(uint64_t nnz, const uint32_t* __restrict__ vox, size_t nv,
           uint32_t* __restrict__ mj) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const uint32_t j = vox[k];
  if (j < nv) atomicAdd(&mj[j], 1u);
}

// Per-voxel entry counts and the empty / out-of-grid counters from a walk
// of each path, for plans that keep no path lists.
__global__ void __launch_bounds__(256)
k_mj_walk(GridD g, const double* __restrict__ geo, size_t n_trk, int straight,
          uint32_t* __restrict__ mj, size_t nv, unsigned long long* __restrict__ cnt) {
  size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool empty = false;
  uint32_t nout = 0;
  if (s < n_trk) {
    uint32_t ne = 0;
    walk_path(g, geo + 9 * s, straight, [&](uint32_t j, double, double, bool oob) {
      ++ne;
      if (oob) ++nout;
      if (j < nv) atomicAdd(&mj[j], 1u);
    });
    empty = ne == 0;
  }
  wave_count(&cnt[C_EMPTY], empty);
  if (nout) atomicAdd(&cnt[C_OUT], (unsigned long long)nout);
}

// voxel-major slots for the order statistics (allocateSValueMatrix,
// MutoMLSD.cpp:432-436).  Slot order inside a voxel does not affect the
// median or the sorted trimmed mean.
__global__ void __launch_bounds__(256)
k_slots(uint64_t nnz, const uint32_t* __restrict__ vox, size_t nv,
        const uint64_t* __restrict__ col_ptr, uint32_t* __restrict__ fill,
        uint64_t* __restrict__ slot) {
  uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const uint32_t j = vox[k];
  slot[k] = j < nv ? col_ptr[j] + atomicAdd(&fill[j], 1u) : ~0ull;
}

// ------------------------------------------------------------------ //
// FP64 iteration                                                     //
// ------------------------------------------------------------------ //
// mode 0: accumulate sum_i S_ij into acc (mean of all);
// mode 1: write S_ij into its voxel-major slot (median / trimmed mean).
template <int MODE>
__global__ void __launch_bounds__(256)
k_em64(size_t n_trk, const uint64_t* __restrict__ row_ptr,
       const uint32_t* __restrict__ vox, const double* __restrict__ len,
       const double* __restrict__ tail, const double* __restrict__ trk4,
       const double* __restrict__ pr, const double* __restrict__ lam,
       size_t nv, ErrM E, double* __restrict__ acc,
       const uint64_t* __restrict__ slot, double* __restrict__ S,
       double* __restrict__ cost_part, unsigned long long* __restrict__ cnt) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double cost = 0.0;
  bool sing = false;
  if (i < n_trk) {
    const double p = pr[i];
    const uint64_t b = row_ptr[i], e = row_ptr[i + 1];
    // getCovarianceMatrix (MutoMLSD.cpp:442-449) + ErrorM
    Mat2 cov = {0.0, 0.0, 0.0, 0.0};
    for (uint64_t k = b; k < e; ++k) {
      Mat2 W;
      weight(len[k], tail[k], W);
      const uint32_t j = vox[k];
      const double lj = j < nv ? lam[j] : 0.0;
      cov.m00 = cov.m00 + lj * W.m00;
      cov.m01 = cov.m01 + lj * W.m01;
      cov.m10 = cov.m10 + lj * W.m10;
      cov.m11 = cov.m11 + lj * W.m11;
    }
    const double pp = p * p;
    cov.m00 = pp * cov.m00 + E.e00;
    cov.m01 = pp * cov.m01 + E.e01;
    cov.m10 = pp * cov.m10 + E.e01;
    cov.m11 = pp * cov.m11 + E.e11;
    // updateSValueMatrix (MutoMLSD.cpp:451-491)
    const double Sx = trk4[4 * i], Sy = trk4[4 * i + 1];
    const double Dx = trk4[4 * i + 2], Dy = trk4[4 * i + 3];
    Mat2 inv;
    double det;
    const bool ok = inv_det(cov, inv, det);
    if (!ok || det <= 0) {
      sing = true;
    } else {
      cost = log(det) + ((Sx * Sx + Sy * Sy) * inv.m00 +
                         (Dx * Dx + Dy * Dy) * inv.m11 +
                         (Sx * Dx + Sy * Dy) * 2.0 * inv.m01) * 0.5;
    }
    for (uint64_t k = b; k < e; ++k) {
      const uint32_t j = vox[k];
      if (j >= nv) continue;
      Mat2 W;
      weight(len[k], tail[k], W);
      const double lj = lam[j];
      const double pr2lj2 = p * p * lj * lj;
      const double P00 = inv.m00 * W.m00 + inv.m01 * W.m10;
      const double P01 = inv.m00 * W.m01 + inv.m01 * W.m11;
      const double P10 = inv.m10 * W.m00 + inv.m11 * W.m10;
      const double P11 = inv.m10 * W.m01 + inv.m11 * W.m11;
      const double tr = P00 + P11;
      const double C00 = P00 * inv.m00 + P01 * inv.m10;
      const double C01 = P00 * inv.m01 + P01 * inv.m11;
      const double C10 = P10 * inv.m00 + P11 * inv.m10;
      const double C11 = P10 * inv.m01 + P11 * inv.m11;
      const double xy = (Sx * Sx + Sy * Sy) * C00 + (Dx * Dx + Dy * Dy) * C11 +
                        (Sx * Dx + Sy * Dy) * (C01 + C10);
      const double sij = 2.0 * lj + (0.5 * xy - tr) * pr2lj2;
      if (MODE == 0)
        atomicAdd(&acc[j], sij);
      else
        S[slot[k]] = sij;
    }
  }
  // per-block cost sums, added up in a fixed order by k_sum_parts_f64 (one
  // atomic per track or wave on one address would serialise)
  block_sum_to(cost_part, cost);
  wave_count(&cnt[C_SINGULAR], sing);
}

// FP64 iteration, one wavefront per track (MUTO_OPT_FP64_KERNEL 1): the
// lanes read the track's row with coalesced loads, the covariance sums are
// combined by a wavefront butterfly (a different association from the
// reference's running sum, so per-track values differ by FP64 rounding), the
// 2x2 algebra is done by every lane on the reduced sums, and the second sweep
// over the row (now in cache) forms S_ij as k_em64 does.
template <int MODE>
__global__ void __launch_bounds__(256)
k_em64w(size_t n_trk, const uint64_t* __restrict__ row_ptr,
        const uint32_t* __restrict__ vox, const double* __restrict__ len,
        const double* __restrict__ tail, const double* __restrict__ trk4,
        const double* __restrict__ pr, const double* __restrict__ lam,
        size_t nv, ErrM E, double* __restrict__ acc,
        const uint64_t* __restrict__ slot, double* __restrict__ S,
        double* __restrict__ cost_part, unsigned long long* __restrict__ cnt) {
  const unsigned lane = threadIdx.x & 63;
  const size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  double cost = 0.0;
  bool sing = false;
  if (i < n_trk) {  // uniform over the wavefront
    const double p = pr[i];
    const uint64_t b = row_ptr[i], e = row_ptr[i + 1];
    double s00 = 0.0, s01 = 0.0, s11 = 0.0;
    for (uint64_t k = b + lane; k < e; k += 64) {
      Mat2 W;
      weight(len[k], tail[k], W);
      const uint32_t j = vox[k];
      const double lj = j < nv ? lam[j] : 0.0;
      s00 = s00 + lj * W.m00;
      s01 = s01 + lj * W.m01;
      s11 = s11 + lj * W.m11;
    }
    s00 = wave_sum(s00);
    s01 = wave_sum(s01);
    s11 = wave_sum(s11);
    const double pp = p * p;
    Mat2 cov;
    cov.m00 = pp * s00 + E.e00;
    cov.m01 = pp * s01 + E.e01;
    cov.m10 = cov.m01;
    cov.m11 = pp * s11 + E.e11;
    const double Sx = trk4[4 * i], Sy = trk4[4 * i + 1];
    const double Dx = trk4[4 * i + 2], Dy = trk4[4 * i + 3];
    Mat2 inv;
    double det;
    const bool ok = inv_det(cov, inv, det);
    if (!ok || det <= 0) {
      sing = lane == 0;
    } else if (lane == 0) {
      cost = log(det) + ((Sx * Sx + Sy * Sy) * inv.m00 +
                         (Dx * Dx + Dy * Dy) * inv.m11 +
                         (Sx * Dx + Sy * Dy) * 2.0 * inv.m01) * 0.5;
    }
    for (uint64_t k = b + lane; k < e; k += 64) {
      const uint32_t j = vox[k];
      if (j >= nv) continue;
      Mat2 W;
      weight(len[k], tail[k], W);
      const double lj = lam[j];
      const double pr2lj2 = p * p * lj * lj;
      const double P00 = inv.m00 * W.m00 + inv.m01 * W.m10;
      const double P01 = inv.m00 * W.m01 + inv.m01 * W.m11;
      const double P10 = inv.m10 * W.m00 + inv.m11 * W.m10;
      const double P11 = inv.m10 * W.m01 + inv.m11 * W.m11;
      const double tr = P00 + P11;
      const double C00 = P00 * inv.m00 + P01 * inv.m10;
      const double C01 = P00 * inv.m01 + P01 * inv.m11;
      const double C10 = P10 * inv.m00 + P11 * inv.m10;
      const double C11 = P10 * inv.m01 + P11 * inv.m11;
      const double xy = (Sx * Sx + Sy * Sy) * C00 + (Dx * Dx + Dy * Dy) * C11 +
                        (Sx * Dx + Sy * Dy) * (C01 + C10);
      const double sij = 2.0 * lj + (0.5 * xy - tr) * pr2lj2;
      if (MODE == 0)
        atomicAdd(&acc[j], sij);
      else
        S[slot[k]] = sij;
    }
  }
  // per-block cost sums, added up in a fixed order by k_sum_parts_f64 (one
  // atomic per track or wave on one address would serialise)
  block_sum_to(cost_part, cost);
  wave_count(&cnt[C_SINGULAR], sing);
}

// mean of all S (updateImageVoxel, fPercentileMean == 100)
__global__ void k_vox_mean64(size_t nv, const uint32_t* __restrict__ mj,
                             double* __restrict__ acc,
                             double* __restrict__ out) {
  size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nv) return;
  const uint32_t m = mj[j];
  out[j] = m == 0 ? 0.0 : acc[j] / (double)m * 0.5;
  acc[j] = 0.0;
}

// Bitonic sort of a[0..n) (n a power of two) by one block.
__device__ void block_bitonic(double* a, uint32_t n) {
  for (uint32_t k = 2; k <= n; k <<= 1) {
    for (uint32_t jj = k >> 1; jj > 0; jj >>= 1) {
      for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
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

// result of the order-statistic update from sorted a[0..m)
__device__ double order_update(const double* a, uint32_t m, int use_mean,
                               int pct) {
  if (use_mean) {  // trimmed mean: lowest floor(pct/100*M) values, >= 1
    uint32_t jm = x86_u32(pct / 100.0 * (double)m);
    if (jm < 1) jm = 1;
    double s = 0.0;
    for (uint32_t k = 0; k < jm; ++k) s = s + a[k];
    return s / (double)jm * 0.5;
  }
  const double med = a[m / 2];
  if (m % 2 == 0) return (med + a[m / 2 - 1]) * 0.25;
  return med * 0.5;
}

constexpr uint32_t SEL_CAP = 4096;

// one block per voxel, M_j <= SEL_CAP: sort in LDS
__global__ void __launch_bounds__(256)
k_vox_select(size_t nv, const uint32_t* __restrict__ mj,
             const uint64_t* __restrict__ col_ptr,
             const double* __restrict__ S, int use_mean, int pct,
             double* __restrict__ out) {
  __shared__ double a[SEL_CAP];
  for (size_t j = blockIdx.x; j < nv; j += gridDim.x) {
    const uint32_t m = mj[j];
    if (m == 0 || m > SEL_CAP) {
      if (threadIdx.x == 0 && m == 0) out[j] = 0.0;
      continue;
    }
    uint32_t n = 1;
    while (n < m) n <<= 1;
    const double* src = S + col_ptr[j];
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
      a[i] = i < m ? src[i] : INFINITY;
    __syncthreads();
    block_bitonic(a, n);
    if (threadIdx.x == 0) out[j] = order_update(a, m, use_mean, pct);
    __syncthreads();
  }
}

// voxels with M_j > SEL_CAP, listed in big[0..n_big): one block per listed
// voxel sorts its values in its own part of the global scratch
// (scratch + off[b], next power of two above M_j doubles)
__global__ void __launch_bounds__(1024)
k_vox_select_big(uint32_t n_big, const uint32_t* __restrict__ big,
                 const uint64_t* __restrict__ off, const uint32_t* __restrict__ mj,
                 const uint64_t* __restrict__ col_ptr, const double* __restrict__ S,
                 int use_mean, int pct, double* __restrict__ scratch,
                 double* __restrict__ out) {
  for (uint32_t b = blockIdx.x; b < n_big; b += gridDim.x) {
    const uint32_t j = big[b];
    const uint32_t m = mj[j];
    uint32_t n = 1;
    while (n < m) n <<= 1;
    double* a = scratch + off[b];
    const double* src = S + col_ptr[j];
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
      a[i] = i < m ? src[i] : INFINITY;
    __syncthreads();
    block_bitonic(a, n);
    if (threadIdx.x == 0) out[j] = order_update(a, m, use_mean, pct);
    __syncthreads();
  }
}

__global__ void k_copy_f64_to_f32(size_t n, const double* __restrict__ a,
                                  float* __restrict__ b) {
  size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) b[j] = (float)a[j];
}

}  // namespace

// ------------------------------------------------------------------ //
// MAP voxel step (src/MutoMAP.cpp:162-184, 207-407), one lane / voxel //
// ------------------------------------------------------------------ //
// pw[k] = pow(t_k^2 + delta^2, (p - 2) / 2) of each neighbour, the factor
// the coefficients and every candidate share (computed once per voxel;
// the expressions are the reference's, in its order)
template <int NN>
__device__ __forceinline__ double map_surrogate(const MapCfgD& c, uint32_t mj, double ljn,
                                                double ljml, double lj, const double (&lm)[NN],
                                                const double (&pw)[NN], double pw_obj) {
  double value = 0.0;  // surrogateFunctionValue, MutoMAP.cpp:261-281
  if (lj < 0) return 1.7976931348623157e308;
  value += 0 < lj ? (double)mj * (log(lj) + ljml / lj) : 0.0;
  const double w2 = 1.0 / sqrt(2.0), w3 = 1.0 / sqrt(3.0);
#pragma unroll
  for (int k = 0; k < NN; ++k) {
    const double w = k < 6 ? 1.0 : (k < 18 ? w2 : w3);
    const double u = 2.0 * lj - ljn - lm[k];
    value += 0.5 * c.beta * c.p * w * pw[k] * (u * u);
  }
  if (c.obj_prior)
    value += 0.5 * c.beta * c.p * lj * lj * pw_obj;
  return value;
}

// cubicEqnSolver (MutoMAP.cpp:353-407); Newton capped at 10000 steps
__device__ uint32_t cubic_solve(double* roots, double p3, double p2, double p1,
                                double p0, double tol, int* warn) {
  *warn = fabs(p3) < tol ? 1 : 0;
  const double a = p2 / p3, b = p1 / p3, c = p0 / p3;
  const double r = 1 + smax(smax(fabs(a), fabs(b)), fabs(c));
  const double xinfl = -a / 3;
  const double gx = xinfl * xinfl * xinfl + a * xinfl * xinfl + b * xinfl + c;
  double xk = 0 < gx ? -r : r;
  double rel = tol * 2;
  int guard = 0;
  while (rel > tol) {
    const double xk2 = xk * xk;
    const double gxk = xk2 * xk + a * xk2 + b * xk + c;
    const double gpxk = 3 * xk2 + 2 * a * xk + b;
    const double xn = xk - gxk / gpxk;
    rel = fabs((xn - xk) / xk);
    xk = xn;
    if (++guard >= 10000) { *warn |= 2; break; }
  }
  const double x1 = xk;
  const double D = x1 + a;
  const double E = D * x1 + b;
  const double signb = 0 <= D ? 1 : -1;
  const double diak = D * D - 4 * E;
  if (diak < 0) { roots[0] = x1; return 1; }
  const double q = -0.5 * (D + signb * sqrt(diak));
  roots[0] = x1;
  roots[1] = q;
  roots[2] = tol < fabs(q) ? E / q : q;
  return 3;
}

template <int NN>
__global__ void __launch_bounds__(256)
k_map_step(MapCfgD c, size_t nv, const uint32_t* __restrict__ mj,
           const double* __restrict__ last, const double* __restrict__ ml,
           double* __restrict__ out, unsigned long long* __restrict__ cnt) {
  size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  int warn = 0;
  if (j < nv) {
    const uint32_t m = mj[j];
    if (m == 0) {
      out[j] = 0.0;
    } else {
      double p0 = -ml[j] * MAP_SCALE, p1 = 1.0, p2 = 0.0, p3 = 0.0;
      // the neighbours' l and pow factors in registers (NN fixed, loops
      // unrolled)
      double lm[NN], pw[NN];
      const double lj = last[j] * MAP_SCALE;
      const double ex = (c.p - 2.0) / 2.0, dd = c.delta * c.delta;
      for_each_neighbour<NN>(c, (uint32_t)j, [&](int k, size_t idx, double w) {
        lm[k] = last[idx] * MAP_SCALE;
        const double t = lj - lm[k];
        pw[k] = pow(t * t + dd, ex);
        const double t1 = w * c.p * pw[k];
        p2 += t1 * 2.0 * (lj + lm[k]);
        p3 += t1 * 4.0;
      });
      const double pw_obj = c.obj_prior ? pow(lj * lj + dd, ex) : 0.0;
      if (c.obj_prior) p3 += c.p * pw_obj;
      p2 *= -c.beta / (double)m;
      p3 *= c.beta / (double)m;
      double roots[3];
      const uint32_t nr = cubic_solve(roots, p3, p2, p1, p0, c.eps, &warn);
      const double ljn = last[j] * MAP_SCALE, ljml = ml[j] * MAP_SCALE;
      double best = 0.0;
      uint32_t bi = 0;
      for (uint32_t r = 0; r <= nr; ++r) {
        const double cand = r < nr ? roots[r] : ljml;
        const double f = map_surrogate<NN>(c, m, ljn, ljml, cand, lm, pw, pw_obj);
        if (r == 0 || f < best) { best = f; bi = r; }
      }
      out[j] = bi < nr ? roots[bi] / MAP_SCALE : ml[j];
    }
  }
  wave_count(&cnt[C_CUBIC], (warn & 1) != 0);
  wave_count(&cnt[C_NEWTON], (warn & 2) != 0);
}

MapCfgD map_cfg(const muto_em_plan* pl) {
  MapCfgD c;
  c.g = pl->g;
  c.beta = pl->cfg.beta;
  c.p = pl->cfg.p;
  c.delta = pl->cfg.delta;
  c.eps = pl->cfg.eps;
  c.obj_prior = pl->cfg.add_object_prior;
  c.all_nb = pl->cfg.all_neighbours;
  return c;
}

// FP32C iteration (em32.hip)
int fp32c_build(muto_em_plan* pl);
int fp32c_run(muto_em_plan* pl, int iters, double* sums, int* done);
int fp32c_autotune(muto_em_plan* pl, int iters, int n_cand, const uint32_t* cand, double* ms);
int fp32c_release(muto_em_plan* pl);
int fp64s_run(muto_em_plan* pl, int iters, double* sums, int* done);
int fp64s_build(muto_em_plan* pl);

}  // namespace muto

using namespace muto;

// ------------------------------------------------------------------ //
// plan creation                                                      //
// ------------------------------------------------------------------ //
int muto_plan_alloc_common(muto_em_plan* pl) {
  muto_ctx* ctx = pl->ctx;
  MUTO_CHECK_HIP(ctx, pl->lam64.alloc(pl->nv));
  MUTO_CHECK_HIP(ctx, pl->ml64.alloc(pl->nv));
  MUTO_CHECK_HIP(ctx, pl->acc64.alloc(pl->nv));
  MUTO_CHECK_HIP(ctx, pl->cost_acc.alloc(4));
  MUTO_CHECK_HIP(ctx, pl->dcount.alloc(8));
  MUTO_CHECK_HIP(ctx, pl->work.alloc(4));
  hipStream_t s = ctx->stream;
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(pl->acc64.p, 0, pl->nv * sizeof(double), s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(pl->lam64.p, 0, pl->nv * sizeof(double), s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(pl->cost_acc.p, 0, 4 * sizeof(double), s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(pl->work.p, 0, 4 * sizeof(unsigned int), s));
  for (int k = 0; k < 8; ++k) MUTO_CHECK_HIP(ctx, hipEventCreate(&pl->ev[k]));
  return MUTO_OK;
}

static int validate_cfg(muto_ctx* ctx, const muto_mlsd_cfg* cfg) {
  if (!cfg) { ctx->set_error("null config"); return MUTO_ERR_ARG; }
  const muto_grid& g = cfg->grid;
  if (g.nx == 0 || g.ny == 0 || g.nz == 0) { ctx->set_error("empty grid"); return MUTO_ERR_ARG; }
  if ((uint64_t)g.nx * g.ny * g.nz > 0xFFFFFFFFull) {
    ctx->set_error("grid larger than 2^32 voxels (32-bit voxel index)");
    return MUTO_ERR_LIMIT;
  }
  if (!(g.dx > 0 && g.dy > 0 && g.dz > 0)) { ctx->set_error("voxel sizes must be positive"); return MUTO_ERR_ARG; }
  if (cfg->precision != MUTO_FP64 && cfg->precision != MUTO_FP32C && cfg->precision != MUTO_FP64_STREAM) {
    ctx->set_error("precision must be MUTO_FP64, MUTO_FP32C or MUTO_FP64_STREAM");
    return MUTO_ERR_ARG;
  }
  return MUTO_OK;
}

int muto_read_counters(muto_em_plan* pl) {
  unsigned long long c[8];
  MUTO_CHECK_HIP(pl->ctx, hipMemcpyAsync(c, pl->dcount.p, sizeof c, hipMemcpyDeviceToHost, pl->ctx->stream));
  MUTO_CHECK_HIP(pl->ctx, hipStreamSynchronize(pl->ctx->stream));
  pl->counters.n_singular = (int64_t)c[C_SINGULAR];
  pl->counters.n_cubic_warn = (int64_t)c[C_CUBIC];
  pl->counters.n_newton_cap = (int64_t)c[C_NEWTON];
  pl->counters.n_out_of_grid = (int64_t)c[C_OUT];
  pl->counters.n_empty_path = (int64_t)c[C_EMPTY];
  pl->counters.n_median_miss = (int64_t)c[C_MISS];
  return MUTO_OK;
}

extern "C" int muto_em_plan_create(muto_ctx* ctx, const muto_mlsd_cfg* cfg,
                                   size_t n, const muto_ray* rays,
                                   muto_em_plan** out) {
  return muto_em_plan_create_ex(ctx, cfg, n, rays, UINT64_MAX, out);
}

extern "C" int muto_em_plan_create_ex(muto_ctx* ctx, const muto_mlsd_cfg* cfg,
                                      size_t n, const muto_ray* rays, uint64_t list_budget,
                                      muto_em_plan** out) {
  if (!ctx) return MUTO_ERR_ARG;
  if (!out) { ctx->set_error("muto_em_plan_create: null out"); return MUTO_ERR_ARG; }
  *out = nullptr;
  if (int rc = validate_cfg(ctx, cfg)) return rc;
  if (n && !rays) { ctx->set_error("muto_em_plan_create: null rays"); return MUTO_ERR_ARG; }
  auto* pl = new muto_em_plan();
  pl->ctx = ctx;
  pl->cfg = *cfg;
  pl->g = to_dev(cfg->grid);
  pl->nv = n_vox(cfg->grid);
  pl->list_budget = list_budget;
  int rc = muto_plan_alloc_common(pl);
  if (rc) { delete pl; return rc; }
  hipStream_t s = ctx->stream;
  size_t ntot = cfg->total_num_rays > 0 ? (size_t)cfg->total_num_rays : n;
  if (ntot > n) ntot = n;
  pl->counters.n_total = (int64_t)ntot;
  MUTO_CHECK_HIP(ctx, pl->dcount.alloc(8));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(pl->dcount.p, 0, 8 * sizeof(unsigned long long), s));
  MUTO_CHECK_HIP(ctx, pl->mj.alloc(pl->nv));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(pl->mj.p, 0, pl->nv * sizeof(uint32_t), s));
  if (ntot == 0) {
    MUTO_CHECK_HIP(ctx, pl->row_ptr.alloc(1));
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(pl->row_ptr.p, 0, sizeof(uint64_t), s));
    MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
    *out = pl;
    return MUTO_OK;
  }
  {
    DBuf<muto_ray> dr;
    DBuf<uint32_t> flag, sel;
    DBuf<uint64_t> pos, counts;
    DBuf<double> pscat, sel_p;
    MUTO_CHECK_HIP(ctx, dr.alloc(ntot));
    MUTO_CHECK_HIP(ctx, flag.alloc(ntot));
    MUTO_CHECK_HIP(ctx, pos.alloc(ntot + 1));
    MUTO_CHECK_HIP(ctx, pscat.alloc(3 * ntot));
    MUTO_CHECK_HIP(ctx, hipMemcpyAsync(dr.p, rays, ntot * sizeof(muto_ray), hipMemcpyHostToDevice, s));
    k_select<<<grid_for(ntot, 256), 256, 0, s>>>(pl->g, dr.p, ntot, cfg->eps, flag.p, pscat.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
    MUTO_CHECK_HIP(ctx, exclusive_scan_u32_to_u64(flag.p, pos.p, ntot, s));
    uint64_t last_pos = 0;
    uint32_t last_flag = 0;
    MUTO_CHECK_HIP(ctx, copy_sync(ctx, &last_pos, pos.p + ntot - 1, sizeof last_pos, hipMemcpyDeviceToHost));
    MUTO_CHECK_HIP(ctx, copy_sync(ctx, &last_flag, flag.p + ntot - 1, sizeof last_flag, hipMemcpyDeviceToHost));
    const size_t nsel = (size_t)(last_pos + last_flag);
    pl->n_trk = nsel;
    pl->counters.n_scat_in_grid = (int64_t)nsel;
    MUTO_CHECK_HIP(ctx, sel.alloc(nsel));
    MUTO_CHECK_HIP(ctx, sel_p.alloc(3 * nsel));
    MUTO_CHECK_HIP(ctx, pl->trk4.alloc(4 * nsel));
    MUTO_CHECK_HIP(ctx, pl->pr.alloc(nsel));
    MUTO_CHECK_HIP(ctx, pl->geo.alloc(9 * nsel));
    k_compact<<<grid_for(ntot, 256), 256, 0, s>>>(pl->g, dr.p, ntot, flag.p, pos.p, pscat.p,
                                                  cfg->has_energy, sel.p, sel_p.p, pl->trk4.p, pl->pr.p,
                                                  pl->geo.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
    MUTO_CHECK_HIP(ctx, counts.alloc(nsel + 1));
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(counts.p, 0, (nsel + 1) * sizeof(uint64_t), s));
    k_paths_count<<<grid_for(nsel, 256), 256, 0, s>>>(pl->g, dr.p, sel.p, sel_p.p, nsel,
        cfg->straight_line, counts.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
    MUTO_CHECK_HIP(ctx, pl->row_ptr.alloc(nsel + 1));
    MUTO_CHECK_HIP(ctx, exclusive_scan_u64(counts.p, pl->row_ptr.p, nsel + 1, s));
    MUTO_CHECK_HIP(ctx, copy_sync(ctx, &pl->nnz, pl->row_ptr.p + nsel, sizeof(uint64_t), hipMemcpyDeviceToHost));
    if ((double)pl->nnz * 20.0 > (double)list_budget) {
      // over the budget: no lists; per-voxel counts from a walk, paths
      // recomputed by the MAP objective operator
      pl->has_lists = false;
      k_mj_walk<<<grid_for(nsel, 256), 256, 0, s>>>(pl->g, pl->geo.p, nsel, cfg->straight_line,
                                                    pl->mj.p, pl->nv, pl->dcount.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
      pl->mj_ready = true;
      MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
    } else {
    MUTO_CHECK_HIP(ctx, pl->vox.alloc(pl->nnz));
    MUTO_CHECK_HIP(ctx, pl->len.alloc(pl->nnz));
    MUTO_CHECK_HIP(ctx, pl->tail.alloc(pl->nnz));
    k_paths_fill<<<grid_for(nsel, PATH_BS), PATH_BS, 0, s>>>(pl->g, dr.p, sel.p, sel_p.p, nsel,
        cfg->straight_line, pl->row_ptr.p, pl->vox.p, pl->len.p, pl->dcount.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
    k_tails<<<grid_for(nsel, TAIL_BS), TAIL_BS, 0, s>>>(nsel, pl->row_ptr.p, pl->len.p, pl->tail.p,
                                                pl->dcount.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
    MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
    }
  }
  pl->counters.nnz = (int64_t)pl->nnz;
  if (int rc2 = muto_read_counters(pl)) { delete pl; return rc2; }
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(pl->dcount.p, 0, 8 * sizeof(unsigned long long), s));
  *out = pl;
  return MUTO_OK;
}

extern "C" int muto_em_plan_create_prepared(
    muto_ctx* ctx, const muto_mlsd_cfg* cfg, size_t n_trk, const double* trk4,
    const double* p_r, const uint64_t* row_ptr, const uint32_t* vox,
    const double* len, muto_em_plan** out) {
  if (!ctx) return MUTO_ERR_ARG;
  if (!out) { ctx->set_error("muto_em_plan_create_prepared: null out"); return MUTO_ERR_ARG; }
  *out = nullptr;
  if (int rc = validate_cfg(ctx, cfg)) return rc;
  if (!row_ptr || (n_trk && !trk4)) { ctx->set_error("muto_em_plan_create_prepared: null argument"); return MUTO_ERR_ARG; }
  const uint64_t nnz = row_ptr[n_trk];
  if (nnz && (!vox || !len)) { ctx->set_error("muto_em_plan_create_prepared: null vox/len"); return MUTO_ERR_ARG; }
  for (size_t i = 0; i < n_trk; ++i)
    if (row_ptr[i + 1] < row_ptr[i]) { ctx->set_error("muto_em_plan_create_prepared: row_ptr not monotone"); return MUTO_ERR_ARG; }
  auto* pl = new muto_em_plan();
  pl->ctx = ctx;
  pl->cfg = *cfg;
  pl->g = to_dev(cfg->grid);
  pl->nv = n_vox(cfg->grid);
  int rc = muto_plan_alloc_common(pl);
  if (rc) { delete pl; return rc; }
  hipStream_t s = ctx->stream;
  pl->n_trk = n_trk;
  pl->nnz = nnz;
  pl->counters.n_total = pl->counters.n_scat_in_grid = (int64_t)n_trk;
  pl->counters.nnz = (int64_t)nnz;
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(pl->dcount.p, 0, 8 * sizeof(unsigned long long), s));
  MUTO_CHECK_HIP(ctx, pl->mj.alloc(pl->nv));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(pl->mj.p, 0, pl->nv * sizeof(uint32_t), s));
  MUTO_CHECK_HIP(ctx, pl->trk4.alloc(4 * n_trk));
  MUTO_CHECK_HIP(ctx, pl->pr.alloc(n_trk));
  MUTO_CHECK_HIP(ctx, pl->row_ptr.alloc(n_trk + 1));
  MUTO_CHECK_HIP(ctx, pl->vox.alloc(nnz));
  MUTO_CHECK_HIP(ctx, pl->len.alloc(nnz));
  MUTO_CHECK_HIP(ctx, pl->tail.alloc(nnz));
  std::vector<double> ph(n_trk, 1.0);
  if (p_r) std::copy(p_r, p_r + n_trk, ph.begin());
  if (n_trk) {
    MUTO_CHECK_HIP(ctx, copy_sync(ctx, pl->trk4.p, trk4, 4 * n_trk * sizeof(double), hipMemcpyHostToDevice));
    MUTO_CHECK_HIP(ctx, copy_sync(ctx, pl->pr.p, ph.data(), n_trk * sizeof(double), hipMemcpyHostToDevice));
  }
  MUTO_CHECK_HIP(ctx, copy_sync(ctx, pl->row_ptr.p, row_ptr, (n_trk + 1) * sizeof(uint64_t), hipMemcpyHostToDevice));
  if (nnz) {
    MUTO_CHECK_HIP(ctx, copy_sync(ctx, pl->vox.p, vox, nnz * sizeof(uint32_t), hipMemcpyHostToDevice));
    MUTO_CHECK_HIP(ctx, copy_sync(ctx, pl->len.p, len, nnz * sizeof(double), hipMemcpyHostToDevice));
  }
  if (n_trk) {
    k_tails<<<grid_for(n_trk, TAIL_BS), TAIL_BS, 0, s>>>(n_trk, pl->row_ptr.p, pl->len.p, pl->tail.p,
                                                 pl->dcount.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
  }
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  if (int rc2 = muto_read_counters(pl)) { delete pl; return rc2; }
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(pl->dcount.p, 0, 8 * sizeof(unsigned long long), s));
  *out = pl;
  return MUTO_OK;
}

extern "C" int muto_em_plan_destroy(muto_em_plan* pl) {
  if (!pl) return MUTO_OK;
  if (pl->ctx) (void)hipStreamSynchronize(pl->ctx->stream);
  fp32c_release(pl);
  for (int k = 0; k < 8; ++k)
    if (pl->ev[k]) (void)hipEventDestroy(pl->ev[k]);
  delete pl;
  return MUTO_OK;
}

extern "C" int muto_em_plan_set_image_ex(muto_em_plan* pl, const double* img, int mem) {
  if (!pl || !img) return MUTO_ERR_ARG;
  muto_ctx* ctx = pl->ctx;
  if (mem != MUTO_MEM_HOST && mem != MUTO_MEM_DEVICE) {
    ctx->set_error("muto_em_plan_set_image_ex: mem must be MUTO_MEM_HOST or MUTO_MEM_DEVICE");
    return MUTO_ERR_ARG;
  }
  MUTO_CHECK_HIP(ctx, copy_sync(ctx, pl->lam64.p, img, pl->nv * sizeof(double),
                                mem == MUTO_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
  pl->cur_prec = MUTO_FP64;
  return MUTO_OK;
}

extern "C" int muto_em_plan_set_image(muto_em_plan* pl, const double* img) {
  return muto_em_plan_set_image_ex(pl, img, MUTO_MEM_HOST);
}

// lam64 is the image in both precisions (FP32C updates it in FP64 too);
// FP32C's forward pass reads a float mirror, lam32, refreshed here when
// FP32C takes over.  cur_prec: -1 no image, else the precision of the last
// run (MUTO_FP32C: the mirror is current).
static int sync_image(muto_em_plan* pl, int prec) {
  muto_ctx* ctx = pl->ctx;
  hipStream_t s = ctx->stream;
  if (pl->cur_prec < 0 || pl->cur_prec == prec) return MUTO_OK;
  if (prec == MUTO_FP32C) {
    if (!pl->lam32.p) MUTO_CHECK_HIP(ctx, pl->lam32.alloc(pl->nv));
    k_copy_f64_to_f32<<<grid_for(pl->nv, 256), 256, 0, s>>>(pl->nv, pl->lam64.p, pl->lam32.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
  }
  pl->cur_prec = prec;
  return MUTO_OK;
}

extern "C" int muto_em_plan_get_image_ex(muto_em_plan* pl, double* img, int mem) {
  if (!pl || !img) return MUTO_ERR_ARG;
  muto_ctx* ctx = pl->ctx;
  if (mem != MUTO_MEM_HOST && mem != MUTO_MEM_DEVICE) {
    ctx->set_error("muto_em_plan_get_image_ex: mem must be MUTO_MEM_HOST or MUTO_MEM_DEVICE");
    return MUTO_ERR_ARG;
  }
  MUTO_CHECK_HIP(ctx, copy_sync(ctx, img, pl->lam64.p, pl->nv * sizeof(double),
                                mem == MUTO_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost));
  return MUTO_OK;
}

extern "C" int muto_em_plan_get_image(muto_em_plan* pl, double* img) {
  return muto_em_plan_get_image_ex(pl, img, MUTO_MEM_HOST);
}

namespace muto {
int ensure_mj(muto_em_plan* pl) {
  if (pl->mj_ready) return MUTO_OK;
  muto_ctx* ctx = pl->ctx;
  if (pl->nnz) {
    k_mj_count<<<grid_for(pl->nnz, 256), 256, 0, ctx->stream>>>(pl->nnz, pl->vox.p, pl->nv, pl->mj.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
  }
  pl->mj_ready = true;
  return MUTO_OK;
}
}  // namespace muto

static int build_slots(muto_em_plan* pl) {
  if (pl->slots_built) return MUTO_OK;
  if (int rc = ensure_mj(pl)) return rc;
  muto_ctx* ctx = pl->ctx;
  hipStream_t s = ctx->stream;
  MUTO_CHECK_HIP(ctx, pl->col_ptr.alloc(pl->nv + 1));
  MUTO_CHECK_HIP(ctx, pl->slot.alloc(pl->nnz));
  MUTO_CHECK_HIP(ctx, pl->S.alloc(pl->nnz));
  std::vector<uint32_t> h(pl->nv);
  {
    DBuf<uint32_t> mj1, fill;
    MUTO_CHECK_HIP(ctx, mj1.alloc(pl->nv + 1));
    MUTO_CHECK_HIP(ctx, fill.alloc(pl->nv));
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(mj1.p, 0, (pl->nv + 1) * sizeof(uint32_t), s));
    MUTO_CHECK_HIP(ctx, hipMemcpyAsync(mj1.p, pl->mj.p, pl->nv * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    MUTO_CHECK_HIP(ctx, exclusive_scan_u32_to_u64(mj1.p, pl->col_ptr.p, pl->nv + 1, s));
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(fill.p, 0, pl->nv * sizeof(uint32_t), s));
    if (pl->nnz) {
      k_slots<<<grid_for(pl->nnz, 256), 256, 0, s>>>(pl->nnz, pl->vox.p, pl->nv, pl->col_ptr.p, fill.p, pl->slot.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    }
    MUTO_CHECK_HIP(ctx, hipMemcpyAsync(h.data(), pl->mj.p, pl->nv * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
    uint64_t mx = 0;
    for (uint32_t v : h) mx = std::max<uint64_t>(mx, v);
    pl->max_mj = mx;
  }
  if (pl->max_mj > SEL_CAP) {
    // the voxels too long for the LDS sort, each with its own scratch range
    std::vector<uint32_t> ids;
    std::vector<uint64_t> off;
    uint64_t tot = 0;
    for (size_t j = 0; j < pl->nv; ++j) {
      const uint32_t m = h[j];
      if (m <= SEL_CAP) continue;
      uint64_t n = 1;
      while (n < m) n <<= 1;
      ids.push_back((uint32_t)j);
      off.push_back(tot);
      tot += n;
    }
    pl->n_big = (uint32_t)ids.size();
    MUTO_CHECK_HIP(ctx, pl->big_vox.alloc(ids.size()));
    MUTO_CHECK_HIP(ctx, pl->big_off.alloc(off.size()));
    MUTO_CHECK_HIP(ctx, pl->sel_scratch.alloc(tot));
    MUTO_CHECK_HIP(ctx, copy_sync(ctx, pl->big_vox.p, ids.data(), ids.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    MUTO_CHECK_HIP(ctx, copy_sync(ctx, pl->big_off.p, off.data(), off.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
  }
  pl->slots_built = true;
  return MUTO_OK;
}

// One FP64 ML iteration: lam64 -> dst
static int iterate64(muto_em_plan* pl, double* dst, bool time_it) {
  muto_ctx* ctx = pl->ctx;
  hipStream_t s = ctx->stream;
  const muto_mlsd_cfg& c = pl->cfg;
  const bool mean_all = c.update_with_mean && c.mean_percentile == 100;
  if (!mean_all)
    if (int rc = build_slots(pl)) return rc;
  const ErrM E = error_matrix(c);
  if (time_it) MUTO_CHECK_HIP(ctx, hipEventRecord(pl->ev[0], s));
  const unsigned nblk = pl->fp64_kernel == 1 ? grid_for(pl->n_trk, 4) : grid_for(pl->n_trk, 256);
  if (pl->n_trk && pl->cost_part.n < nblk) MUTO_CHECK_HIP(ctx, pl->cost_part.alloc(nblk));
  if (pl->n_trk && pl->fp64_kernel == 1) {
    // 4 tracks per 256-thread block
    if (mean_all)
      k_em64w<0><<<nblk, 256, 0, s>>>(
          pl->n_trk, pl->row_ptr.p, pl->vox.p, pl->len.p, pl->tail.p, pl->trk4.p, pl->pr.p,
          pl->lam64.p, pl->nv, E, pl->acc64.p, nullptr, nullptr, pl->cost_part.p, pl->dcount.p);
    else
      k_em64w<1><<<nblk, 256, 0, s>>>(
          pl->n_trk, pl->row_ptr.p, pl->vox.p, pl->len.p, pl->tail.p, pl->trk4.p, pl->pr.p,
          pl->lam64.p, pl->nv, E, nullptr, pl->slot.p, pl->S.p, pl->cost_part.p, pl->dcount.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
  } else if (pl->n_trk) {
    if (mean_all)
      k_em64<0><<<grid_for(pl->n_trk, 256), 256, 0, s>>>(
          pl->n_trk, pl->row_ptr.p, pl->vox.p, pl->len.p, pl->tail.p, pl->trk4.p, pl->pr.p,
          pl->lam64.p, pl->nv, E, pl->acc64.p, nullptr, nullptr, pl->cost_part.p, pl->dcount.p);
    else
      k_em64<1><<<grid_for(pl->n_trk, 256), 256, 0, s>>>(
          pl->n_trk, pl->row_ptr.p, pl->vox.p, pl->len.p, pl->tail.p, pl->trk4.p, pl->pr.p,
          pl->lam64.p, pl->nv, E, nullptr, pl->slot.p, pl->S.p, pl->cost_part.p, pl->dcount.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
  }
  if (pl->n_trk) {
    k_sum_parts_f64<<<1, 256, 0, s>>>(pl->cost_part.p, nblk, pl->cost_acc.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
  }
  if (time_it) MUTO_CHECK_HIP(ctx, hipEventRecord(pl->ev[1], s));
  if (mean_all) {
    k_vox_mean64<<<grid_for(pl->nv, 256), 256, 0, s>>>(pl->nv, pl->mj.p, pl->acc64.p, dst);
  } else {
    const unsigned nb = (unsigned)std::min<size_t>(pl->nv, 65536);
    k_vox_select<<<nb, 256, 0, s>>>(pl->nv, pl->mj.p, pl->col_ptr.p, pl->S.p,
                                    c.update_with_mean, c.mean_percentile, dst);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
    if (pl->n_big) {
      MUTO_CHECK_HIP(ctx, hipGetLastError());
      k_vox_select_big<<<std::min<uint32_t>(pl->n_big, 1024), 1024, 0, s>>>(
          pl->n_big, pl->big_vox.p, pl->big_off.p, pl->mj.p, pl->col_ptr.p, pl->S.p,
          c.update_with_mean, c.mean_percentile, pl->sel_scratch.p, dst);
    }
  }
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  if (time_it) MUTO_CHECK_HIP(ctx, hipEventRecord(pl->ev[2], s));
  return MUTO_OK;
}

namespace muto {
int map_step(muto_em_plan* pl, const double* last, const double* ml, double* out) {
  muto_ctx* ctx = pl->ctx;
  const MapCfgD mc = map_cfg(pl);
  if (mc.all_nb)
    k_map_step<26><<<grid_for(pl->nv, 256), 256, 0, ctx->stream>>>(mc, pl->nv, pl->mj.p, last, ml, out,
                                                                   pl->dcount.p);
  else
    k_map_step<6><<<grid_for(pl->nv, 256), 256, 0, ctx->stream>>>(mc, pl->nv, pl->mj.p, last, ml, out,
                                                                  pl->dcount.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  return MUTO_OK;
}
}  // namespace muto

static int plan_run(muto_em_plan* pl, int iters, int precision, double* cost);

extern "C" int muto_em_plan_run(muto_em_plan* pl, int iters, int precision,
                                double* cost) {
  if (!pl) return MUTO_ERR_ARG;
  pl->is_map = false;
  return plan_run(pl, iters, precision, cost);
}

extern "C" int muto_map_plan_run(muto_em_plan* pl, int iters, int precision,
                                 double* cost) {
  if (!pl) return MUTO_ERR_ARG;
  pl->is_map = true;
  return plan_run(pl, iters, precision, cost);
}

static int plan_run(muto_em_plan* pl, int iters, int precision, double* cost) {
  muto_ctx* ctx = pl->ctx;
  if (iters < 0) { ctx->set_error("muto_em_plan_run: iters < 0"); return MUTO_ERR_ARG; }
  // MUTO_FP64_STREAM: FP64 on the brick stream for this run, whatever
  // MUTO_OPT_FP64_KERNEL says
  const bool stream64 = precision == MUTO_FP64_STREAM || (precision == MUTO_FP64 && pl->fp64_kernel == 2);
  if (precision == MUTO_FP64_STREAM) precision = MUTO_FP64;
  if (precision != MUTO_FP64 && precision != MUTO_FP32C) {
    ctx->set_error("muto_em_plan_run: precision must be MUTO_FP64, MUTO_FP32C or MUTO_FP64_STREAM");
    return MUTO_ERR_ARG;
  }
  const muto_mlsd_cfg& c = pl->cfg;
  if (pl->cur_prec < 0) { ctx->set_error("muto_em_plan_run: no image set"); return MUTO_ERR_STATE; }
  if (!pl->has_lists) {
    ctx->set_error("muto_em_plan_run: the plan keeps no path lists (they exceeded its list budget); "
                   "only muto_map_forward/backward run on it, recomputing the paths");
    return MUTO_ERR_STATE;
  }
  hipStream_t s = ctx->stream;
  if (pl->n_trk == 0) {
    // no selected track: every voxel has M_j = 0 and becomes 0; the mean
    // cost divides by zero as in the reference (MutoMLSD.cpp:202)
    if (iters > 0) {
      if (int rc = sync_image(pl, MUTO_FP64)) return rc;
      MUTO_CHECK_HIP(ctx, hipMemsetAsync(pl->lam64.p, 0, pl->nv * sizeof(double), s));
      MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
    }
    const int done = (c.early_stop && iters > 0) ? 1 : iters;
    if (cost) for (int k = 0; k < done; ++k) cost[k] = NAN;
    pl->counters.iterations_done = done;
    return MUTO_OK;
  }
  if (int rc = sync_image(pl, precision)) return rc;
  if (precision == MUTO_FP32C)
    if (int rc = fp32c_build(pl)) return rc;  // counts mj from the brick stream
  if (stream64)
    if (int rc = fp64s_build(pl)) return rc;
  if (int rc = ensure_mj(pl)) return rc;
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(pl->dcount.p, 0, 8 * sizeof(unsigned long long), s));
  std::vector<double> costs(iters, 0.0);
  double cost_prev = INFINITY;
  int done = 0;
  double tsum[5] = {0, 0, 0, 0, 0};
  const bool time_it = true;
  if (precision == MUTO_FP32C || stream64) {
    std::vector<double> sums(iters, 0.0);
    if (int rc = precision == MUTO_FP32C ? fp32c_run(pl, iters, sums.data(), &done)
                                         : fp64s_run(pl, iters, sums.data(), &done))
      return rc;
    for (int it = 0; it < done; ++it) {
      const double cst = sums[it] / (double)pl->n_trk;  // MutoMLSD.cpp:202
      if (c.log_results) printf("Iteration #%d finished.\n\tcost function = %f\n", it, cst);
      if (c.early_stop && it + 1 == done && it > 0 && costs[it - 1] < cst && c.log_results)
        printf("Early stop at iteration#%d. Minimum cost function = %f\n", it, costs[it - 1]);
      costs[it] = cst;
    }
    if (cost) for (int k = 0; k < done; ++k) cost[k] = costs[k];
    pl->counters.iterations_done = done;
    return muto_read_counters(pl);
  }
  for (int it = 0; it < iters; ++it) {
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(pl->cost_acc.p, 0, sizeof(double), s));
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(pl->dcount.p, 0, 3 * sizeof(unsigned long long), s));
    if (precision == MUTO_FP64) {
      if (!pl->tmp64.p) MUTO_CHECK_HIP(ctx, pl->tmp64.alloc(pl->nv));
      if (pl->is_map) {
        if (int rc = iterate64(pl, pl->ml64.p, time_it)) return rc;
        if (int rc = map_step(pl, pl->lam64.p, pl->ml64.p, pl->tmp64.p)) return rc;
      } else {
        if (int rc = iterate64(pl, pl->tmp64.p, time_it)) return rc;
      }
      std::swap(pl->lam64.p, pl->tmp64.p);
      if (time_it) {
        MUTO_CHECK_HIP(ctx, hipEventRecord(pl->ev[3], s));
        MUTO_CHECK_HIP(ctx, hipEventSynchronize(pl->ev[3]));
        float a = 0, d = 0;
        MUTO_CHECK_HIP(ctx, hipEventElapsedTime(&a, pl->ev[0], pl->ev[1]));
        MUTO_CHECK_HIP(ctx, hipEventElapsedTime(&d, pl->ev[0], pl->ev[3]));
        tsum[2] += a;      // fused covariance + S values kernel
        tsum[3] += d - a;  // voxel update (+ MAP step)
        tsum[4] += d;
      }
    }
    double csum = 0.0;
    MUTO_CHECK_HIP(ctx, hipMemcpyAsync(&csum, pl->cost_acc.p, sizeof(double), hipMemcpyDeviceToHost, s));
    MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
    const double cst = csum / (double)pl->n_trk;  // MutoMLSD.cpp:202
    costs[it] = cst;
    done = it + 1;
    if (c.log_results) printf("Iteration #%d finished.\n\tcost function = %f\n", it, cst);
    if (c.early_stop && cost_prev < cst) {
      if (c.log_results) printf("Early stop at iteration#%d. Minimum cost function = %f\n", it, cost_prev);
      break;
    }
    cost_prev = cst;
  }
  if (cost) for (int k = 0; k < done; ++k) cost[k] = costs[k];
  for (int k = 0; k < 5; ++k) pl->t_ms[k] = done ? tsum[k] / done : 0.0;
  pl->counters.iterations_done = done;
  if (int rc = muto_read_counters(pl)) return rc;
  return MUTO_OK;
}

extern "C" int muto_em_plan_set_option(muto_em_plan* pl, int option, int value) {
  if (!pl) return MUTO_ERR_ARG;
  switch (option) {
    case MUTO_OPT_GRAPH:
      pl->use_graph = value != 0;
      return MUTO_OK;
    case MUTO_OPT_TIMING:
      pl->timing = value != 0;
      return MUTO_OK;
    case MUTO_OPT_BRICK: {
      const BrickShape b = BrickShape::from_code((uint32_t)value);
      if (b.bx == 0 || b.by == 0 || b.bz == 0 || b.vol() > 16384 || b.block == 0) break;
      if (b != pl->f32.bs) {
        pl->f32.reset();
        pl->f32.bs = b;
      }
      return MUTO_OK;
    }
    case MUTO_OPT_FP64_KERNEL:
      if (value < 0 || value > 2) break;
      pl->fp64_kernel = value;
      return MUTO_OK;
    case MUTO_OPT_MEDIAN_BRACKETS:
      if (value != 0 && value != 1) break;
      if (pl->f32.med_brk != (value != 0) && pl->f32.graph) {
        (void)hipGraphExecDestroy(pl->f32.graph);
        pl->f32.graph = nullptr;
      }
      pl->f32.med_brk = value != 0;
      return MUTO_OK;
    case MUTO_OPT_ACC_MODE:
      if (value != 0 && value != 1) break;
      if (pl->f32.acc_mode != value && pl->f32.graph) {
        (void)hipGraphExecDestroy(pl->f32.graph);
        pl->f32.graph = nullptr;
      }
      pl->f32.acc_mode = value;
      pl->f32.acc_mode_set = true;
      return MUTO_OK;
    default:
      break;
  }
  pl->ctx->set_error("muto_em_plan_set_option: unknown option or value");
  return MUTO_ERR_ARG;
}

extern "C" int muto_em_plan_autotune(muto_em_plan* pl, int iters, int n_cand,
                                     const uint32_t* shapes, double* ms, uint32_t* best) {
  if (!pl) return MUTO_ERR_ARG;
  muto_ctx* ctx = pl->ctx;
  // 256-thread blocks for bricks of up to 4096 voxels, 512 for 8192, 1024
  // for 16384 (the larger LDS tile with the same waves per CU)
  static const uint32_t kDefault[] = {
      16u | 16u << 8 | 16u << 16,           16u | 8u << 8 | 32u << 16,
      8u | 16u << 8 | 32u << 16,            8u | 8u << 8 | 64u << 16,
      16u | 16u << 8 | 32u << 16 | 1u << 24, 32u | 16u << 8 | 16u << 16 | 1u << 24,
      16u | 32u << 8 | 16u << 16 | 1u << 24, 16u | 16u << 8 | 64u << 16 | 2u << 24,
      32u | 16u << 8 | 32u << 16 | 2u << 24, 16u | 16u << 8 | 32u << 16};
  if (!shapes || n_cand <= 0) {
    shapes = kDefault;
    n_cand = (int)(sizeof kDefault / sizeof kDefault[0]);
  }
  if (pl->cur_prec < 0) { ctx->set_error("muto_em_plan_autotune: no image set"); return MUTO_ERR_STATE; }
  if (!pl->has_lists || pl->n_trk == 0) { ctx->set_error("muto_em_plan_autotune: nothing to tune"); return MUTO_ERR_STATE; }
  std::vector<double> t(n_cand);
  pl->is_map = false;
  if (int rc = fp32c_autotune(pl, iters, n_cand, shapes, t.data())) return rc;
  pl->cur_prec = MUTO_FP64;
  if (ms) std::copy(t.begin(), t.end(), ms);
  if (best) *best = pl->f32.bs.code();
  return MUTO_OK;
}

extern "C" int muto_em_plan_counters(const muto_em_plan* pl, muto_counters* out) {
  if (!pl || !out) return MUTO_ERR_ARG;
  *out = pl->counters;
  out->n_segments = (int64_t)pl->f32.n_seg;
  out->n_map_tile_records = (int64_t)pl->mt_n_rec;
  return MUTO_OK;
}

extern "C" int muto_em_plan_voxel_counts(muto_em_plan* pl, uint32_t* mj) {
  if (!pl || !mj) return MUTO_ERR_ARG;
  if (int rc = ensure_mj(pl)) return rc;
  MUTO_CHECK_HIP(pl->ctx, copy_sync(pl->ctx, mj, pl->mj.p, pl->nv * sizeof(uint32_t), hipMemcpyDeviceToHost));
  return MUTO_OK;
}

extern "C" int muto_em_plan_export(muto_em_plan* pl, uint64_t* row_ptr, uint32_t* vox,
                                   double* len, double* tail, double* trk4, double* p_r) {
  if (!pl) return MUTO_ERR_ARG;
  muto_ctx* ctx = pl->ctx;
  const size_t nt = pl->n_trk;
  const uint64_t nnz = pl->nnz;
  if (!pl->has_lists && (vox || len || tail)) {
    ctx->set_error("muto_em_plan_export: the plan keeps no path lists");
    return MUTO_ERR_STATE;
  }
  if (row_ptr) MUTO_CHECK_HIP(ctx, copy_sync(ctx, row_ptr, pl->row_ptr.p, (nt + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost));
  if (nnz && vox) MUTO_CHECK_HIP(ctx, copy_sync(ctx, vox, pl->vox.p, nnz * sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (nnz && len) MUTO_CHECK_HIP(ctx, copy_sync(ctx, len, pl->len.p, nnz * sizeof(double), hipMemcpyDeviceToHost));
  if (nnz && tail) MUTO_CHECK_HIP(ctx, copy_sync(ctx, tail, pl->tail.p, nnz * sizeof(double), hipMemcpyDeviceToHost));
  if (nt && trk4) MUTO_CHECK_HIP(ctx, copy_sync(ctx, trk4, pl->trk4.p, 4 * nt * sizeof(double), hipMemcpyDeviceToHost));
  if (nt && p_r) MUTO_CHECK_HIP(ctx, copy_sync(ctx, p_r, pl->pr.p, nt * sizeof(double), hipMemcpyDeviceToHost));
  return MUTO_OK;
}

extern "C" int muto_em_plan_timing(const muto_em_plan* pl, double* t_ms) {
  if (!pl || !t_ms) return MUTO_ERR_ARG;
  for (int k = 0; k < 5; ++k) t_ms[k] = pl->t_ms[k];
  return MUTO_OK;
}

// ------------------------------------------------------------------ //
// one-shot MLSD / MAP                                                 //
// ------------------------------------------------------------------ //
// initailizeImage (MutoMLSD.cpp:226-266) for methods 1/2 and a missing
// file.  Method 1 uses std::random_device in the reference (not
// reproducible); here a fixed-seed LCG + Box-Muller draws N(2, 0.5)*1e-7
// floored at 1e-10.
static void init_image_host(const muto_mlsd_cfg* cfg, const double* init, double* img, size_t nv) {
  if (cfg->initialize_method == 0 && init) { memcpy(img, init, nv * sizeof(double)); return; }
  if (cfg->initialize_method == 1 || (cfg->initialize_method == 0 && !init)) {
    uint64_t s = 0x9E3779B97F4A7C15ull;
    for (size_t j = 0; j < nv; ++j) {
      s = s * 6364136223846793005ull + 1442695040888963407ull;
      double u1 = ((s >> 11) + 0.5) * (1.0 / 9007199254740992.0);
      s = s * 6364136223846793005ull + 1442695040888963407ull;
      double u2 = ((s >> 11) + 0.5) * (1.0 / 9007199254740992.0);
      double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
      double v = (2.0 + 0.5 * z) * 1.0e-7;
      img[j] = 1.0e-10 <= v ? v : 1.0e-10;
    }
    return;
  }
  for (size_t j = 0; j < nv; ++j) img[j] = cfg->initialize_value;
}

static int one_shot(muto_ctx* ctx, const muto_mlsd_cfg* cfg, size_t n, const muto_ray* rays,
                    const double* init_img, double* img_out, double* cost_per_iter,
                    muto_counters* counters, bool is_map) {
  if (!ctx) return MUTO_ERR_ARG;
  if (!img_out) { ctx->set_error("null img_out"); return MUTO_ERR_ARG; }
  muto_em_plan* pl = nullptr;
  int rc = muto_em_plan_create(ctx, cfg, n, rays, &pl);
  if (rc) return rc;
  std::vector<double> img(pl->nv);
  init_image_host(cfg, init_img, img.data(), pl->nv);
  if (cfg->initialize_method == 0 && !init_img && cfg->log_results)
    printf("Warning: no initial image; using random initialisation\n");
  rc = muto_em_plan_set_image(pl, img.data());
  if (!rc)
    rc = is_map ? muto_map_plan_run(pl, cfg->num_iteration, cfg->precision, cost_per_iter)
                : muto_em_plan_run(pl, cfg->num_iteration, cfg->precision, cost_per_iter);
  if (!rc) rc = muto_em_plan_get_image(pl, img_out);
  if (!rc && counters) muto_em_plan_counters(pl, counters);
  muto_em_plan_destroy(pl);
  return rc;
}

extern "C" int muto_mlsd_reconstruct(muto_ctx* ctx, const muto_mlsd_cfg* cfg, size_t n,
                                     const muto_ray* rays, const double* init_img,
                                     double* img_out, double* cost_per_iter,
                                     muto_counters* counters) {
  return one_shot(ctx, cfg, n, rays, init_img, img_out, cost_per_iter, counters, false);
}

extern "C" int muto_map_reconstruct(muto_ctx* ctx, const muto_mlsd_cfg* cfg, size_t n,
                                    const muto_ray* rays, const double* init_img,
                                    double* img_out, double* cost_per_iter,
                                    muto_counters* counters) {
  return one_shot(ctx, cfg, n, rays, init_img, img_out, cost_per_iter, counters, true);
}
