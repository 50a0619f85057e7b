// FP32C EM iteration: brick-major segment stream, LDS-staged gathers and
// backprojection, FP32 sums with two-sum compensation, FP64 2x2 algebra.
//
// Layout (built once per plan from the track-major FP64 lists):
//   - the grid is cut into bricks (LDS tiles) of bx*by*bz voxels;
//   - each track's merged path is cut into segments: maximal runs of
//     consecutive entries inside one brick;
//   - segments are sorted by (brick, track rank), tracks ranked by a
//     spatial key (first brick, last brick) so that neighbouring lanes
//     handle neighbouring, nearly parallel tracks;
//   - entries are stored brick-major as (u16 voxel-in-brick, f32 length):
//     6 bytes per track-voxel entry.
// Per iteration (src/MutoMLSD.cpp:285-308, mean update):
//   k_fwd32   per segment: partial covariance sums sum lambda_j W_ij with
//             lambda staged in LDS, W from (L, tail) in registers;
//   k_track32 per track: Sigma_i in FP64 from its segment partials, Eigen's
//             2x2 det/inverse in FP64, cost_i, G_i = p^2 (Si^-1 D Si^-1/2
//             - Si^-1) in FP32;
//   k_bwd32   per segment: c_ij = <W_ij, G_i> accumulated per voxel in LDS
//             with atomic two-sum (hi, lo), flushed once per task;
//   k_upd32   per voxel: lambda += lambda^2/(2 M_j) sum_i c_ij, which is
//             half the mean of S_ij = 2 lambda_j + c_ij lambda_j^2.
// The persistent kernels take tasks (brick, segment range) from a queue
// sorted by size, largest first, through an atomic counter.
#include <math.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "em_plan.hpp"

namespace muto {

enum { C_SINGULAR = 0 };
struct ErrM {
  double e00, e01, e11;
};
ErrM error_matrix(const muto_mlsd_cfg& c);
int map_step(muto_em_plan* pl, const double* last, const double* ml, double* out);

namespace {

struct BrickGeo {
  uint32_t n[3];   // grid
  uint32_t b[3];   // brick shape
  uint32_t nb[3];  // bricks per axis
};

__device__ __forceinline__ uint32_t brick_of(const BrickGeo& G, uint32_t j,
                                             uint32_t* local) {
  const uint32_t x = j % G.n[0];
  const uint32_t t = j / G.n[0];
  const uint32_t y = t % G.n[1];
  const uint32_t z = t / G.n[1];
  const uint32_t bx = x / G.b[0], by = y / G.b[1], bz = z / G.b[2];
  if (local)
    *local = (x - bx * G.b[0]) + G.b[0] * ((y - by * G.b[1]) + G.b[1] * (z - bz * G.b[2]));
  return bx + G.nb[0] * (by + G.nb[1] * bz);
}

// ---------------- build ------------------------------------------------
__global__ void __launch_bounds__(256)
k_seg_count(BrickGeo G, size_t n_trk, size_t nv,
            const uint64_t* __restrict__ row_ptr,
            const uint32_t* __restrict__ vox, uint32_t* __restrict__ nseg,
            uint64_t* __restrict__ key, uint32_t* __restrict__ val) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_trk) return;
  const uint64_t b = row_ptr[t], e = row_ptr[t + 1];
  uint32_t cnt = 0, prev = 0xFFFFFFFFu, first = 0, last = 0;
  bool have = false;
  for (uint64_t k = b; k < e; ++k) {
    const uint32_t j = vox[k];
    if (j >= nv) { prev = 0xFFFFFFFFu; continue; }
    const uint32_t br = brick_of(G, j, nullptr);
    if (br != prev) ++cnt;
    if (!have) { first = br; have = true; }
    last = br;
    prev = br;
  }
  nseg[t] = cnt;
  const uint64_t nbk = (uint64_t)G.nb[0] * G.nb[1] * G.nb[2];
  key[t] = have ? (uint64_t)first * nbk + last : ~0ull;
  val[t] = (uint32_t)t;
}

__global__ void k_invert(size_t n, const uint32_t* __restrict__ order,
                         uint32_t* __restrict__ rank) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) rank[order[r]] = (uint32_t)r;
}

__global__ void __launch_bounds__(256)
k_seg_fill(BrickGeo G, size_t n_trk, size_t nv,
           const uint64_t* __restrict__ row_ptr,
           const uint32_t* __restrict__ vox, const double* __restrict__ tail,
           const uint64_t* __restrict__ sstart,
           const uint32_t* __restrict__ rank, uint64_t* __restrict__ skey,
           uint32_t* __restrict__ sval, uint64_t* __restrict__ st_start,
           uint32_t* __restrict__ st_len, float* __restrict__ st_tend,
           uint32_t* __restrict__ st_trk) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_trk) return;
  const uint64_t b = row_ptr[t], e = row_ptr[t + 1];
  uint64_t s = sstart[t];
  uint32_t prev = 0xFFFFFFFFu;
  const uint32_t r = rank[t];
  uint64_t cur = ~0ull;  // current segment id
  for (uint64_t k = b; k < e; ++k) {
    const uint32_t j = vox[k];
    if (j >= nv) { prev = 0xFFFFFFFFu; continue; }
    const uint32_t br = brick_of(G, j, nullptr);
    if (br != prev) {
      cur = s++;
      skey[cur] = ((uint64_t)br << 32) | r;
      sval[cur] = (uint32_t)cur;
      st_start[cur] = k;
      st_len[cur] = 0;
      st_trk[cur] = (uint32_t)t;
    }
    st_len[cur] += 1;
    st_tend[cur] = (float)tail[k];
    prev = br;
  }
}

__global__ void __launch_bounds__(256)
k_seg_gather(uint64_t n_seg, const uint32_t* __restrict__ sval,
             const uint32_t* __restrict__ st_len,
             const float* __restrict__ st_tend,
             const uint32_t* __restrict__ st_trk,
             const uint32_t* __restrict__ rank, const uint64_t* __restrict__ skey,
             uint32_t* __restrict__ seg_trk, float* __restrict__ seg_tend,
             uint32_t* __restrict__ lenq, uint32_t* __restrict__ bcount) {
  uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_seg) return;
  const uint32_t s = sval[q];
  seg_trk[q] = rank[st_trk[s]];
  seg_tend[q] = st_tend[s];
  lenq[q] = st_len[s];
  atomicAdd(&bcount[(uint32_t)(skey[q] >> 32)], 1u);
}

__global__ void __launch_bounds__(256)
k_ent_scatter(BrickGeo G, uint64_t n_seg, const uint32_t* __restrict__ sval,
              const uint64_t* __restrict__ st_start,
              const uint32_t* __restrict__ st_len,
              const uint64_t* __restrict__ seg_off,
              const uint32_t* __restrict__ vox, const double* __restrict__ len,
              uint16_t* __restrict__ ent_vox, float* __restrict__ ent_len) {
  uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_seg) return;
  const uint32_t s = sval[q];
  const uint64_t src = st_start[s];
  const uint32_t n = st_len[s];
  uint64_t dst = seg_off[q];
  // entries with j >= nv end segments, so a segment is contiguous in the
  // track-major list
  for (uint32_t k = 0; k < n; ++k) {
    uint32_t loc;
    brick_of(G, vox[src + k], &loc);
    ent_vox[dst + k] = (uint16_t)loc;
    ent_len[dst + k] = (float)len[src + k];
  }
}

__global__ void __launch_bounds__(256)
k_tseg_fill(uint64_t n_seg, const uint32_t* __restrict__ sval,
            const uint32_t* __restrict__ st_trk,
            const uint64_t* __restrict__ sstart,
            const uint32_t* __restrict__ rank,
            const uint32_t* __restrict__ tseg_ptr,
            uint32_t* __restrict__ tseg_pos) {
  uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_seg) return;
  const uint32_t s = sval[q];
  const uint32_t t = st_trk[s];
  tseg_pos[tseg_ptr[rank[t]] + (uint32_t)(s - sstart[t])] = (uint32_t)q;
}

__global__ void k_trk_d(size_t n_trk, const uint32_t* __restrict__ order,
                        const double* __restrict__ trk4,
                        const double* __restrict__ pr,
                        const uint32_t* __restrict__ nseg,
                        double* __restrict__ trk_d,
                        uint32_t* __restrict__ nseg_rank) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_trk) return;
  const uint32_t t = order[r];
  const double Sx = trk4[4 * t], Sy = trk4[4 * t + 1];
  const double Dx = trk4[4 * t + 2], Dy = trk4[4 * t + 3];
  trk_d[4 * r] = Sx * Sx + Sy * Sy;
  trk_d[4 * r + 1] = Sx * Dx + Sy * Dy;
  trk_d[4 * r + 2] = Dx * Dx + Dy * Dy;
  trk_d[4 * r + 3] = pr[t];
  nseg_rank[r] = nseg[t];
}

__global__ void k_brick_entries(uint64_t n_seg, const uint64_t* __restrict__ skey,
                                const uint64_t* __restrict__ seg_off,
                                unsigned long long* __restrict__ bent) {
  uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_seg) return;
  atomicAdd(&bent[(uint32_t)(skey[q] >> 32)],
            (unsigned long long)(seg_off[q + 1] - seg_off[q]));
}

// ---------------- iteration --------------------------------------------
__device__ __forceinline__ void two_sum_acc(float& hi, float& lo, float x) {
  const float s = hi + x;
  const float bp = s - hi;
  const float e = (hi - (s - bp)) + (x - bp);
  hi = s;
  lo += e;
}


struct IterArgs {
  BrickGeo G;
  const Task* tasks;
  uint32_t n_tasks;
  unsigned int* work;
  uint64_t n_seg;
  const uint64_t* seg_off;
  const float* seg_tend;
  const uint32_t* seg_trk;
  const uint16_t* ent_vox;
  const float* ent_len;
};

__device__ __forceinline__ void brick_origin(const BrickGeo& G, uint32_t brick,
                                             uint32_t* o) {
  const uint32_t bx = brick % G.nb[0];
  const uint32_t t = brick / G.nb[0];
  o[0] = bx * G.b[0];
  o[1] = (t % G.nb[1]) * G.b[1];
  o[2] = (t / G.nb[1]) * G.b[2];
}

// lambda of the brick into LDS (zero outside the grid)
__device__ __forceinline__ void load_brick(const BrickGeo& G, uint32_t brick,
                                           const float* __restrict__ lam,
                                           float* lds) {
  uint32_t o[3];
  brick_origin(G, brick, o);
  const uint32_t vol = G.b[0] * G.b[1] * G.b[2];
  for (uint32_t v = threadIdx.x; v < vol; v += blockDim.x) {
    const uint32_t lx = v % G.b[0], r = v / G.b[0];
    const uint32_t ly = r % G.b[1], lz = r / G.b[1];
    const uint32_t x = o[0] + lx, y = o[1] + ly, z = o[2] + lz;
    float val = 0.0f;
    if (x < G.n[0] && y < G.n[1] && z < G.n[2])
      val = lam[(size_t)x + (size_t)G.n[0] * ((size_t)y + (size_t)G.n[1] * z)];
    lds[v] = val;
  }
}

// Forward: per segment, sum_j lambda_j W_ij (three components), walking
// the segment backwards so the tail T grows by each entry's length.
__global__ void __launch_bounds__(256)
k_fwd32(IterArgs a, const float* __restrict__ lam, float* __restrict__ part) {
  extern __shared__ float lds_lam[];
  __shared__ uint32_t task_s;
  for (;;) {
    if (threadIdx.x == 0) task_s = atomicAdd(&a.work[0], 1u);
    __syncthreads();
    const uint32_t ti = task_s;
    if (ti >= a.n_tasks) break;
    const Task tk = a.tasks[ti];
    load_brick(a.G, tk.brick, lam, lds_lam);
    __syncthreads();
    for (uint64_t q = tk.s0 + threadIdx.x; q < tk.s1; q += blockDim.x) {
      const uint64_t o0 = a.seg_off[q], o1 = a.seg_off[q + 1];
      float T = a.seg_tend[q];
      float h0 = 0.f, l0 = 0.f, h1 = 0.f, l1 = 0.f, h2 = 0.f, l2 = 0.f;
      for (uint64_t k = o1; k > o0; --k) {
        const float L = a.ent_len[k - 1];
        const float lj = lds_lam[a.ent_vox[k - 1]];
        const float LT = L * T;
        const float w01 = LT + L * L * 0.5f;
        const float w11 = L * L * L * (1.0f / 3.0f) + (L + T) * LT;
        two_sum_acc(h0, l0, lj * L);
        two_sum_acc(h1, l1, lj * w01);
        two_sum_acc(h2, l2, lj * w11);
        T += L;
      }
      part[q] = h0 + l0;
      part[a.n_seg + q] = h1 + l1;
      part[2 * a.n_seg + q] = h2 + l2;
    }
    __syncthreads();
  }
}

// Per track (rank order): Sigma from segment partials, 2x2 algebra in FP64.
__global__ void __launch_bounds__(256)
k_track32(size_t n_trk, const uint32_t* __restrict__ tseg_ptr,
          const uint32_t* __restrict__ tseg_pos, uint64_t n_seg,
          const float* __restrict__ part, const double* __restrict__ trk_d,
          ErrM E, float* __restrict__ trk_g, double* __restrict__ cost_acc,
          unsigned long long* __restrict__ cnt) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double cost = 0.0;
  bool sing = false;
  if (r < n_trk) {
    double s00 = 0.0, s01 = 0.0, s11 = 0.0;
    for (uint32_t k = tseg_ptr[r]; k < tseg_ptr[r + 1]; ++k) {
      const uint32_t q = tseg_pos[k];
      s00 += (double)part[q];
      s01 += (double)part[n_seg + q];
      s11 += (double)part[2 * n_seg + q];
    }
    const double D00 = trk_d[4 * r], D01 = trk_d[4 * r + 1];
    const double D11 = trk_d[4 * r + 2], p = trk_d[4 * r + 3];
    const double pp = p * p;
    const double m00 = pp * s00 + E.e00, m01 = pp * s01 + E.e01;
    const double m11 = pp * s11 + E.e11;
    const double det = m00 * m11 - m01 * m01;
    double i00 = 0, i01 = 0, i11 = 0;
    const bool ok = fabs(det) > 1e-12;
    if (ok) {
      const double invdet = 1.0 / det;
      i00 = m11 * invdet;
      i01 = -m01 * invdet;
      i11 = m00 * invdet;
    }
    if (!ok || det <= 0) {
      sing = true;
    } else {
      cost = log(det) + (D00 * i00 + D11 * i11 + D01 * 2.0 * i01) * 0.5;
    }
    // G = 1/2 Si^-1 D Si^-1 - Si^-1  (symmetric)
    const double A00 = i00 * D00 + i01 * D01, A01 = i00 * D01 + i01 * D11;
    const double A10 = i01 * D00 + i11 * D01, A11 = i01 * D01 + i11 * D11;
    const double M00 = A00 * i00 + A01 * i01;
    const double M01 = A00 * i01 + A01 * i11;
    const double M11 = A10 * i01 + A11 * i11;
    trk_g[4 * r] = (float)(pp * (0.5 * M00 - i00));
    trk_g[4 * r + 1] = (float)(pp * (0.5 * M01 - i01));
    trk_g[4 * r + 2] = (float)(pp * (0.5 * M11 - i11));
    trk_g[4 * r + 3] = 0.0f;
  }
  double v = cost;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0 && v != 0.0) atomicAdd(cost_acc, v);
  unsigned long long m = __ballot(sing);
  if (m && (threadIdx.x & 63) == (unsigned)(__ffsll((long long)m) - 1))
    atomicAdd(&cnt[C_SINGULAR], (unsigned long long)__popcll(m));
}

// Backprojection: c_ij = <W_ij, G_i> accumulated per voxel in LDS with an
// atomic two-sum: the returned old value o and the stored o + c give the
// exact rounding error, which goes into the low word.
__global__ void __launch_bounds__(256)
k_bwd32(IterArgs a, const float* __restrict__ trk_g,
        double* __restrict__ acc) {
  extern __shared__ float lds_acc[];
  __shared__ uint32_t task_s;
  const uint32_t vol = a.G.b[0] * a.G.b[1] * a.G.b[2];
  float* hi = lds_acc;
  float* lo = lds_acc + vol;
  for (;;) {
    if (threadIdx.x == 0) task_s = atomicAdd(&a.work[1], 1u);
    for (uint32_t v = threadIdx.x; v < 2 * vol; v += blockDim.x) lds_acc[v] = 0.0f;
    __syncthreads();
    const uint32_t ti = task_s;
    if (ti >= a.n_tasks) break;
    const Task tk = a.tasks[ti];
    for (uint64_t q = tk.s0 + threadIdx.x; q < tk.s1; q += blockDim.x) {
      const uint64_t o0 = a.seg_off[q], o1 = a.seg_off[q + 1];
      const uint32_t r = a.seg_trk[q];
      const float4 g = *reinterpret_cast<const float4*>(trk_g + 4 * (size_t)r);
      float T = a.seg_tend[q];
      for (uint64_t k = o1; k > o0; --k) {
        const float L = a.ent_len[k - 1];
        const uint32_t v = a.ent_vox[k - 1];
        const float LT = L * T;
        const float w01 = LT + L * L * 0.5f;
        const float w11 = L * L * L * (1.0f / 3.0f) + (L + T) * LT;
        const float c = L * g.x + 2.0f * w01 * g.y + w11 * g.z;
        const float old = atomicAdd(&hi[v], c);
        const float s = old + c;
        const float bp = s - old;
        const float e = (old - (s - bp)) + (c - bp);
        if (e != 0.0f) atomicAdd(&lo[v], e);
        T += L;
      }
    }
    __syncthreads();
    uint32_t o[3];
    brick_origin(a.G, tk.brick, o);
    for (uint32_t v = threadIdx.x; v < vol; v += blockDim.x) {
      const uint32_t lx = v % a.G.b[0], rr = v / a.G.b[0];
      const uint32_t ly = rr % a.G.b[1], lz = rr / a.G.b[1];
      const uint32_t x = o[0] + lx, y = o[1] + ly, z = o[2] + lz;
      if (x >= a.G.n[0] || y >= a.G.n[1] || z >= a.G.n[2]) continue;
      const size_t j = (size_t)x + (size_t)a.G.n[0] * ((size_t)y + (size_t)a.G.n[1] * z);
      const double val = (double)hi[v] + (double)lo[v];
      if (tk.owner)
        acc[j] = val;
      else if (val != 0.0)
        atomicAdd(&acc[j], val);
    }
    __syncthreads();
  }
}

// lambda_j <- lambda_j + lambda_j^2 / (2 M_j) * sum_i c_ij; clears the
// accumulator and the work counters for the next iteration.
__global__ void __launch_bounds__(256)
k_upd32(size_t nv, const uint32_t* __restrict__ mj, double* __restrict__ acc,
        float* __restrict__ lam, double* __restrict__ ml,
        double* __restrict__ last, unsigned int* __restrict__ work) {
  size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j == 0) { work[0] = 0; work[1] = 0; }
  if (j >= nv) return;
  const uint32_t m = mj[j];
  const double l = (double)lam[j];
  const double out = m == 0 ? 0.0 : l + l * l / (2.0 * (double)m) * acc[j];
  acc[j] = 0.0;
  if (ml) {  // MAP: keep lambda_last (FP64) and lambda_ML for the voxel step
    ml[j] = out;
    last[j] = l;
  } else {
    lam[j] = (float)out;
  }
}

__global__ void k_f64_to_f32(size_t n, const double* __restrict__ a,
                             float* __restrict__ b) {
  size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) b[j] = (float)a[j];
}

}  // namespace

// ---------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------
static BrickGeo brick_geo(const muto_em_plan* pl) {
  BrickGeo G;
  for (int k = 0; k < 3; ++k) G.n[k] = pl->g.n[k];
  G.b[0] = pl->f32.bs.bx;
  G.b[1] = pl->f32.bs.by;
  G.b[2] = pl->f32.bs.bz;
  for (int k = 0; k < 3; ++k) G.nb[k] = (G.n[k] + G.b[k] - 1) / G.b[k];
  return G;
}

static IterArgs iter_args(muto_em_plan* pl) {
  Fp32cData& f = pl->f32;
  IterArgs a;
  a.G = brick_geo(pl);
  a.tasks = f.tasks.p;
  a.n_tasks = f.n_tasks;
  a.work = pl->work.p;
  a.n_seg = f.n_seg;
  a.seg_off = f.seg_off.p;
  a.seg_tend = f.seg_tend.p;
  a.seg_trk = f.seg_trk.p;
  a.ent_vox = f.ent_vox.p;
  a.ent_len = f.ent_len.p;
  return a;
}

int fp32c_release(muto_em_plan* pl) {
  Fp32cData& f = pl->f32;
  if (f.graph) (void)hipGraphExecDestroy(f.graph);
  f.graph = nullptr;
  return MUTO_OK;
}

static int bits_for(uint64_t v) {
  int b = 0;
  while ((1ull << b) <= v && b < 63) ++b;
  return b;
}

int fp32c_build(muto_em_plan* pl) {
  Fp32cData& f = pl->f32;
  if (f.built) return MUTO_OK;
  muto_ctx* ctx = pl->ctx;
  hipStream_t s = ctx->stream;
  const BrickGeo G = brick_geo(pl);
  f.nb[0] = G.nb[0]; f.nb[1] = G.nb[1]; f.nb[2] = G.nb[2];
  const uint64_t nbk = (uint64_t)G.nb[0] * G.nb[1] * G.nb[2];
  if (f.bs.vol() > 32768) { ctx->set_error("brick volume above 32768 voxels"); return MUTO_ERR_LIMIT; }
  const size_t nt = pl->n_trk;
  DBuf<uint32_t> nseg, order, rank;
  DBuf<uint64_t> tkey, sstart;
  MUTO_CHECK_HIP(ctx, nseg.alloc(nt + 1));
  MUTO_CHECK_HIP(ctx, order.alloc(nt));
  MUTO_CHECK_HIP(ctx, rank.alloc(nt));
  MUTO_CHECK_HIP(ctx, tkey.alloc(nt));
  MUTO_CHECK_HIP(ctx, sstart.alloc(nt + 1));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(nseg.p, 0, (nt + 1) * sizeof(uint32_t), s));
  k_seg_count<<<grid_for(nt, 256), 256, 0, s>>>(G, nt, pl->nv, pl->row_ptr.p, pl->vox.p,
                                                nseg.p, tkey.p, order.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  MUTO_CHECK_HIP(ctx, sort_pairs_u64_u32(tkey.p, order.p, nt, 0, 64, s));
  k_invert<<<grid_for(nt, 256), 256, 0, s>>>(nt, order.p, rank.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  MUTO_CHECK_HIP(ctx, exclusive_scan_u32_to_u64(nseg.p, sstart.p, nt + 1, s));
  uint64_t n_seg = 0;
  MUTO_CHECK_HIP(ctx, hipMemcpy(&n_seg, sstart.p + nt, sizeof n_seg, hipMemcpyDeviceToHost));
  if (n_seg >= 0xFFFFFFFFull) { ctx->set_error("more than 2^32 segments"); return MUTO_ERR_LIMIT; }
  f.n_seg = n_seg;
  {
    DBuf<uint64_t> skey, st_start;
    DBuf<uint32_t> sval, st_len, st_trk, lenq, bcount;
    DBuf<float> st_tend;
    MUTO_CHECK_HIP(ctx, skey.alloc(n_seg));
    MUTO_CHECK_HIP(ctx, sval.alloc(n_seg));
    MUTO_CHECK_HIP(ctx, st_start.alloc(n_seg));
    MUTO_CHECK_HIP(ctx, st_len.alloc(n_seg));
    MUTO_CHECK_HIP(ctx, st_tend.alloc(n_seg));
    MUTO_CHECK_HIP(ctx, st_trk.alloc(n_seg));
    k_seg_fill<<<grid_for(nt, 256), 256, 0, s>>>(G, nt, pl->nv, pl->row_ptr.p, pl->vox.p,
                                                 pl->tail.p, sstart.p, rank.p, skey.p, sval.p,
                                                 st_start.p, st_len.p, st_tend.p, st_trk.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
    MUTO_CHECK_HIP(ctx, sort_pairs_u64_u32(skey.p, sval.p, n_seg, 0, 32 + bits_for(nbk), s));
    MUTO_CHECK_HIP(ctx, f.seg_trk.alloc(n_seg));
    MUTO_CHECK_HIP(ctx, f.seg_tend.alloc(n_seg));
    MUTO_CHECK_HIP(ctx, f.seg_off.alloc(n_seg + 1));
    MUTO_CHECK_HIP(ctx, lenq.alloc(n_seg + 1));
    MUTO_CHECK_HIP(ctx, bcount.alloc(nbk + 1));
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(lenq.p, 0, (n_seg + 1) * sizeof(uint32_t), s));
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(bcount.p, 0, (nbk + 1) * sizeof(uint32_t), s));
    if (n_seg) {
      k_seg_gather<<<grid_for(n_seg, 256), 256, 0, s>>>(n_seg, sval.p, st_len.p, st_tend.p, st_trk.p,
                                                        rank.p, skey.p, f.seg_trk.p, f.seg_tend.p,
                                                        lenq.p, bcount.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    }
    MUTO_CHECK_HIP(ctx, exclusive_scan_u32_to_u64(lenq.p, f.seg_off.p, n_seg + 1, s));
    MUTO_CHECK_HIP(ctx, hipMemcpy(&f.n_ent, f.seg_off.p + n_seg, sizeof(uint64_t), hipMemcpyDeviceToHost));
    MUTO_CHECK_HIP(ctx, f.ent_vox.alloc(f.n_ent));
    MUTO_CHECK_HIP(ctx, f.ent_len.alloc(f.n_ent));
    if (n_seg) {
      k_ent_scatter<<<grid_for(n_seg, 256), 256, 0, s>>>(G, n_seg, sval.p, st_start.p, st_len.p,
                                                         f.seg_off.p, pl->vox.p, pl->len.p,
                                                         f.ent_vox.p, f.ent_len.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    }
    // per-rank segment lists
    DBuf<uint32_t> nseg_rank;
    DBuf<uint64_t> tptr64;
    MUTO_CHECK_HIP(ctx, nseg_rank.alloc(nt + 1));
    MUTO_CHECK_HIP(ctx, tptr64.alloc(nt + 1));
    MUTO_CHECK_HIP(ctx, f.trk_d.alloc(4 * nt));
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(nseg_rank.p, 0, (nt + 1) * sizeof(uint32_t), s));
    k_trk_d<<<grid_for(nt, 256), 256, 0, s>>>(nt, order.p, pl->trk4.p, pl->pr.p, nseg.p,
                                              f.trk_d.p, nseg_rank.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
    MUTO_CHECK_HIP(ctx, exclusive_scan_u32_to_u64(nseg_rank.p, tptr64.p, nt + 1, s));
    std::vector<uint64_t> h64(nt + 1);
    MUTO_CHECK_HIP(ctx, hipMemcpy(h64.data(), tptr64.p, (nt + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost));
    std::vector<uint32_t> h32(h64.begin(), h64.end());
    MUTO_CHECK_HIP(ctx, f.tseg_ptr.alloc(nt + 1));
    MUTO_CHECK_HIP(ctx, hipMemcpy(f.tseg_ptr.p, h32.data(), (nt + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
    MUTO_CHECK_HIP(ctx, f.tseg_pos.alloc(n_seg));
    if (n_seg) {
      k_tseg_fill<<<grid_for(n_seg, 256), 256, 0, s>>>(n_seg, sval.p, st_trk.p, sstart.p, rank.p,
                                                       f.tseg_ptr.p, f.tseg_pos.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    }
    // bricks -> tasks
    DBuf<uint64_t> bstart;
    DBuf<unsigned long long> bent;
    MUTO_CHECK_HIP(ctx, bstart.alloc(nbk + 1));
    MUTO_CHECK_HIP(ctx, bent.alloc(nbk));
    MUTO_CHECK_HIP(ctx, exclusive_scan_u32_to_u64(bcount.p, bstart.p, nbk + 1, s));
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(bent.p, 0, nbk * sizeof(unsigned long long), s));
    if (n_seg) {
      k_brick_entries<<<grid_for(n_seg, 256), 256, 0, s>>>(n_seg, skey.p, f.seg_off.p, bent.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    }
    MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
    std::vector<uint64_t> hb(nbk + 1);
    std::vector<unsigned long long> he(nbk);
    MUTO_CHECK_HIP(ctx, hipMemcpy(hb.data(), bstart.p, (nbk + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost));
    MUTO_CHECK_HIP(ctx, hipMemcpy(he.data(), bent.p, nbk * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    // occupancy-derived persistent grid
    int occ_f = 1, occ_b = 1;
    const size_t lds_f = f.bs.vol() * sizeof(float), lds_b = 2 * f.bs.vol() * sizeof(float);
    MUTO_CHECK_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_f, k_fwd32, f.block, lds_f));
    MUTO_CHECK_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_b, k_bwd32, f.block, lds_b));
    f.grid_fwd = std::max(1, occ_f) * ctx->n_cu;
    f.grid_bwd = std::max(1, occ_b) * ctx->n_cu;
    // task size: at least 16 entries per brick voxel (flush amortised),
    // at most total / (4 * workers) so the queue balances
    const uint64_t workers = (uint64_t)std::max(f.grid_fwd, f.grid_bwd);
    uint64_t target = std::max<uint64_t>(16ull * f.bs.vol(), f.n_ent / (4 * workers) + 1);
    std::vector<Task> tasks;
    std::vector<uint64_t> tent;
    for (uint64_t b = 0; b < nbk; ++b) {
      const uint64_t s0 = hb[b], s1 = hb[b + 1];
      if (s1 == s0) continue;
      const uint64_t e = he[b];
      const uint64_t k = std::max<uint64_t>(1, (e + target - 1) / target);
      const uint64_t ns = s1 - s0;
      const uint64_t kk = std::min(k, ns);
      for (uint64_t t = 0; t < kk; ++t) {
        Task tk;
        tk.brick = (uint32_t)b;
        tk.owner = kk == 1 ? 1u : 0u;
        tk.s0 = s0 + ns * t / kk;
        tk.s1 = s0 + ns * (t + 1) / kk;
        tasks.push_back(tk);
        tent.push_back(e / kk);
      }
    }
    std::vector<size_t> idx(tasks.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](size_t x, size_t y) { return tent[x] > tent[y]; });
    std::vector<Task> sorted(tasks.size());
    for (size_t i = 0; i < idx.size(); ++i) sorted[i] = tasks[idx[i]];
    f.n_tasks = (uint32_t)sorted.size();
    MUTO_CHECK_HIP(ctx, f.tasks.alloc(sorted.size()));
    if (!sorted.empty())
      MUTO_CHECK_HIP(ctx, hipMemcpy(f.tasks.p, sorted.data(), sorted.size() * sizeof(Task), hipMemcpyHostToDevice));
    // bricks that no task owns completely get their accumulators by atomics
    // into acc64, which is zero between iterations
    MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  }
  MUTO_CHECK_HIP(ctx, f.seg_part.alloc(3 * n_seg));
  MUTO_CHECK_HIP(ctx, f.trk_g.alloc(4 * nt));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(pl->work.p, 0, 4 * sizeof(unsigned int), s));
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  f.built = true;
  pl->counters.n_segments = (int64_t)n_seg;
  return MUTO_OK;
}

int fp32c_iterate(muto_em_plan* pl, bool time_it) {
  Fp32cData& f = pl->f32;
  muto_ctx* ctx = pl->ctx;
  hipStream_t s = ctx->stream;
  const IterArgs a = iter_args(pl);
  const size_t lds_f = f.bs.vol() * sizeof(float), lds_b = 2 * f.bs.vol() * sizeof(float);
  const ErrM E = error_matrix(pl->cfg);
  const bool map = pl->is_map;
  if (map && !pl->tmp64.p) MUTO_CHECK_HIP(ctx, pl->tmp64.alloc(pl->nv));
  if (time_it) MUTO_CHECK_HIP(ctx, hipEventRecord(pl->ev[0], s));
  k_fwd32<<<f.grid_fwd, f.block, lds_f, s>>>(a, pl->lam32.p, f.seg_part.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  if (time_it) MUTO_CHECK_HIP(ctx, hipEventRecord(pl->ev[1], s));
  k_track32<<<grid_for(pl->n_trk, 256), 256, 0, s>>>(pl->n_trk, f.tseg_ptr.p, f.tseg_pos.p, f.n_seg,
                                                      f.seg_part.p, f.trk_d.p, E, f.trk_g.p,
                                                      pl->cost_acc.p, pl->dcount.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  if (time_it) MUTO_CHECK_HIP(ctx, hipEventRecord(pl->ev[2], s));
  k_bwd32<<<f.grid_bwd, f.block, lds_b, s>>>(a, f.trk_g.p, pl->acc64.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  if (time_it) MUTO_CHECK_HIP(ctx, hipEventRecord(pl->ev[3], s));
  k_upd32<<<grid_for(pl->nv, 256), 256, 0, s>>>(pl->nv, pl->mj.p, pl->acc64.p, pl->lam32.p,
                                                map ? pl->ml64.p : nullptr,
                                                map ? pl->lam64.p : nullptr, pl->work.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  if (map) {
    if (int rc = map_step(pl, pl->lam64.p, pl->ml64.p, pl->tmp64.p)) return rc;
    k_f64_to_f32<<<grid_for(pl->nv, 256), 256, 0, s>>>(pl->nv, pl->tmp64.p, pl->lam32.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
  }
  if (time_it) {
    MUTO_CHECK_HIP(ctx, hipEventRecord(pl->ev[4], s));
    MUTO_CHECK_HIP(ctx, hipEventSynchronize(pl->ev[4]));
    float t[4];
    for (int k = 0; k < 4; ++k)
      MUTO_CHECK_HIP(ctx, hipEventElapsedTime(&t[k], pl->ev[k], pl->ev[k + 1]));
    pl->t_ms[0] = t[0];
    pl->t_ms[1] = t[1];
    pl->t_ms[2] = t[2];
    pl->t_ms[3] = t[3];
    pl->t_ms[4] = t[0] + t[1] + t[2] + t[3];
  }
  return MUTO_OK;
}

}  // namespace muto
