// FP32C EM iteration: brick-major segment stream, LDS-staged gathers and
// backprojection, FP32 entry arithmetic, FP64 2x2 algebra.
//
// Layout (built once per plan from the track-major FP64 lists):
//   - the grid is cut into bricks (LDS tiles) of bx*by*bz voxels;
//   - each track's merged path is cut into segments: maximal runs of
//     consecutive entries inside one brick;
//   - tracks are ranked by a spatial key (direction bin, then the Morton
//     code of where the track crosses the top plane), segments sorted by
//     (brick, track rank): the tracks a block sees together are nearly
//     parallel neighbours;
//   - each brick's segment list is cut into tasks (the persistent kernels'
//     work units) and each task into chunks of at most 256 segments and
//     ECAP entries;
//   - entries are stored brick-major as one 32-bit word each (the FP32
//     length and a 3-bit step from the previous voxel of the segment; a
//     segment's first voxel is kept per segment); each chunk starts on a
//     16-entry boundary and is padded to whole 16-entry groups with zero
//     words, and each group has a descriptor (segment-start mask, the
//     segment holding its first entry, its first voxel, the length of its
//     last segment after the group): about 4.6 bytes per entry in all;
//   - every segment starts on a SEG_ALIGN-entry boundary: a segment is
//     padded to a multiple of SEG_ALIGN entries with padding words (step
//     code 7, zero length payload; zero FP64 length), so the walks test for
//     a segment start at one entry in SEG_ALIGN only.
// The passes over entries are entry-parallel: thread t of a block takes the
// chunk's group t (16-byte vector loads) and walks it backwards, the tail of
// each entry inside its segment following from the group's carry; every
// thread does the same work, whatever the segment lengths.
// Per iteration (src/MutoMLSD.cpp:179-213, 442-517, mean update):
//   k_fwd32   per segment: partial covariance sums sum lambda_j W_ij with
//             lambda staged in LDS, W from (L, tail) in registers, in the
//             segment's local tail basis;
//   k_track32 per track: Sigma_i in FP64 from its segment partials, Eigen's
//             2x2 det/inverse in FP64, cost_i, G_i in FP32;
//   k_bwd32   per entry: c_ij = <W_ij, G_i> accumulated per voxel in LDS,
//             flushed once per task;
//   k_upd32   per voxel: lambda += lambda^2/(2 M_j) sum_i c_ij, which is
//             half the mean of S_ij = 2 lambda_j + c_ij lambda_j^2.
// Accumulation (acc_mode): 1 (default) FP64 sums of the FP32 entry terms
// per segment piece (at most 16 entries) in the forward walk, each rounded
// once to FP32 and combined per segment in FP64, and one int64 fixed-point
// LDS atomic per entry in the backprojection; 0 FP32 with
// two-sum compensation -- error-free transformation in the forward walk,
// atomic two-sum (returned old value + error word) in the backprojection.
// The persistent kernels take tasks from a queue sorted by size, largest
// first, through an atomic counter.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <type_traits>
#include <vector>

#include "em_dev.hpp"
#include "em_plan.hpp"

namespace muto {

int map_step(muto_em_plan* pl, const double* last, const double* ml, double* out);
int fp32c_mirror(muto_em_plan* pl);
int fp32c_run(muto_em_plan* pl, int iters, double* sums, int* done);
int fp32c_build(muto_em_plan* pl);

namespace {

// threads per block of the persistent passes: 256, 512 or 1024 (the brick
// shape's `block`); a chunk holds up to one segment per thread and GE
// entries per thread
constexpr int GE = 16;                   // entries per thread and chunk
// segments start on multiples of SEG_ALIGN entries (padding words inside
// the segment's last SEG_ALIGN run): the walks' start tests, voxel selects
// and first-voxel reads then exist at one entry position in SEG_ALIGN
constexpr int SEG_ALIGN = 2;
static_assert(GE % SEG_ALIGN == 0, "groups hold whole alignment runs");
constexpr uint32_t ENT_PAD_CODE = 7u;
constexpr uint32_t ENT_PAD = ENT_PAD_CODE << 29;  // a padding word
__host__ __device__ constexpr uint32_t seg_padded(uint32_t n) {
  return (n + (uint32_t)SEG_ALIGN - 1u) / (uint32_t)SEG_ALIGN * (uint32_t)SEG_ALIGN;
}
// FP32C median by brackets: values kept inside a voxel's bracket, and ranks
// kept on either side of the median for the next bracket
constexpr uint32_t MED_K = 32;
constexpr uint32_t MED_H = 6;
constexpr uint32_t DIR_BINS = 8;         // track-order direction bins per unit slope
// backprojection sums (ACC 1) are int64 fixed point: an entry is at most
// 2^fx_bits, fx_bits = min(FX_BITS, 62 - bits(max M_j)) per plan; a task
// holds at most MAX_TASK_SEG segments
constexpr int FX_BITS = 46;
constexpr uint32_t MAX_TASK_SEG = 1u << 16;
// static LDS of both passes is a multiple of 16 bytes, so the dynamic region
// after it is 16-byte aligned
// (forward: piece tables, first voxels, step table, task word; backward:
// G' table, step table, task word)
constexpr size_t fwd_static_lds(uint32_t bs) { return 2 * (size_t)bs * 16 + (size_t)bs * 4 + 64 + 16; }
// A chunk holds at most block - 2 segments.  The backprojection can take
// two chunks per workgroup of twice the block size: its G' table for them
// (2 (block - 2) entries of 16 bytes), step table and task word then come to
// 16368 bytes at block 512, so two such workgroups of a 16 x 16 x 32 brick
// (64 KiB of int64 sums each) fit the 160 KiB of LDS of one CU.
constexpr uint32_t chunk_seg_cap(uint32_t bs) { return bs - 2; }
constexpr size_t bwd_static_lds(uint32_t bs, uint32_t sub) {
  return (size_t)sub * chunk_seg_cap(bs) * 16 + 32 + 16;
}
static_assert(fwd_static_lds(256) % 16 == 0 && fwd_static_lds(512) % 16 == 0 &&
              fwd_static_lds(1024) % 16 == 0, "forward static LDS");
static_assert(bwd_static_lds(256, 1) % 16 == 0 && bwd_static_lds(512, 1) % 16 == 0 &&
              bwd_static_lds(1024, 1) % 16 == 0 && bwd_static_lds(256, 2) % 16 == 0 &&
              bwd_static_lds(512, 2) % 16 == 0, "backward static LDS");
static_assert(bwd_static_lds(512, 2) + 65536 <= 81920, "two backprojection workgroups per CU");

// Unsigned division by a divisor fixed for a launch: with l = ceil(log2 d)
// and m = floor(2^32 (2^l - d) / d) + 1, n / d = (umulhi(n, m) + n) >> l for
// every 32-bit n (the sum taken in 64 bits).  Four VALU operations instead
// of the ~30 of a general 32-bit division.
struct FastDiv {
  uint32_t d = 1, m = 1, l = 0;
  static FastDiv make(uint32_t d) {
    FastDiv f;
    f.d = d;
    f.l = 0;
    while (f.l < 32 && (1ull << f.l) < d) ++f.l;
    f.m = (uint32_t)(((((1ull << f.l) - d) << 32) / d) + 1);
    return f;
  }
  __device__ __forceinline__ uint32_t div(uint32_t n) const {
    const uint32_t t = __umulhi(n, m);
    return (uint32_t)(((uint64_t)t + n) >> l);
  }
};

struct BrickGeo {
  uint32_t n[3];   // grid
  uint32_t b[3];   // brick shape
  uint32_t nb[3];  // bricks per axis
  FastDiv dn0, dn1, db0, db1, db2;  // by n[0], n[1], b[0], b[1], b[2]
};

// Coordinates of voxel j (x fastest), its brick and its index in the brick.
struct Vox {
  uint32_t x, y, z, brick, local;
};
__device__ __forceinline__ Vox locate(const BrickGeo& G, uint32_t j) {
  Vox v;
  const uint32_t t = G.dn0.div(j);
  v.x = j - t * G.n[0];
  v.z = G.dn1.div(t);
  v.y = t - v.z * G.n[1];
  const uint32_t bx = G.db0.div(v.x), by = G.db1.div(v.y), bz = G.db2.div(v.z);
  v.local = (v.x - bx * G.b[0]) + G.b[0] * ((v.y - by * G.b[1]) + G.b[1] * (v.z - bz * G.b[2]));
  v.brick = bx + G.nb[0] * (by + G.nb[1] * bz);
  return v;
}

// Packed track-voxel entry of the FP32C stream (4 bytes).
//   bits 28..0: the path length L in units of U (a power of two above every
//               length: twice the voxel diagonal, so a merged two-leg entry
//               fits) as the FP32 bit pattern of L/U minus that of 2^-63.
//               L/U lies in [2^-63, 1), so the sign and the top exponent
//               bit are constant and the 23 mantissa bits are kept: the
//               length is L rounded to FP32, as an FP32 path list would
//               hold it.  One AND-OR decodes it.  Shorter lengths, and the
//               zero words that pad a chunk, read as 2^-63 U.
//   bits 31..29: the step from the segment's previous voxel: 0 none, 1/2
//               +-x, 3/4 +-y, 5/6 +-z.  Other steps start a new segment.
// A segment's first voxel is kept per segment.
constexpr uint32_t ENT_PMASK = 0x1FFFFFFFu;
constexpr uint32_t ENT_BASE = 0x20000000u;  // FP32 bits of 2^-63
__host__ __device__ __forceinline__ float ent_len(uint32_t w) {  // in units of U
  return __uint_as_float((w & ENT_PMASK) | ENT_BASE);
}
__host__ __device__ __forceinline__ uint32_t ent_code(uint32_t w) { return w >> 29; }
struct EntCode {
  double unit;  // U in mm
  int bx, bxy;  // brick-local index strides
  // brick-local index step of a step code
  __host__ __device__ int step(uint32_t code) const {
    switch (code) {
      case 1: return 1;
      case 2: return -1;
      case 3: return bx;
      case 4: return -bx;
      case 5: return bxy;
      case 6: return -bxy;
      default: return 0;
    }
  }
};

// A track's entry continues the segment of the entry before when both lie
// in one brick and the step between them moves along at most one axis by
// one voxel (the steps a packed entry can hold).
__device__ __forceinline__ bool seg_continues(const Vox& p, const Vox& c) {
  if (p.brick != c.brick) return false;
  return abs((int)c.x - (int)p.x) + abs((int)c.y - (int)p.y) + abs((int)c.z - (int)p.z) <= 1;
}

__device__ __forceinline__ uint32_t spread12(uint32_t v) {
  // interleave the 12 low bits of v with zeros (Morton)
  v &= 0xFFFu;
  v = (v | (v << 8)) & 0x00FF00FFu;
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

// ---------------- build ------------------------------------------------
// Segments per track and the track's spatial key: direction bin of the
// chord from its first to its last voxel (bins of 1/DIR_BINS in dx/|dz| and
// dy/|dz| over [-2, 2]) above the Morton code of where that chord crosses
// the top plane of the grid; in a batch of scans the scan comes first, so
// each scan's tracks form one run of ranks.
// a voxel's coordinates and brick from another lane of the wavefront
__device__ __forceinline__ Vox shfl_vox(const Vox& v, int src) {
  Vox r;
  r.x = __shfl(v.x, src, 64);
  r.y = __shfl(v.y, src, 64);
  r.z = __shfl(v.z, src, 64);
  r.brick = __shfl(v.brick, src, 64);
  r.local = 0;
  return r;
}
__device__ __forceinline__ Vox shfl_up_vox(const Vox& v) {
  Vox r;
  r.x = __shfl_up(v.x, 1u, 64);
  r.y = __shfl_up(v.y, 1u, 64);
  r.z = __shfl_up(v.z, 1u, 64);
  r.brick = __shfl_up(v.brick, 1u, 64);
  r.local = 0;
  return r;
}

// One wavefront per track: the lanes read 64 consecutive entries of the
// track's list at a time; each entry's predecessor comes from the lane below
// (lane 0: the last entry of the previous 64).
__global__ void __launch_bounds__(256)
k_seg_count(BrickGeo G, size_t n_trk, size_t nv,
            const uint64_t* __restrict__ row_ptr,
            const uint32_t* __restrict__ vox, const uint16_t* __restrict__ scan,
            uint32_t* __restrict__ nseg, uint64_t* __restrict__ key, uint32_t* __restrict__ val) {
  const int lane = (int)(threadIdx.x & 63u);
  const size_t t = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (t >= n_trk) return;  // uniform over the wavefront
  const uint64_t b = row_ptr[t], e = row_ptr[t + 1];
  uint32_t cnt = 0;
  bool have = false, pcv = false;
  Vox vf{}, vl{}, pc{};
  for (uint64_t c0 = b; c0 < e; c0 += 64) {
    const uint64_t k = c0 + (uint64_t)lane;
    const bool in = k < e;
    const uint32_t j = in ? vox[k] : 0xFFFFFFFFu;
    const bool valid = in && j < nv;
    const Vox cv = valid ? locate(G, j) : Vox{};
    Vox pv = shfl_up_vox(cv);
    bool pval = __shfl_up((int)valid, 1u, 64) != 0;
    if (lane == 0) { pv = pc; pval = pcv; }
    const bool start = valid && !(pval && seg_continues(pv, cv));
    cnt += (uint32_t)__popcll(__ballot(start));
    const unsigned long long vm = __ballot(valid);
    if (vm) {
      if (!have) {
        vf = shfl_vox(cv, __ffsll((long long)vm) - 1);
        have = true;
      }
      vl = shfl_vox(cv, 63 - __clzll((long long)vm));
    }
    pc = shfl_vox(cv, 63);
    pcv = __shfl((int)valid, 63, 64) != 0;
  }
  if (lane != 0) return;
  nseg[t] = cnt;
  val[t] = (uint32_t)t;
  if (!have) { key[t] = ~0ull; return; }
  const float x0 = (float)vf.x, y0 = (float)vf.y, z0 = (float)vf.z;
  const float x1 = (float)vl.x, y1 = (float)vl.y, z1 = (float)vl.z;
  const float dz = fmaxf(fabsf(z1 - z0), 1.0f);
  const float sx = fminf(fmaxf((x1 - x0) / dz, -2.0f), 2.0f);
  const float sy = fminf(fmaxf((y1 - y0) / dz, -2.0f), 2.0f);
  // slope bins of 1/DIR_BINS: over a 256-voxel grid height the tracks of
  // one bin spread by about two bricks, so the tracks crossing a brick in
  // one bin are a dense run of ranks, and the track kernel's gathers of
  // neighbouring ranks share lines (1/8 measured 9% faster there than 1/16)
  const uint32_t dbin = (uint32_t)((sx + 2.0f) * DIR_BINS + 0.5f) * (4u * DIR_BINS + 1u) +
                        (uint32_t)((sy + 2.0f) * DIR_BINS + 0.5f);
  const float s = (z1 != z0) ? ((float)G.n[2] - z0) / (z1 - z0) : 0.0f;
  const float xt = fminf(fmaxf(x0 + s * (x1 - x0) + 1024.0f, 0.0f), 4095.0f);
  const float yt = fminf(fmaxf(y0 + s * (y1 - y0) + 1024.0f, 0.0f), 4095.0f);
  const uint32_t mort = spread12((uint32_t)xt) | (spread12((uint32_t)yt) << 1);
  key[t] = ((scan ? (uint64_t)scan[t] : 0ull) << 48) | ((uint64_t)dbin << 24) | mort;
}

__global__ void k_invert(size_t n, const uint32_t* __restrict__ order,
                         uint32_t* __restrict__ rank) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) rank[order[r]] = (uint32_t)r;
}

// Track-major segments: key (brick, rank), first entry, length, tail after
// the last entry (FP64), track.
// Segment records, one wavefront per track as in k_seg_count: a segment
// start writes its key, first entry and track; the segment's last entry
// (followed by a start, an entry outside the grid or the end of the list)
// writes its length and the tail after it.
__global__ void __launch_bounds__(256)
k_seg_fill(BrickGeo G, size_t n_trk, size_t nv,
           const uint64_t* __restrict__ row_ptr,
           const uint32_t* __restrict__ vox, const double* __restrict__ tail,
           const uint64_t* __restrict__ sstart,
           const uint32_t* __restrict__ rank, uint64_t* __restrict__ skey,
           uint32_t* __restrict__ sval, uint64_t* __restrict__ st_start,
           uint32_t* __restrict__ st_len, double* __restrict__ st_tend,
           uint32_t* __restrict__ st_trk) {
  const int lane = (int)(threadIdx.x & 63u);
  const size_t t = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (t >= n_trk) return;  // uniform over the wavefront
  const uint64_t b = row_ptr[t], e = row_ptr[t + 1];
  const uint64_t s0 = sstart[t];
  const uint32_t r = rank[t];
  const unsigned long long upto = lane == 63 ? ~0ull : (2ull << lane) - 1ull;  // lanes <= this one
  uint64_t before = 0;      // segments started in earlier chunks
  uint64_t last_start = b;  // list position of the latest start in earlier chunks
  bool pcv = false;
  Vox pc{};
  for (uint64_t c0 = b; c0 < e; c0 += 64) {
    const uint64_t k = c0 + (uint64_t)lane;
    const bool in = k < e;
    const uint32_t j = in ? vox[k] : 0xFFFFFFFFu;
    const bool valid = in && j < nv;
    const Vox cv = valid ? locate(G, j) : Vox{};
    Vox pv = shfl_up_vox(cv);
    bool pval = __shfl_up((int)valid, 1u, 64) != 0;
    if (lane == 0) { pv = pc; pval = pcv; }
    const bool start = valid && !(pval && seg_continues(pv, cv));
    const unsigned long long sm = __ballot(start);
    // whether entry k + 1 starts a segment or lies outside the grid / list
    bool nvalid = __shfl_down((int)valid, 1u, 64) != 0;
    bool nstart = __shfl_down((int)start, 1u, 64) != 0;
    if (lane == 63) {
      const bool in2 = k + 1 < e;
      const uint32_t j2 = in2 ? vox[k + 1] : 0xFFFFFFFFu;
      nvalid = in2 && j2 < nv;
      nstart = nvalid && !(valid && seg_continues(cv, locate(G, j2)));
    }
    const bool last = valid && (!nvalid || nstart);
    // position of the latest start at or before this entry: a max-scan of
    // (offset + 1) over the lanes, 0 where none in this chunk
    uint32_t sp = start ? (uint32_t)(k - b + 1) : 0u;
#pragma unroll
    for (unsigned o = 1; o < 64; o <<= 1) {
      const uint32_t w = __shfl_up(sp, o, 64);
      if ((unsigned)lane >= o) sp = max(sp, w);
    }
    const uint64_t spos = sp ? b + sp - 1 : last_start;
    const uint64_t cur = s0 + before + (uint64_t)__popcll(sm & upto) - 1;  // this entry's segment
    if (start) {
      skey[cur] = ((uint64_t)cv.brick << 32) | r;
      sval[cur] = (uint32_t)cur;
      st_start[cur] = k;
      st_trk[cur] = (uint32_t)t;
    }
    if (last) {
      st_len[cur] = (uint32_t)(k - spos + 1);
      st_tend[cur] = tail[k];
    }
    before += (uint64_t)__popcll(sm);
    const uint32_t sp63 = __shfl(sp, 63, 64);
    if (sp63) last_start = b + sp63 - 1;
    pc = shfl_vox(cv, 63);
    pcv = __shfl((int)valid, 63, 64) != 0;
  }
}

// lengths in brick-major order and the first segment of each brick
__global__ void __launch_bounds__(256)
k_seg_len(uint64_t n_seg, const uint32_t* __restrict__ sval,
          const uint32_t* __restrict__ st_len, const uint64_t* __restrict__ skey,
          uint32_t* __restrict__ lenq, uint64_t* __restrict__ bstart) {
  uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_seg) return;
  lenq[q] = st_len[sval[q]];
  const uint32_t b = (uint32_t)(skey[q] >> 32);
  if (q == 0 || (uint32_t)(skey[q - 1] >> 32) != b) bstart[b] = q;
}

// a segment's length rounded up to whole alignment runs (its stream size)
__global__ void k_pad_len(uint64_t n_seg, uint32_t* __restrict__ lenq) {
  uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n_seg) lenq[q] = seg_padded(lenq[q]);
}

// Track rank and tail after the segment, as unsigned fixed point with
// 2^-tail_bits mm resolution (tails are >= 0 and below the fixed-point
// range, which the host sizes from the grid diagonal).
__global__ void __launch_bounds__(256)
k_seg_gather(uint64_t n_seg, const uint32_t* __restrict__ sval,
             const double* __restrict__ st_tend,
             const uint32_t* __restrict__ st_trk,
             const uint32_t* __restrict__ rank, double tail_scale,
             uint32_t* __restrict__ seg_trk, uint32_t* __restrict__ seg_tfx,
             double* __restrict__ seg_tail64) {
  uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_seg) return;
  const uint32_t s = sval[q];
  seg_trk[q] = rank[st_trk[s]];
  seg_tfx[q] = (uint32_t)llrint(st_tend[s] * tail_scale);
  if (seg_tail64) seg_tail64[q] = st_tend[s];  // the FP64 stream: the exact tail
}

// Entry offset of every segment: its chunk's base (16-entry aligned, set by
// the host) plus its offset inside the chunk.
__global__ void k_chunk_offsets(uint32_t n, const Chunk* __restrict__ chunks,
                                const uint64_t* __restrict__ seg_off,
                                uint64_t* __restrict__ seg_off_p) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const Chunk ch = chunks[c];
  const uint64_t b = seg_off[ch.q0];
  for (uint32_t t = 0; t < ch.cnt; ++t) seg_off_p[ch.q0 + t] = ch.base + (seg_off[ch.q0 + t] - b);
}

// Descriptor of every 16-entry group (one block per chunk, one thread per
// group): segment-start mask, the chunk segment holding its first entry and
// the length of its last segment after the group (summed in FP64).
template <uint32_t BS>
__global__ void __launch_bounds__(BS)
k_group_desc(const Chunk* __restrict__ chunks, const uint64_t* __restrict__ seg_off_p,
             const uint32_t* __restrict__ ent, const uint16_t* __restrict__ seg_v0, EntCode ec,
             uint32_t* __restrict__ gdesc, float* __restrict__ gcarry, uint16_t* __restrict__ gv0) {
  __shared__ uint32_t so[BS + 1];
  const Chunk ch = chunks[blockIdx.x];
  const uint32_t t = threadIdx.x;
  if (t < ch.cnt) so[t] = (uint32_t)(seg_off_p[ch.q0 + t] - ch.base);
  if (t == 0) so[ch.cnt] = ch.ne;
  __syncthreads();
  const uint32_t e0 = t * GE;
  if (e0 >= ch.ne) return;
  uint32_t lo = 0, hi = ch.cnt;  // so[lo] <= e0 < so[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (so[mid] <= e0) lo = mid; else hi = mid;
  }
  uint32_t mask = 0, s = lo;
  for (uint32_t k = lo; k < ch.cnt && so[k] < e0 + GE; ++k) {
    if (so[k] >= e0) mask |= 1u << (so[k] - e0);
    s = k;
  }
  double cr = 0.0;  // in units of U
  for (uint32_t e = e0 + GE; e < so[s + 1]; ++e) {
    const uint32_t w = ent[ch.base + e];
    if (ent_code(w) != ENT_PAD_CODE) cr += (double)ent_len(w);  // padding: zero length
  }
  // voxel of entry e0: its segment's first voxel plus the steps up to it
  uint32_t v = seg_v0[ch.q0 + lo];
  for (uint32_t e = so[lo] + 1; e <= e0; ++e) v += ec.step(ent_code(ent[ch.base + e]));
  const uint64_t gi = ch.base / GE + t;
  gdesc[gi] = mask | (lo << 16);
  gcarry[gi] = (float)cr;
  gv0[gi] = (uint16_t)v;
}

// Packed entries of one chunk per block, one thread per entry (strided by
// the block size, so the loads of a wavefront cover a few whole segments and
// the stores are contiguous): the chunk's segments are listed in LDS with
// their source position in the track-major lists, and each entry finds its
// segment through an entry -> segment table.
template <uint32_t BS>
__global__ void __launch_bounds__(BS)
k_ent_fill(BrickGeo G, const Chunk* __restrict__ chunks, const uint32_t* __restrict__ sval,
           const uint64_t* __restrict__ st_start, const uint32_t* __restrict__ st_len,
           const uint64_t* __restrict__ seg_off_p,
           const uint32_t* __restrict__ vox, const double* __restrict__ len, double inv_unit,
           uint32_t* __restrict__ ent, uint16_t* __restrict__ seg_v0,
           unsigned long long* __restrict__ n_clamped, const double* __restrict__ tail,
           double* __restrict__ len64, double* __restrict__ gtail64) {
  __shared__ uint64_t s_src[BS];
  __shared__ uint32_t s_off[BS], s_len[BS];
  __shared__ uint16_t s_id[GE * BS];
  const Chunk ch = chunks[blockIdx.x];
  const uint32_t t = threadIdx.x;
  if (t < ch.cnt) {
    const uint64_t q = ch.q0 + t;
    const uint32_t off = (uint32_t)(seg_off_p[q] - ch.base);
    const uint32_t end = t + 1 < ch.cnt ? (uint32_t)(seg_off_p[q + 1] - ch.base) : ch.ne;
    s_src[t] = st_start[sval[q]];
    s_off[t] = off;
    s_len[t] = st_len[sval[q]];  // real entries; the rest up to `end` is padding
    for (uint32_t e = off; e < end; ++e) s_id[e] = (uint16_t)t;
  }
  __syncthreads();
  const int64_t nx = G.n[0], nxy = (int64_t)G.n[0] * G.n[1];
  uint32_t clamped = 0;
  for (uint32_t e = t; e < ch.ne; e += BS) {
    const uint32_t i = s_id[e];
    const uint32_t k = e - s_off[i];
    if (k >= s_len[i]) {
      // padding of the segment's last alignment run: no step, zero length
      // (len64 stays 0); a group anchored here takes the tail of the
      // segment's last real entry (the same tail after a zero length)
      ent[ch.base + e] = ENT_PAD;
      if (len64 && ((e % GE) == GE - 1 || e + 1 == ch.ne))
        gtail64[(ch.base + e) / GE] = tail[s_src[i] + s_len[i] - 1];
      continue;
    }
    const uint64_t src = s_src[i] + k;
    const uint32_t j = vox[src];
    uint32_t code = 0;
    if (k == 0) {
      // a segment's first voxel goes to seg_v0; every later entry stores
      // its step from the one before
      seg_v0[ch.q0 + i] = (uint16_t)locate(G, j).local;
    } else {
      // inside a segment the step moves along at most one axis by one voxel
      // inside one brick, so the index difference names it: +-nx*ny is z,
      // else +-nx is y, else +-1 is x (an axis of one voxel never steps, so
      // the order settles nx = 1 or ny = 1)
      const int64_t dd = (int64_t)j - (int64_t)vox[src - 1];
      code = dd == nxy ? 5u : dd == -nxy ? 6u : dd == nx ? 3u : dd == -nx ? 4u
           : dd == 1 ? 1u : dd == -1 ? 2u : 0u;
    }
    float lu = (float)(len[src] * inv_unit);  // L / U rounded to FP32
    if (!(lu >= 0.0f && lu < 1.0f)) {  // outside the packed range: reported
      ++clamped;
      lu = lu > 0.0f ? 0.99999994f : 0.0f;
    }
    const uint32_t bits = __float_as_uint(lu);
    const uint32_t pay = bits >= ENT_BASE ? bits - ENT_BASE : 0u;
    ent[ch.base + e] = code << 29 | pay;
    if (len64) {
      // the FP64 stream: the exact length, and per 16-entry group the exact
      // tail (the reference's tail array) of its entry 15, or of the chunk's
      // last entry when entry 15 is padding (zero length, the same tail)
      len64[ch.base + e] = len[src];
      if ((e % GE) == GE - 1 || e + 1 == ch.ne) gtail64[(ch.base + e) / GE] = tail[src];
    }
  }
  if (clamped) atomicAdd(n_clamped, (unsigned long long)clamped);
}

// Brick-major position of each track's segments, listed in rank order.
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

// longest path over the tracks (tail of the first entry plus its length),
// as the bit pattern of a non-negative double (ordered like the values)
__global__ void __launch_bounds__(256)
k_max_path(size_t n_trk, const uint64_t* __restrict__ row_ptr, const double* __restrict__ len,
           const double* __restrict__ tail, unsigned long long* __restrict__ out) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (t < n_trk && row_ptr[t + 1] > row_ptr[t]) v = fmax(tail[row_ptr[t]] + len[row_ptr[t]], 0.0);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  // one atomic per block (same-address atomics serialise)
  __shared__ double wm[4];
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    v = fmax(fmax(wm[0], wm[1]), fmax(wm[2], wm[3]));
    if (v > 0.0) atomicMax(out, (unsigned long long)__double_as_longlong(v));
  }
}

// Per-track data in rank order.
__global__ void k_trk_d(size_t n_trk, const uint32_t* __restrict__ order,
                        const double* __restrict__ trk4,
                        const double* __restrict__ pr,
                        const uint32_t* __restrict__ nseg,
                        double* __restrict__ trk_d,
                        uint32_t* __restrict__ nseg_rank,
                        const uint16_t* __restrict__ scan_in,
                        uint16_t* __restrict__ scan_rank) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_trk) return;
  const uint32_t t = order[r];
  if (scan_rank) scan_rank[r] = scan_in[t];
  const double Sx = trk4[4 * t], Sy = trk4[4 * t + 1];
  const double Dx = trk4[4 * t + 2], Dy = trk4[4 * t + 3];
  trk_d[4 * r] = Sx * Sx + Sy * Sy;
  trk_d[4 * r + 1] = Sx * Dx + Sy * Dy;
  trk_d[4 * r + 2] = Dx * Dx + Dy * Dy;
  trk_d[4 * r + 3] = pr[t];
  nseg_rank[r] = nseg[t];
}

// ---------------- iteration --------------------------------------------
__device__ __forceinline__ void two_sum_acc(float& hi, float& lo, float x) {
  const float s = hi + x;
  const float bp = s - hi;
  const float e = (hi - (s - bp)) + (x - bp);
  hi = s;
  lo += e;
}

// Fixed-point scale of the backprojection sums: 2^fx_bits / (the bound on
// |c_ij| the track kernel found this pass), so every scaled entry is below
// 2^fx_bits.  fx_bits is set per plan so that M_j entries of one voxel stay
// inside int64.
__device__ __forceinline__ double fx_scale(const unsigned long long* cmax, int fx_bits) {
  const double cb = __longlong_as_double((long long)cmax[0]);
  return cb > 0.0 ? ldexp(1.0, fx_bits) / cb : 1.0;
}

struct IterArgs {
  BrickGeo G;
  const Task* tasks;
  const Chunk* chunks;
  uint32_t n_tasks;
  unsigned int* work;
  int fx_bits;
  uint64_t n_seg;
  const uint32_t* seg_tfx;
  const uint32_t* gdesc;
  const float* gcarry;
  double inv_scale;
  const uint32_t* seg_trk;
  const uint16_t* seg_v0;
  const uint32_t* ent;
  const uint16_t* gv0;
  EntCode ec;
};

// Thread t's segment of a chunk, and its group.  Threads past the chunk's
// segments (groups) read the chunk's last segment (first group) again and
// never use it: the loads are then unconditional, so the compiler counts
// exactly which of them a later use has to wait for, and the repeated
// addresses add no memory traffic (a wavefront's equal addresses are one
// request).
__device__ __forceinline__ uint64_t seg_slot(const IterArgs& a, const Chunk& ch, uint32_t t) {
  const uint32_t last = ch.cnt > 0 ? ch.cnt - 1 : 0u;
  return min(ch.q0 + min(t, last), a.n_seg - 1);
}
__device__ __forceinline__ uint64_t group_slot(const Chunk& ch, uint32_t t) {
  return ch.base / GE + (t * GE < ch.ne ? t : 0u);
}

// A chunk's descriptor, read through the constant address space: with a
// wave-uniform index that is a scalar load, which takes no vector registers
// and is not ordered with the vector loads, so the forward pass issues the
// next chunk's while it walks the current one.
__device__ __forceinline__ Chunk chunk_at(const IterArgs& a, uint32_t c) {
  typedef const __attribute__((address_space(4))) Chunk* ConstChunkPtr;
  const ConstChunkPtr p = (ConstChunkPtr)a.chunks + c;
  Chunk r;
  r.q0 = p->q0;
  r.base = p->base;
  r.cnt = p->cnt;
  r.ne = p->ne;
  return r;
}

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
    const uint32_t r = G.db0.div(v), lx = v - r * G.b[0];
    const uint32_t lz = G.db1.div(r), ly = r - lz * G.b[1];
    const uint32_t x = o[0] + lx, y = o[1] + ly, z = o[2] + lz;
    float val = 0.0f;
    if (x < G.n[0] && y < G.n[1] && z < G.n[2])
      val = lam[(size_t)x + (size_t)G.n[0] * ((size_t)y + (size_t)G.n[1] * z)];
    lds[v] = val;
  }
}

// per bit: m ? a : b (v_bfi_b32).  Written as an instruction so the
// compiler cannot turn the select of an LDS value into a branch around the
// LDS read, which would put the read's latency into a dependency chain.
__device__ __forceinline__ uint32_t bit_select(uint32_t m, uint32_t a, uint32_t b) {
  uint32_t r;
  asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "v"(m), "v"(a), "v"(b));
  return r;
}

// L/U of a packed entry in one v_and_or_b32 (the mask from an SGPR, the
// constant exponent bits from a VGPR: a gfx9 VOP3 instruction reads at
// most one scalar operand)
__device__ __forceinline__ float ent_len_v(uint32_t w, uint32_t base_v) {
  uint32_t r;
  asm("v_and_or_b32 %0, %1, %2, %3" : "=v"(r) : "v"(w), "s"(ENT_PMASK), "v"(base_v));
  return __uint_as_float(r);
}

// all ones when bit i of m is set, else zero (one v_bfe_i32)
__device__ __forceinline__ uint32_t bit_mask(uint32_t m, int i) {
  return (uint32_t)((int)(m << (31 - i)) >> 31);
}

// all ones when entry i of a group starts a segment (i a constant of the
// unrolled walk: zero without a test where no segment can start)
__device__ __forceinline__ uint32_t start_mask(uint32_t m, int i) {
  return i % SEG_ALIGN == 0 ? bit_mask(m, i) : 0u;
}

// One thread's group of 16 consecutive entries of a chunk (chunks start on
// 16-entry boundaries, so the loads are 16-byte vectors), with the group's
// descriptor: bit i of `mask` is set when entry i starts a segment, `s0` is
// the chunk segment holding entry 0, `v0` the voxel of entry 0 and `carry`
// the length (units of U) of entry 15's segment after the group.
struct Group {
  uint32_t w[GE];  // packed entries
  uint32_t mask, s0, v0;
  float carry;
  __device__ __forceinline__ void load(const IterArgs& a, uint64_t gi) {
    const uint4* pw = reinterpret_cast<const uint4*>(a.ent + gi * GE);
#pragma unroll
    for (int k = 0; k < GE / 4; ++k) {
      const uint4 v = pw[k];
      w[4 * k] = v.x; w[4 * k + 1] = v.y; w[4 * k + 2] = v.z; w[4 * k + 3] = v.w;
    }
    mask = a.gdesc[gi];  // raw until decode()
    carry = a.gcarry[gi];
    v0 = a.gv0[gi];
  }
  // Splits the descriptor word into the start mask and the first segment.
  // Kept apart from load() so that a pass can issue its other loads before
  // it waits for the group's.
  __device__ __forceinline__ void decode() {
    s0 = mask >> 16;
    mask &= 0xFFFFu;
  }
  __device__ __forceinline__ bool start(int i) const { return (mask >> i) & 1u; }
  // chunk segment of entry i: s0 plus the segment starts in 1..i
  __device__ __forceinline__ uint32_t seg(int i) const {
    return s0 + (uint32_t)__popc(mask & ((2u << i) - 2u));
  }
  // voxel (inside the brick) of every entry, handed to sink(i, v): entry 0
  // from the descriptor, a segment start from the chunk's first-voxel
  // table, else the previous voxel plus the step of the entry's code (LDS
  // table of 8 steps).  At the entries where a segment may start (one in
  // SEG_ALIGN) both reads are made and the choice is a bitwise select, so
  // the reads of a half group issue together and the chain is three VALU
  // ops per entry; elsewhere the chain is one add.  Halves are kept apart
  // (scheduling barrier) so that fewer reads are in flight per thread.
  template <class First, class Sink>
  __device__ __forceinline__ void voxels(const int* stab, First first, Sink sink) const {
    uint32_t v = v0;
    sink(0, v);
#pragma unroll
    for (int h = 0; h < GE; h += GE / 2) {
      uint32_t f[GE / 2];
      int d[GE / 2];
#pragma unroll
      for (int k = 0; k < GE / 2; ++k) {
        const int i = h + k;
        if (i == 0) continue;
        if (i % SEG_ALIGN == 0) f[k] = first(seg(i));
        d[k] = stab[ent_code(w[i])];
      }
#pragma unroll
      for (int k = 0; k < GE / 2; ++k) {
        const int i = h + k;
        if (i == 0) continue;
        if (i % SEG_ALIGN == 0)
          v = bit_select(bit_mask(mask, i), f[k], v + (uint32_t)d[k]);  // a start takes f
        else
          v += (uint32_t)d[k];  // no segment starts here
        sink(i, v);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  }
};

// the step table of the brick shape (code 7 unused)
__device__ __forceinline__ void fill_step_table(const IterArgs& a, int* stab) {
  if (threadIdx.x < 8) stab[threadIdx.x] = a.ec.step(threadIdx.x);
}

// Forward: per segment, sum_j lambda_j W_ij (three components) with
// lambda staged in LDS.  Entry-parallel: thread t takes the chunk's 16-entry
// group t and walks it backwards, so an entry's tail inside its segment
// grows by each length from the group's carry and restarts at segment
// starts.  Tails are measured from the segment's own end (at most a brick
// diagonal), so the FP32 terms carry no large offsets; the track kernel
// moves the partials to the track's basis in FP64.  Lengths and tails are
// in units of U (a power of two); the partial is scaled back exactly.
// A group splits into pieces at segment starts.  The piece that starts at
// entry i >= 1 (or at entry 0 when that starts a segment) is a segment's
// first piece: the thread stores its sums and start entry in the segment's
// LDS slot.  The group's head piece (entries 0.. up to its first start in
// 1..15) goes to the group's slot.  A segment is then its first piece plus
// the head pieces of the later groups it covers, which its own thread adds
// in FP64 and writes as the partial, with the segment's end tail
// (fixed-point bits) as its fourth word.  Plain LDS stores, no atomics: the
// pass is deterministic.  Two barriers per chunk.  (Piece sums restart at
// each start: running sums over the group would make a short piece the
// difference of two larger FP32 sums.)
// ACC = 1: FP64 piece sums of the FP32 entry terms (the products with the
//          tail weights in FP64);
// ACC = 0: FP32 two-sum (compensated) piece sums.
// A piece is stored rounded to FP32 (one rounding per piece; the partial is
// rounded once more), which keeps the LDS tables small.
template <int ACC, uint32_t BS>
__global__ void __launch_bounds__(BS)
k_fwd32(IterArgs a, const float* __restrict__ lam, float4* __restrict__ part) {
  // static LDS is a multiple of 16 bytes and the dynamic region is declared
  // as 16-byte aligned doubles, so the brick array starts 16-byte aligned
  extern __shared__ __align__(16) double lds_dyn[];
  float* lds_lam = reinterpret_cast<float*>(lds_dyn);
  __shared__ float4 spc[BS];  // per segment: first piece (x, y, z), start entry bits (w)
  __shared__ float4 hpc[BS];  // per group: head piece
  __shared__ uint32_t sv0[BS];
  __shared__ int stab[16];
  __shared__ uint32_t sh_[4];
  uint32_t& sh = sh_[0];
  const uint32_t t = threadIdx.x;
  fill_step_table(a, stab);
  const double U = a.ec.unit, U2 = U * U, U3 = U2 * U;
  for (;;) {
    if (t == 0) sh = atomicAdd(&a.work[0], 1u);
    __syncthreads();
    const uint32_t ti = sh;
    if (ti >= a.n_tasks) break;
    const Task tk = a.tasks[ti];
    load_brick(a.G, tk.brick, lam, lds_lam);  // (the first chunk's barrier orders it)
    Chunk nx = chunk_at(a, tk.c0);
    // a segment partial of the chunk before, stored once the next chunk's
    // loads are issued (so that the next chunk's address arithmetic does not
    // wait for the store to read its registers)
    float4 pend = make_float4(0.f, 0.f, 0.f, 0.f);
    uint64_t pend_q = ~0ull;
    for (uint32_t c = tk.c0; c < tk.c1; ++c) {
      // One memory round trip per chunk: the chunk's descriptor came with
      // the chunk before, and its segment words and groups load together.
      const Chunk ch = nx;
      if (c + 1 < tk.c1) nx = chunk_at(a, c + 1);
      const uint64_t q = seg_slot(a, ch, t);
      const uint32_t tfx = a.seg_tfx[q];
      const uint32_t v0w = a.seg_v0[q];
      const bool act = t * GE < ch.ne;
      Group gr;
      gr.load(a, group_slot(ch, t));
      if (pend_q != ~0ull) part[pend_q] = pend;
      pend_q = ~0ull;
      if (t < ch.cnt) sv0[t] = v0w;  // waits for the segment words only
      __syncthreads();
      if (act) {
        gr.decode();
        float lam_e[GE];
        gr.voxels(stab, [&](uint32_t q) { return sv0[q]; },
                  [&](int i, uint32_t v) { lam_e[i] = lds_lam[v]; });
        // copies the compiler cannot see through: the flush branch
        // recomputes segment indices and start entries from them instead of
        // keeping 32 values live from the voxel chain through the walk
        uint32_t e0 = t * GE, mk = gr.mask, sk = gr.s0, base_v = ENT_BASE;
        asm volatile("" : "+v"(e0), "+v"(mk), "+v"(sk), "+v"(base_v));
        float T = gr.carry;
        double d0 = 0.0, d1 = 0.0, d2 = 0.0;  // ACC 1
        float c0 = 0.f, c1 = 0.f, c2 = 0.f;   // ACC 0 high words
        float l0 = 0.f, l1 = 0.f, l2 = 0.f;   // ACC 0 low words
        auto piece = [&](uint32_t pos) {
          return ACC == 1 ? make_float4((float)d0, (float)d1, (float)d2, __uint_as_float(pos))
                          : make_float4((float)((double)c0 + (double)l0), (float)((double)c1 + (double)l1),
                                        (float)((double)c2 + (double)l2), __uint_as_float(pos));
        };
#pragma unroll
        for (int i = GE - 1; i >= 0; --i) {
          // W = [[L, L(L/2 + T)], [., L(L^2/3 + L T + T^2)]]
          const float L = ent_len_v(gr.w[i], base_v);
          const uint32_t st = start_mask(mk, i);  // entry i starts a segment
          const float av = lam_e[i] * L;
          const float h = fmaf(0.5f, L, T);
          const float k = fmaf(L, fmaf(L, 1.0f / 3.0f, T), T * T);
          if (ACC == 1) {
            const double ad = (double)av;
            d0 += ad;
            d1 = fma(ad, (double)h, d1);
            d2 = fma(ad, (double)k, d2);
          } else {
            two_sum_acc(c0, l0, av);
            two_sum_acc(c1, l1, av * h);
            two_sum_acc(c2, l2, av * k);
          }
          if (i > 0 && st != 0u) {  // entry i starts segment seg(i)
            spc[sk + (uint32_t)__popc(mk & ((2u << i) - 2u))] = piece(e0 + i);
            d0 = d1 = d2 = 0.0;
            c0 = c1 = c2 = 0.f;
            l0 = l1 = l2 = 0.f;
          }
          T = __uint_as_float(bit_select(st, 0u, __float_as_uint(T + L)));
        }
        const float4 head = piece(e0);
        hpc[t] = head;
        if (gr.mask & 1u) spc[gr.s0] = head;
      }
      __syncthreads();
      if (t < ch.cnt) {
        const float4 p = spc[t];
        const uint32_t A = __float_as_uint(p.w);
        const uint32_t B = t + 1 < ch.cnt ? __float_as_uint(spc[t + 1].w) : ch.ne;
        double x = p.x, y = p.y, z = p.z;
        const uint32_t g_last = min((B - 1) / GE, (ch.ne - 1) / GE);  // (bounded by the chunk)
        for (uint32_t g = A / GE + 1; g <= g_last; ++g) {
          const float4 hp = hpc[g];
          x += hp.x;
          y += hp.y;
          z += hp.z;
        }
        pend = make_float4((float)(x * U), (float)(y * U2), (float)(z * U3), __uint_as_float(tfx));
        pend_q = ch.q0 + t;
      }
    }
    if (pend_q != ~0ull) part[pend_q] = pend;
  }
}


// Per track (rank order): Sigma from segment partials, 2x2 algebra in FP64.
// A segment's partial P_q is summed in its local basis u(t) = (1, t - tau_q)
// (tau_q: tail after the segment); in a basis at t_c it becomes
// A P_q A^T with A = [[1, 0], [tau_q - t_c, 1]].  The track's basis t_c is
// its scattering centroid for the current image, rounded to the tail
// fixed-point grid so that tau_q - t_c is exact; there Sigma is nearly
// diagonal and the determinant and G keep their accuracy where Sigma itself
// is strongly correlated.  With M = [[1, 0], [t_c, 1]], Sigma = M Sigma' M^T:
// E and D enter as M^-1 E M^-T and M^-1 D M^-T; det, cost and <W, G> are
// basis-free.  Written per track: G' (FP32) and t_c (fixed-point bits).
__global__ void __launch_bounds__(256)
k_track32(size_t n_trk, const uint32_t* __restrict__ tseg_ptr,
          const uint32_t* __restrict__ tseg_pos,
          const float4* __restrict__ part, const double* __restrict__ trk_d,
          ErrM E, double tail_scale, double lmax, double ubound, int grad_mode,
          float* __restrict__ trk_g, double4* __restrict__ trk_g64, unsigned long long* __restrict__ cmax,
          const uint16_t* __restrict__ trk_scan,
          double* __restrict__ cost_acc, unsigned long long* __restrict__ cnt) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double cost = 0.0, bnd = 0.0;
  bool sing = false;
  if (r < n_trk) {
    // this track's segment partials: brick-major positions in track order;
    // neighbouring tracks (adjacent ranks) read neighbouring positions.
    // Moments about the track end (t = 0) in FP64, exact up to FP64 rounding.
    const double inv_scale = 1.0 / tail_scale;
    double S00 = 0.0, S01 = 0.0, S11 = 0.0;
    uint32_t tmin = 0xFFFFFFFFu, tmax = 0u;
    auto add = [&](const float4 pq) {
      const uint32_t tb = __float_as_uint(pq.w);
      tmin = min(tmin, tb);
      tmax = max(tmax, tb);
      const double tau = (double)tb * inv_scale;
      const double p0 = pq.x, p1 = pq.y, p2 = pq.z;
      S00 += p0;
      S01 += fma(tau, p0, p1);
      S11 += fma(tau, fma(tau, p0, 2.0 * p1), p2);
    };
    const uint32_t k0 = tseg_ptr[r], k1 = tseg_ptr[r + 1];
    uint32_t k = k0;
    for (; k + 4 <= k1; k += 4) {
      const uint32_t qa = tseg_pos[k], qb = tseg_pos[k + 1], qc = tseg_pos[k + 2], qd = tseg_pos[k + 3];
      const float4 pa = part[qa], pb = part[qb], pc = part[qc], pd = part[qd];
      add(pa);
      add(pb);
      add(pc);
      add(pd);
    }
    for (; k < k1; ++k) add(part[tseg_pos[k]]);
    // basis at the centroid, on the fixed-point grid (clamped to its range:
    // any basis is exact, the centroid is only the best-conditioned one)
    double xc = S00 > 0.0 ? S01 / S00 * tail_scale : 0.0;
    xc = xc >= 0.0 ? fmin(xc, 4294967295.0) : 0.0;
    const uint32_t tcx = (uint32_t)llrint(xc);
    const double tc = (double)tcx * inv_scale;
    const double s00 = S00, s01 = S01 - tc * S00;
    const double s11 = S11 - tc * (2.0 * S01 - tc * S00);
    const double D00 = trk_d[4 * r], D01 = trk_d[4 * r + 1];
    const double D11 = trk_d[4 * r + 2], p = trk_d[4 * r + 3];
    // shifted-basis measurement and error matrices
    const double d00 = D00, d01 = D01 - tc * D00, d11 = D11 - 2.0 * tc * D01 + tc * tc * D00;
    const double e00 = E.e00, e01 = E.e01 - tc * E.e00;
    const double e11 = E.e11 - 2.0 * tc * E.e01 + tc * tc * E.e00;
    const double pp = p * p;
    const double m00 = pp * s00 + e00, m01 = pp * s01 + e01, m11 = pp * s11 + e11;
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
      cost = log(det) + (d00 * i00 + d11 * i11 + d01 * 2.0 * i01) * 0.5;
    }
    // G' = 1/2 Si'^-1 D' Si'^-1 - Si'^-1  (symmetric)
    const double A00 = i00 * d00 + i01 * d01, A01 = i00 * d01 + i01 * d11;
    const double A10 = i01 * d00 + i11 * d01, A11 = i01 * d01 + i11 * d11;
    const double M00 = A00 * i00 + A01 * i01;
    const double M01 = A00 * i01 + A01 * i11;
    const double M11 = A10 * i01 + A11 * i11;
    float4 g;
    g.x = (float)(pp * (0.5 * M00 - i00));
    g.y = (float)(pp * (0.5 * M01 - i01));
    g.z = (float)(pp * (0.5 * M11 - i11));
    if (grad_mode && sing) g.x = g.y = g.z = 0.0f;  // cost_i = 0: no gradient
    g.w = __uint_as_float(tcx);
    *reinterpret_cast<float4*>(trk_g + 4 * r) = g;
    if (trk_g64) {  // the order statistics' backprojection: G' before its rounding to FP32
      const bool zero = grad_mode && sing;
      trk_g64[r] = make_double4(zero ? 0.0 : pp * (0.5 * M00 - i00), zero ? 0.0 : pp * (0.5 * M01 - i01),
                                zero ? 0.0 : pp * (0.5 * M11 - i11), (double)tcx);
    }
    // bound on |c_ij| over the track's entries for the backprojection's
    // fixed-point sums: an entry lies at most R = max|tau_q - t_c| + ubound
    // from t_c and is at most lmax long, so |<W, G'>| <= lmax (|g00| +
    // 2 R |g01| + R^2 |g11|)
    if (k1 > k0) {
      const double S = fmax((double)tmax * inv_scale - tc, tc - (double)tmin * inv_scale);
      const double R = fmax(S, 0.0) + ubound;
      bnd = lmax * (fabs((double)g.x) + 2.0 * R * fabs((double)g.y) + R * R * fabs((double)g.z)) * 1.001;
    }
  }
  // The |c| bound (a maximum) and the cost per scan, combined over the
  // block first: one atomic per block on each address instead of one per
  // wavefront (same-address atomics serialise).
  __shared__ double red_m[16], red_c[16];
  __shared__ uint32_t red_s[16];
  const uint32_t w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  {
    double m = bnd;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) red_m[w] = m;
  }
  // cost per scan: summed per wavefront when its tracks share a scan, else
  // added per track
  const uint32_t sc = (trk_scan && r < n_trk) ? trk_scan[r] : 0u;
  const uint32_t s0 = __shfl(sc, 0, 64);
  const bool same = __all(sc == s0);
  if (same) {
    double v = cost;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) { red_c[w] = v; red_s[w] = s0; }
  } else {
    if (cost != 0.0) atomicAdd(&cost_acc[sc], cost);
    if ((threadIdx.x & 63) == 0) { red_c[w] = 0.0; red_s[w] = 0xFFFFFFFFu; }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (uint32_t k = 0; k < nw; ++k) m = fmax(m, red_m[k]);
    if (m > 0.0) atomicMax(cmax, (unsigned long long)__double_as_longlong(m));
    // waves of one scan in a row are added together
    uint32_t cs = 0xFFFFFFFFu;
    double cv = 0.0;
    for (uint32_t k = 0; k < nw; ++k) {
      if (red_s[k] == 0xFFFFFFFFu) continue;
      if (red_s[k] != cs) {
        if (cs != 0xFFFFFFFFu && cv != 0.0) atomicAdd(&cost_acc[cs], cv);
        cs = red_s[k];
        cv = 0.0;
      }
      cv += red_c[k];
    }
    if (cs != 0xFFFFFFFFu && cv != 0.0) atomicAdd(&cost_acc[cs], cv);
  }
  unsigned long long m = __ballot(sing);
  if (m && (threadIdx.x & 63) == (unsigned)(__ffsll((long long)m) - 1))
    atomicAdd(&cnt[C_SINGULAR], (unsigned long long)__popcll(m));
}

// Backprojection: c_ij = <W_ij, G_i> accumulated per voxel in LDS, then
// flushed once per task (stored when the task owns the brick, else added).
// One thread per segment first moves G' from the track's basis to the
// segment's local one (FP64) into an LDS table; then entry-parallel like the
// forward pass.  Two barriers per chunk (one table keeps the LDS small).
// SUB chunks of CW threads each per workgroup (CW the plan's block size):
// with SUB = 2 each half of a workgroup of 2 CW threads takes one of two
// consecutive chunks of the task, both adding into the brick's sums, so
// that the brick array is held once for twice the waves.
// ACC = 1: one int64 fixed-point LDS atomic per entry;
// ACC = 0: FP32 atomic two-sum -- the returned old value o and the stored
//          o + c give the exact rounding error, added into a low word.
template <int ACC, uint32_t CW, uint32_t SUB>
__global__ void __launch_bounds__(CW * SUB)
k_bwd32(IterArgs a, const float* __restrict__ trk_g,
        const unsigned long long* __restrict__ cmax, double* __restrict__ acc) {
  constexpr uint32_t TAB = chunk_seg_cap(CW);  // G' entries per chunk
  // static LDS is a multiple of 16 bytes and the dynamic region is declared
  // as 16-byte aligned doubles: ds_add_f64 needs 8-byte aligned addresses
  extern __shared__ __align__(16) double lds_dyn[];
  float* lds_acc = reinterpret_cast<float*>(lds_dyn);
  __shared__ float4 sg[SUB * TAB];
  __shared__ int stab[8];
  __shared__ uint32_t sh_[4];
  uint32_t& sh = sh_[0];
  const uint32_t t = threadIdx.x;
  const uint32_t sub = SUB == 1 ? 0u : t / CW;  // this thread's chunk of the SUB (whole wavefronts)
  const uint32_t tl = SUB == 1 ? t : t % CW;    // and its thread in that chunk
  fill_step_table(a, stab);
  const double U = a.ec.unit, U2 = U * U, U3 = U2 * U;
  const uint32_t vol = a.G.b[0] * a.G.b[1] * a.G.b[2];
  float* hi_s = lds_acc;
  float* lo_s = lds_acc + vol;
  long long* d_s = reinterpret_cast<long long*>(lds_dyn);
  // fixed-point scale of the sums: every |c_ij| is below the bound the
  // track kernel found, so a scaled entry is at most 2^fx_bits, and the
  // M_j entries of a voxel sum inside int64
  const double fx = fx_scale(cmax, a.fx_bits);
  for (;;) {
    if (t == 0) sh = atomicAdd(&a.work[1], 1u);
    for (uint32_t v = t; v < 2 * vol; v += blockDim.x) lds_acc[v] = 0.0f;
    __syncthreads();
    const uint32_t ti = sh;
    if (ti >= a.n_tasks) break;
    const Task tk = a.tasks[ti];
    for (uint32_t c0 = tk.c0; c0 < tk.c1; c0 += SUB) {
      // (the chunk's descriptor, then its segment words and the G' gather,
      // then the stream's groups: measured faster here than issuing the
      // groups first or the descriptor a chunk ahead)
      Chunk ch{};  // no chunk (cnt = ne = 0) past the task's last
      if (c0 + sub < tk.c1) ch = a.chunks[c0 + sub];
      float4* sgb = sg + sub * TAB;
      if (tl < ch.cnt) {
        const uint64_t q = ch.q0 + tl;
        const uint32_t tfx = a.seg_tfx[q];
        const float4 g = *reinterpret_cast<const float4*>(trk_g + 4 * (size_t)a.seg_trk[q]);
        // A^T G' A with A = [[1, 0], [s, 1]], s = tau_q - t_c (exact)
        const double shf = (double)((int64_t)tfx - (int64_t)__float_as_uint(g.w)) * a.inv_scale;
        const double gy = fma(shf, (double)g.z, (double)g.y);
        // in units of U for the entry terms: g00 U, g01 U^2, g11 U^3
        sgb[tl] = make_float4((float)(fma(shf, gy + (double)g.y, (double)g.x) * U), (float)(gy * U2),
                              (float)((double)g.z * U3), __uint_as_float((uint32_t)a.seg_v0[q]));
      }
      const bool act = tl * GE < ch.ne;
      Group gr;
      if (act) gr.load(a, (ch.base / GE) + tl);
      __syncthreads();
      if (act) {
        gr.decode();
        uint32_t vx[GE];
        gr.voxels(stab, [&](uint32_t q) { return __float_as_uint(sgb[q].w); },
                  [&](int i, uint32_t v) { vx[i] = v; });
        float T = gr.carry;
        uint32_t mk = gr.mask, sk = gr.s0, base_v = ENT_BASE;  // opaque copies, as in the forward walk
        asm volatile("" : "+v"(mk), "+v"(sk), "+v"(base_v));
        float4 g = sgb[sk + (uint32_t)__popc(mk & 0xFFFEu)];
#pragma unroll
        for (int i = GE - 1; i >= 0; --i) {
          // <W, G> = L (g00 + (L + 2T) g01 + (L (L/3 + T) + T^2) g11)
          const float L = ent_len_v(gr.w[i], base_v);
          const uint32_t st = start_mask(mk, i);  // entry i starts a segment
          const float k = fmaf(L, fmaf(L, 1.0f / 3.0f, T), T * T);
          const float cv = L * fmaf(k, g.z, fmaf(fmaf(2.0f, T, L), g.y, g.x));
          const uint32_t v = vx[i];
          if (ACC == 1) {
            // round(cv * fx) as int64: adding 1.5 * 2^52 leaves the integer
            // in the low mantissa bits (|cv * fx| < 2^51)
            const double m = fma((double)cv, fx, 6755399441055744.0);
            const long long q = __double_as_longlong(m) - 0x4338000000000000LL;
            atomicAdd(reinterpret_cast<unsigned long long*>(&d_s[v]), (unsigned long long)q);
          } else {
            const float old = atomicAdd(&hi_s[v], cv);
            const float sum = old + cv;
            const float bp = sum - old;
            const float e = (old - (sum - bp)) + (cv - bp);
            if (e != 0.0f) atomicAdd(&lo_s[v], e);
          }
          // entry i - 1 lies in the segment before when entry i starts one
          if (i > 0 && st != 0u) g = sgb[sk + (uint32_t)__popc(mk & ((2u << i) - 2u)) - 1u];
          T = __uint_as_float(bit_select(st, 0u, __float_as_uint(T + L)));
        }
      }
      __syncthreads();  // the table is rewritten for the next chunk
    }
    __syncthreads();
    uint32_t o[3];
    brick_origin(a.G, tk.brick, o);
    // Flush.  ACC 1: the sums stay int64 fixed point in the accumulator
    // words; a task that covers its brick stores them, tasks sharing a brick
    // add them (exact, so the result does not depend on the order the tasks
    // ran in) and the voxel update turns them into FP64.  ACC 0: FP64 sums,
    // stored or added.
    const bool shared = tk.ntask > 1;
    long long* acc_fx = reinterpret_cast<long long*>(acc);
    for (uint32_t v = t; v < vol; v += blockDim.x) {
      const uint32_t rr = a.G.db0.div(v), lx = v - rr * a.G.b[0];
      const uint32_t lz = a.G.db1.div(rr), ly = rr - lz * a.G.b[1];
      const uint32_t x = o[0] + lx, y = o[1] + ly, z = o[2] + lz;
      if (x >= a.G.n[0] || y >= a.G.n[1] || z >= a.G.n[2]) continue;
      const size_t j = (size_t)x + (size_t)a.G.n[0] * ((size_t)y + (size_t)a.G.n[1] * z);
      if (ACC == 1) {
        if (!shared)
          acc_fx[j] = d_s[v];
        else if (d_s[v] != 0)
          atomicAdd(reinterpret_cast<unsigned long long*>(&acc_fx[j]), (unsigned long long)d_s[v]);
      } else {
        const double val = (double)hi_s[v] + (double)lo_s[v];
        if (!shared)
          acc[j] = val;
        else if (val != 0.0)
          atomicAdd(&acc[j], val);
      }
    }
    __syncthreads();
  }
}

// ---------------- FP64 stream iteration --------------------------------
// The same brick stream with the reference's exact FP64 lengths (len64) and
// tails: each group stores the tail of its entry 15 (gtail64) and each
// segment the tail of its last entry (seg_tail64), both taken from the
// tail array of src/MutoMLSD.cpp:403-414, and the walk runs that array's
// own backward recurrence T_{k-1} = T_k + L_k inside a segment.  So every
// W_ij (src/MutoMLSD.cpp:417-430) is bit-identical to the reference's; the
// sums over a track differ from its running order only in association.
// FP32C and FP64 stream medians by brackets (k_bwd32m, k_bwd64s modes 2-3,
// k_sel_brk): the arrays of one pass
struct MedArgs {
  const Task* btask;   // one task per brick, largest first
  uint32_t n_btask;
  unsigned int* work;  // task counter of the pass
  float2* brk;         // [nv] bracket (lo, hi)
  float* bmed;         // [nv] last median value (upper middle value; NaN: none yet)
  uint32_t* bcnt;      // [nv] values below | values inside << 16
  float* buf;          // [nv * MED_K] values inside (FP32C)
  double* buf64;       // ... (FP64 stream)
  uint8_t* miss;       // [nv] the median left the bracket
  uint8_t* bmiss;      // [bricks] the brick holds a miss
  const uint64_t* col; // [nv + 1] slot ranges (misses' whole lists)
  float* cval;         // (FP32C)
  double* cval64;      // (FP64 stream)
  unsigned long long* nmiss;  // misses of the pass (a plan counter)
};

struct Args64 {
  const double* len64;
  const double* gtail64;
  const double* seg_tail64;
  const uint32_t* slot;  // median / trimmed mean: slot of each entry
  double* cval;          // ... and the per-entry c_ij, voxel-major
  // the order statistics' backprojection takes its tasks from one queue per
  // XCD (workgroup b on queue b % nq): small tasks of one brick at a time,
  // so that the brick's slot range stays in that XCD's L2 while its
  // scattered c_ij writes land
  const Task* qtask;
  const uint32_t* qoff;  // [nq + 1]
  unsigned int* qwork;   // [nq]
  uint32_t nq;
};

constexpr size_t fwd64_static_lds(uint32_t bs) {
  return 2 * (size_t)bs * 32 + (size_t)bs * 4 + (size_t)bs * 8 + 64 + 16;
}
constexpr size_t bwd64_static_lds(uint32_t bs) { return (size_t)bs * 32 + (size_t)bs * 4 + 32 + 16; }

// FP64 lambda of the brick into LDS (zero outside the grid)
__device__ __forceinline__ void load_brick64(const BrickGeo& G, uint32_t brick,
                                             const double* __restrict__ lam, double* lds) {
  uint32_t o[3];
  brick_origin(G, brick, o);
  const uint32_t vol = G.b[0] * G.b[1] * G.b[2];
  for (uint32_t v = threadIdx.x; v < vol; v += blockDim.x) {
    const uint32_t r = G.db0.div(v), lx = v - r * G.b[0];
    const uint32_t lz = G.db1.div(r), ly = r - lz * G.b[1];
    const uint32_t x = o[0] + lx, y = o[1] + ly, z = o[2] + lz;
    double val = 0.0;
    if (x < G.n[0] && y < G.n[1] && z < G.n[2])
      val = lam[(size_t)x + (size_t)G.n[0] * ((size_t)y + (size_t)G.n[1] * z)];
    lds[v] = val;
  }
}

// Forward, FP64: per segment, sum_j lambda_j W_ij over its entries (the
// three components), in the track's own tail basis, so no basis shift is
// needed.  Entry-parallel as k_fwd32: thread t walks the chunk's group t
// backwards in two halves of 8 entries (lengths and lambdas loaded per
// half), pieces split at segment starts and put together per segment by
// the segment's own thread.
template <uint32_t BS>
__global__ void __launch_bounds__(BS)
k_fwd64s(IterArgs a, Args64 b, const double* __restrict__ lam, double4* __restrict__ part) {
  extern __shared__ __align__(16) double lds_dyn[];
  double* lds_lam = lds_dyn;
  __shared__ double4 spc[BS];  // per segment: first piece (x, y, z), start entry (w)
  __shared__ double4 hpc[BS];  // per group: head piece
  __shared__ uint32_t sv0[BS];
  __shared__ double stl[BS];   // per segment: tail of its last entry
  __shared__ int stab[16];
  __shared__ uint32_t sh_[4];
  uint32_t& sh = sh_[0];
  const uint32_t t = threadIdx.x;
  fill_step_table(a, stab);
  for (;;) {
    if (t == 0) sh = atomicAdd(&a.work[0], 1u);
    __syncthreads();
    const uint32_t ti = sh;
    if (ti >= a.n_tasks) break;
    const Task tk = a.tasks[ti];
    load_brick64(a.G, tk.brick, lam, lds_lam);  // (the first chunk's barrier orders it)
    Chunk nx = chunk_at(a, tk.c0);
    // as in k_fwd32: one memory round trip per chunk (the descriptor a chunk
    // ahead; segment words, groups, the group's exact tail and lengths issued
    // together), the partials stored after the next chunk's loads
    double4 pend = make_double4(0.0, 0.0, 0.0, 0.0);
    uint64_t pend_q = ~0ull;
    for (uint32_t c = tk.c0; c < tk.c1; ++c) {
      const Chunk ch = nx;
      if (c + 1 < tk.c1) nx = chunk_at(a, c + 1);
      const uint64_t q = seg_slot(a, ch, t);
      const uint32_t v0w = a.seg_v0[q];
      const double tlw = b.seg_tail64[q];
      const bool act = t * GE < ch.ne;
      const uint64_t gi = group_slot(ch, t);
      Group gr;
      gr.load(a, gi);
      double T = b.gtail64[gi];
      double L16[GE];
      {
        const double2* pl2 = reinterpret_cast<const double2*>(b.len64 + gi * GE);
#pragma unroll
        for (int k = 0; k < GE / 2; ++k) {
          const double2 v2 = pl2[k];
          L16[2 * k] = v2.x;
          L16[2 * k + 1] = v2.y;
        }
      }
      if (pend_q != ~0ull) part[pend_q] = pend;
      pend_q = ~0ull;
      if (t < ch.cnt) {  // waits for the segment words only
        sv0[t] = v0w;
        stl[t] = tlw;
      }
      __syncthreads();
      if (act) {
        gr.decode();
        uint32_t vx[GE];
        gr.voxels(stab, [&](uint32_t q) { return sv0[q]; }, [&](int i, uint32_t v) { vx[i] = v; });
        uint32_t e0 = t * GE, mk = gr.mask, sk = gr.s0;
        asm volatile("" : "+v"(e0), "+v"(mk), "+v"(sk));
        double d0 = 0.0, d1 = 0.0, d2 = 0.0;
#pragma unroll
        for (int h = GE / 2; h >= 0; h -= GE / 2) {
          double l8[GE / 2];
#pragma unroll
          for (int k = 0; k < GE / 2; ++k) l8[k] = lds_lam[vx[h + k]];
#pragma unroll
          for (int k = GE / 2 - 1; k >= 0; --k) {
            const int i = h + k;
            const double L = L16[i], lj = l8[k];
            // generateWeightMatrix (MutoMLSD.cpp:417-430), the reference's own expressions
            const double LT = L * T;
            const double w01 = LT + L * L * 0.5;
            const double w11 = L * L * L / 3.0 + (L + T) * LT;
            d0 = d0 + lj * L;
            d1 = d1 + lj * w01;
            d2 = d2 + lj * w11;
            if (i > 0 && i % SEG_ALIGN == 0 && ((mk >> i) & 1u)) {  // entry i starts segment seg(i)
              const uint32_t q = sk + (uint32_t)__popc(mk & ((2u << i) - 2u));
              spc[q] = make_double4(d0, d1, d2, (double)(e0 + i));
              d0 = d1 = d2 = 0.0;
              T = stl[q - 1];  // entry i - 1 is the previous segment's last
            } else {
              T = T + L;
            }
          }
          __builtin_amdgcn_sched_barrier(0);
        }
        const double4 head = make_double4(d0, d1, d2, (double)e0);
        hpc[t] = head;
        if (gr.mask & 1u) spc[gr.s0] = head;
      }
      __syncthreads();
      if (t < ch.cnt) {
        const double4 p = spc[t];
        const uint32_t A = (uint32_t)p.w;
        const uint32_t B = t + 1 < ch.cnt ? (uint32_t)spc[t + 1].w : ch.ne;
        double x = p.x, y = p.y, z = p.z;
        const uint32_t g_last = min((B - 1) / GE, (ch.ne - 1) / GE);
        for (uint32_t g = A / GE + 1; g <= g_last; ++g) {
          const double4 hp = hpc[g];
          x = x + hp.x;
          y = y + hp.y;
          z = z + hp.z;
        }
        pend = make_double4(x, y, z, 0.0);
        pend_q = ch.q0 + t;
      }
    }
    if (pend_q != ~0ull) part[pend_q] = pend;
  }
}

// Per track (rank order), FP64: Sigma from the segment partials, Eigen's
// 2x2 det/inverse (threshold 1e-12), cost_i (src/MutoMLSD.cpp:442-468) and
// p^2 G with G = 1/2 Sigma^-1 D Sigma^-1 - Sigma^-1, so that c_ij =
// <W_ij, p^2 G> and S_ij = 2 lambda_j + c_ij lambda_j^2 (the identity
// 1/2 xy - tr = <W_ij, G> of the reference's :471-483 expressions).  With
// grad_mode, G = 0 where cost_i = 0 (the MAP objective's convention);
// otherwise G follows the computed inverse as the reference's S_ij does.
__global__ void __launch_bounds__(256)
k_track64s(size_t n_trk, const uint32_t* __restrict__ tseg_ptr, const uint32_t* __restrict__ tseg_pos,
           const double4* __restrict__ part, const double* __restrict__ trk_d, ErrM E, int grad_mode,
           double4* __restrict__ trk_g, double* __restrict__ cost_part,
           unsigned long long* __restrict__ cnt) {
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double cost = 0.0;
  bool sing = false;
  if (r < n_trk) {
    double S00 = 0.0, S01 = 0.0, S11 = 0.0;
    const uint32_t k0 = tseg_ptr[r], k1 = tseg_ptr[r + 1];
    uint32_t k = k0;
    for (; k + 4 <= k1; k += 4) {
      const double4 pa = part[tseg_pos[k]], pb = part[tseg_pos[k + 1]];
      const double4 pc = part[tseg_pos[k + 2]], pd = part[tseg_pos[k + 3]];
      S00 = S00 + pa.x; S01 = S01 + pa.y; S11 = S11 + pa.z;
      S00 = S00 + pb.x; S01 = S01 + pb.y; S11 = S11 + pb.z;
      S00 = S00 + pc.x; S01 = S01 + pc.y; S11 = S11 + pc.z;
      S00 = S00 + pd.x; S01 = S01 + pd.y; S11 = S11 + pd.z;
    }
    for (; k < k1; ++k) {
      const double4 pa = part[tseg_pos[k]];
      S00 = S00 + pa.x; S01 = S01 + pa.y; S11 = S11 + pa.z;
    }
    const double d00 = trk_d[4 * r], d01 = trk_d[4 * r + 1];
    const double d11 = trk_d[4 * r + 2], p = trk_d[4 * r + 3];
    const double pp = p * p;
    Mat2 cov;
    cov.m00 = pp * S00 + E.e00;
    cov.m01 = pp * S01 + E.e01;
    cov.m10 = pp * S01 + E.e01;
    cov.m11 = pp * S11 + E.e11;
    Mat2 inv;
    double det;
    const bool ok = inv_det(cov, inv, det);
    if (!ok || det <= 0) {
      sing = true;
    } else {
      cost = log(det) + (d00 * inv.m00 + d11 * inv.m11 + d01 * 2.0 * inv.m01) * 0.5;
    }
    const double A00 = inv.m00 * d00 + inv.m01 * d01, A01 = inv.m00 * d01 + inv.m01 * d11;
    const double A10 = inv.m10 * d00 + inv.m11 * d01, A11 = inv.m10 * d01 + inv.m11 * d11;
    double4 g;
    g.x = pp * (0.5 * (A00 * inv.m00 + A01 * inv.m10) - inv.m00);
    g.y = pp * (0.5 * (A00 * inv.m01 + A01 * inv.m11) - inv.m01);
    g.z = pp * (0.5 * (A10 * inv.m01 + A11 * inv.m11) - inv.m11);
    g.w = 0.0;
    if (grad_mode && sing) g.x = g.y = g.z = 0.0;
    trk_g[r] = g;
  }
  block_sum_to(cost_part, cost);
  wave_count(&cnt[C_SINGULAR], sing);
}

// Backprojection, FP64: c_ij = <W_ij, p^2 G_i> per entry, from the same
// exact lengths and tails as the forward walk.  MODE 0 (mean): added per
// voxel in LDS (FP64 atomics), flushed once per task (stored when the task
// covers its brick, else added); MODE 1 (median, trimmed mean): stored in
// the entry's voxel-major slot.
template <int MODE, uint32_t BS>
__global__ void __launch_bounds__(BS)
k_bwd64s(IterArgs a, Args64 b, MedArgs md, const double4* __restrict__ trk_g, double* __restrict__ acc) {
  // MODE 2: the median's bracket pass, MODE 3: its bricks with misses (as
  // k_bwd32m, on the exact FP64 values)
  constexpr bool MED = MODE >= 2;
  extern __shared__ __align__(16) double lds_dyn[];
  double* acc_s = lds_dyn;
  float2* brk_s = reinterpret_cast<float2*>(lds_dyn);
  __shared__ double4 sg[BS];  // per segment: p^2 G of its track (x, y, z), tail of its last entry (w)
  __shared__ uint32_t sv0[BS];
  __shared__ int stab[8];
  __shared__ uint32_t sh_[4];
  uint32_t& sh = sh_[0];
  const uint32_t t = threadIdx.x;
  fill_step_table(a, stab);
  const uint32_t vol = a.G.b[0] * a.G.b[1] * a.G.b[2];
  uint32_t* cnt_s = reinterpret_cast<uint32_t*>(brk_s + vol);
  const uint32_t qi = MODE == 1 ? blockIdx.x % b.nq : 0u;
  for (;;) {
    if (t == 0)
      sh = MED ? atomicAdd(md.work, 1u) : MODE == 1 ? b.qoff[qi] + atomicAdd(&b.qwork[qi], 1u) : atomicAdd(&a.work[1], 1u);
    if (MODE == 0)
      for (uint32_t v = t; v < vol; v += BS) acc_s[v] = 0.0;
    __syncthreads();
    const uint32_t ti = sh;
    if (MED ? ti >= md.n_btask : MODE == 1 ? ti >= b.qoff[qi + 1] : ti >= a.n_tasks) break;
    const Task tk = MED ? md.btask[ti] : MODE == 1 ? b.qtask[ti] : a.tasks[ti];
    if (MODE == 3 && !md.bmiss[tk.brick]) {
      __syncthreads();  // every thread has read sh before it is rewritten
      continue;
    }
    uint32_t o[3];
    brick_origin(a.G, tk.brick, o);
    if (MED) {
      for (uint32_t v = t; v < vol; v += BS) {
        const uint32_t rr = a.G.db0.div(v), lx = v - rr * a.G.b[0];
        const uint32_t lz = a.G.db1.div(rr), ly = rr - lz * a.G.b[1];
        const uint32_t x = o[0] + lx, y = o[1] + ly, z = o[2] + lz;
        float2 br = make_float2(INFINITY, -INFINITY);
        if (x < a.G.n[0] && y < a.G.n[1] && z < a.G.n[2]) {
          const size_t j = (size_t)x + (size_t)a.G.n[0] * ((size_t)y + (size_t)a.G.n[1] * z);
          if (MODE == 3)
            br = md.miss[j] ? make_float2(-INFINITY, INFINITY) : make_float2(INFINITY, -INFINITY);
          else
            br = md.brk[j];
        }
        brk_s[v] = br;
        cnt_s[v] = 0u;
      }
    }
    for (uint32_t c = tk.c0; c < tk.c1; ++c) {
      const Chunk ch = a.chunks[c];
      if (t < ch.cnt) {
        const uint64_t q = ch.q0 + t;
        const double4 g = trk_g[a.seg_trk[q]];
        sg[t] = make_double4(g.x, g.y, g.z, b.seg_tail64[q]);
        sv0[t] = a.seg_v0[q];
      }
      // MODE 1: the group, its exact tail, lengths and slots issued together
      // before the barrier, after the G' gathers above (25.1 -> 23.0 ms at
      // the benchmark size); MODE 0 measured faster loading the tail and
      // lengths in the walk (5.34 against 5.64 ms)
      constexpr bool PRE = MODE == 1;
      const bool act = t * GE < ch.ne;
      const uint64_t gi = PRE ? group_slot(ch, t) : (ch.base / GE) + t;
      Group gr;
      if (PRE || act) gr.load(a, gi);
      double T = 0.0;
      double L16[GE];
      uint32_t sl[GE];
      if (PRE) {
        T = b.gtail64[gi];
        const double2* pl2 = reinterpret_cast<const double2*>(b.len64 + gi * GE);
#pragma unroll
        for (int k = 0; k < GE / 2; ++k) {
          const double2 v2 = pl2[k];
          L16[2 * k] = v2.x;
          L16[2 * k + 1] = v2.y;
        }
        const uint4* ps = reinterpret_cast<const uint4*>(b.slot + gi * GE);
#pragma unroll
        for (int k = 0; k < GE / 4; ++k) {
          const uint4 v4 = ps[k];
          sl[4 * k] = v4.x; sl[4 * k + 1] = v4.y; sl[4 * k + 2] = v4.z; sl[4 * k + 3] = v4.w;
        }
      }
      __syncthreads();
      if (act) {
        gr.decode();
        uint32_t vx[GE];
        gr.voxels(stab, [&](uint32_t q) { return sv0[q]; }, [&](int i, uint32_t v) { vx[i] = v; });
        uint32_t mk = gr.mask, sk = gr.s0;
        asm volatile("" : "+v"(mk), "+v"(sk));
        if (!PRE) T = b.gtail64[gi];
        double4 g = sg[sk + (uint32_t)__popc(mk & 0xFFFEu)];
        const double2* pl2 = reinterpret_cast<const double2*>(b.len64 + gi * GE);
#pragma unroll
        for (int h = GE / 2; h >= 0; h -= GE / 2) {
          if (!PRE) {
#pragma unroll
            for (int k = 0; k < GE / 4; ++k) {
              const double2 v2 = pl2[h / 2 + k];
              L16[h + 2 * k] = v2.x;
              L16[h + 2 * k + 1] = v2.y;
            }
          }
#pragma unroll
          for (int k = GE / 2 - 1; k >= 0; --k) {
            const int i = h + k;
            const double L = L16[i];
            const double LT = L * T;
            const double w01 = LT + L * L * 0.5;
            const double w11 = L * L * L / 3.0 + (L + T) * LT;
            const double cv = L * g.x + 2.0 * w01 * g.y + w11 * g.z;
            if (MODE == 0) {
              atomicAdd(&acc_s[vx[i]], cv);
            } else if (MODE == 1) {
              if (sl[i] != 0xFFFFFFFFu) b.cval[sl[i]] = cv;
            } else if (t * GE + (uint32_t)i < ch.ne && ent_code(gr.w[i]) != ENT_PAD_CODE) {
              // the bracket's FP32 bounds against the exact value
              const uint32_t v = vx[i];
              const float2 br = brk_s[v];
              if (cv < (double)br.x) {
                atomicAdd(&cnt_s[v], 1u);
              } else if (!(cv > (double)br.y)) {
                const uint32_t pos = atomicAdd(&cnt_s[v], 0x10000u) >> 16;
                const uint32_t rr = a.G.db0.div(v), lx = v - rr * a.G.b[0];
                const uint32_t lz = a.G.db1.div(rr), ly = rr - lz * a.G.b[1];
                const size_t j = (size_t)(o[0] + lx) +
                                 (size_t)a.G.n[0] * ((size_t)(o[1] + ly) + (size_t)a.G.n[1] * (o[2] + lz));
                if (MODE == 3)
                  md.cval64[md.col[j] + pos] = cv;
                else if (pos < MED_K)
                  md.buf64[j * MED_K + pos] = cv;
              }
            }
            if (i > 0 && i % SEG_ALIGN == 0 && ((mk >> i) & 1u)) {  // entry i - 1 ends the segment before
              const uint32_t q = sk + (uint32_t)__popc(mk & ((2u << i) - 2u)) - 1u;
              g = sg[q];
              T = g.w;
            } else {
              T = T + L;
            }
          }
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      __syncthreads();  // the tables are rewritten for the next chunk
    }
    if (MODE == 2) {
      for (uint32_t v = t; v < vol; v += BS) {
        const uint32_t rr = a.G.db0.div(v), lx = v - rr * a.G.b[0];
        const uint32_t lz = a.G.db1.div(rr), ly = rr - lz * a.G.b[1];
        const uint32_t x = o[0] + lx, y = o[1] + ly, z = o[2] + lz;
        if (x >= a.G.n[0] || y >= a.G.n[1] || z >= a.G.n[2]) continue;
        md.bcnt[(size_t)x + (size_t)a.G.n[0] * ((size_t)y + (size_t)a.G.n[1] * z)] = cnt_s[v];
      }
    }
    if (MODE == 0) {
      const bool shared = tk.ntask > 1;
      for (uint32_t v = t; v < vol; v += BS) {
        const uint32_t rr = a.G.db0.div(v), lx = v - rr * a.G.b[0];
        const uint32_t lz = a.G.db1.div(rr), ly = rr - lz * a.G.b[1];
        const uint32_t x = o[0] + lx, y = o[1] + ly, z = o[2] + lz;
        if (x >= a.G.n[0] || y >= a.G.n[1] || z >= a.G.n[2]) continue;
        const size_t j = (size_t)x + (size_t)a.G.n[0] * ((size_t)y + (size_t)a.G.n[1] * z);
        const double val = acc_s[v];
        if (!shared)
          acc[j] = val;
        else if (val != 0.0)
          atomicAdd(&acc[j], val);
      }
    }
    __syncthreads();
  }
}

// Order-statistic backprojection, FP32C (median, trimmed mean): c_ij =
// <W_ij, G_i> per entry from the stream's FP32 lengths, with the track's
// FP64 G' moved to the segment's basis and the tail and the product in FP64
// (a selection takes single values, whose errors do not average out over
// the voxel's entries as the mean's sums do), rounded once to FP32 and
// written to the entry's voxel-major slot (the FP64 stream's slots: the two streams
// share their layout) for the selection that follows.  Tasks of one chunk
// from one queue per XCD, brick by brick, so that a brick's slot range
// stays in that XCD's L2 while its scattered writes land.  No voxel chain:
// the slot names the destination.
template <uint32_t BS>
__global__ void __launch_bounds__(BS)
k_bwd32s(IterArgs a, Args64 b, const double4* __restrict__ trk_g, float* __restrict__ cval) {
  __shared__ double4 sg[BS];
  __shared__ uint32_t sh_[4];
  uint32_t& sh = sh_[0];
  const uint32_t t = threadIdx.x;
  const double U = a.ec.unit, U2 = U * U, U3 = U2 * U;
  const uint32_t qi = blockIdx.x % b.nq;
  for (;;) {
    if (t == 0) sh = b.qoff[qi] + atomicAdd(&b.qwork[qi], 1u);
    __syncthreads();
    const uint32_t ti = sh;
    if (ti >= b.qoff[qi + 1]) break;
    const Task tk = b.qtask[ti];
    for (uint32_t c = tk.c0; c < tk.c1; ++c) {
      const Chunk ch = a.chunks[c];
      if (t < ch.cnt) {
        const uint64_t q = ch.q0 + t;
        const uint32_t tfx = a.seg_tfx[q];
        const double4 g = trk_g[a.seg_trk[q]];
        // A^T G' A with A = [[1, 0], [s, 1]], s = tau_q - t_c (exact), in
        // units of U for the entry terms, all in FP64: rounding G' before
        // this shift would be amplified where s is large
        const double shf = (double)((int64_t)tfx - (int64_t)g.w) * a.inv_scale;
        const double gy = fma(shf, g.z, g.y);
        sg[t] = make_double4(fma(shf, gy + g.y, g.x) * U, gy * U2, g.z * U3, 0.0);
      }
      const bool act = t * GE < ch.ne;
      const uint64_t gi = (ch.base / GE) + t;
      Group gr;
      if (act) gr.load(a, gi);
      __syncthreads();
      if (act) {
        gr.decode();
        uint32_t sl[GE];  // slots of the group's entries (padding: none, all ones)
        const uint4* ps = reinterpret_cast<const uint4*>(b.slot + gi * GE);
#pragma unroll
        for (int k = 0; k < GE / 4; ++k) {
          const uint4 v4 = ps[k];
          sl[4 * k] = v4.x; sl[4 * k + 1] = v4.y; sl[4 * k + 2] = v4.z; sl[4 * k + 3] = v4.w;
        }
        double T = gr.carry;
        uint32_t mk = gr.mask, sk = gr.s0, base_v = ENT_BASE;  // opaque copies, as in k_bwd32
        asm volatile("" : "+v"(mk), "+v"(sk), "+v"(base_v));
        double4 g = sg[sk + (uint32_t)__popc(mk & 0xFFFEu)];
#pragma unroll
        for (int i = GE - 1; i >= 0; --i) {
          // the FP32 length; the tail and <W, G> in FP64
          const double L = ent_len_v(gr.w[i], base_v);
          const uint32_t st = start_mask(mk, i);  // entry i starts a segment
          const double k = fma(L, fma(L, 1.0 / 3.0, T), T * T);
          const double cv = L * fma(k, g.z, fma(fma(2.0, T, L), g.y, g.x));
          if (sl[i] != 0xFFFFFFFFu) cval[sl[i]] = (float)cv;
          if (i > 0 && st != 0u) g = sg[sk + (uint32_t)__popc(mk & ((2u << i) - 2u)) - 1u];
          T = st != 0u ? 0.0 : T + L;
        }
      }
      __syncthreads();  // the table is rewritten for the next chunk
    }
  }
}

// Slot of every stream entry in the voxel-major c_ij array (one block per
// task, the passes' voxel chain): slot = col[j] + a running count per
// voxel.  The order inside a voxel does not matter to the median or to
// the sorted trimmed mean.
template <uint32_t BS>
__global__ void __launch_bounds__(BS)
k_slot64s(IterArgs a, const uint64_t* __restrict__ col, uint32_t* __restrict__ fill,
          uint32_t* __restrict__ slot) {
  __shared__ uint32_t sv0[BS];
  __shared__ int stab[16];
  const uint32_t t = threadIdx.x;
  fill_step_table(a, stab);
  const Task tk = a.tasks[blockIdx.x];
  uint32_t o[3];
  brick_origin(a.G, tk.brick, o);
  for (uint32_t c = tk.c0; c < tk.c1; ++c) {
    const Chunk ch = a.chunks[c];
    if (t < ch.cnt) sv0[t] = a.seg_v0[ch.q0 + t];
    const bool act = t * GE < ch.ne;
    Group gr;
    if (act) gr.load(a, (ch.base / GE) + t);
    __syncthreads();
    if (act) {
      gr.decode();
      const uint32_t e0 = t * GE;
      gr.voxels(stab, [&](uint32_t q) { return sv0[q]; }, [&](int i, uint32_t v) {
        if (e0 + (uint32_t)i >= ch.ne || ent_code(gr.w[i]) == ENT_PAD_CODE) return;  // padding: no slot
        const uint32_t rr = a.G.db0.div(v), lx = v - rr * a.G.b[0];
        const uint32_t lz = a.G.db1.div(rr), ly = rr - lz * a.G.b[1];
        const size_t j = (size_t)(o[0] + lx) +
                         (size_t)a.G.n[0] * ((size_t)(o[1] + ly) + (size_t)a.G.n[1] * (o[2] + lz));
        slot[ch.base + e0 + i] = (uint32_t)(col[j] + atomicAdd(&fill[j], 1u));
      });
    }
    __syncthreads();
  }
}

// lambda_j + lambda_j^2 / (2 M_j) sum_i c_ij: the mean of S_ij over 2
// (src/MutoMLSD.cpp:497-500); MAP: lambda_ML to ml
__global__ void __launch_bounds__(256)
k_upd64s(size_t nv, const uint32_t* __restrict__ mj, double* __restrict__ acc,
         double* __restrict__ lam, double* __restrict__ ml) {
  const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nv) return;
  const uint32_t m = mj[j];
  if (m == 0) return;  // made 0 before the run's first pass, as in k_upd32
  const double l = lam[j];
  const double out = l + l * l / (2.0 * (double)m) * acc[j];
  acc[j] = 0.0;
  if (ml) ml[j] = out; else lam[j] = out;
}

// the order-statistic update from the sorted c values of one voxel (the
// reference sorts S_ij = 2 lambda + c_ij lambda^2, the same order for
// lambda != 0 and all equal for lambda = 0): median (src/MutoMLSD.cpp:
// 505-515) or the mean of the lowest floor(pct/100 M), at least one
// (:501-504); `get(k)` returns the k-th smallest c
template <class Get>
__device__ __forceinline__ double order_stat_c(double l, uint32_t m, int use_mean, int pct, Get get) {
  const double l2 = l * l;
  if (use_mean) {
    uint32_t jm = x86_u32(pct / 100.0 * (double)m);
    if (jm < 1) jm = 1;
    double s = 0.0;
    for (uint32_t k = 0; k < jm; ++k) s = s + (2.0 * l + get(k) * l2);
    return s / (double)jm * 0.5;
  }
  const double med = 2.0 * l + get(m / 2) * l2;
  if (m % 2 == 0) return (med + (2.0 * l + get(m / 2 - 1) * l2)) * 0.25;
  return med * 0.5;
}

// Bitonic sort of the 64 R values of one wavefront, element e = r 64 + lane
// in v[r]; partners one register apart (distances >= 64) are exchanged in
// the lane, the others through a lane shuffle.
template <int R>
__device__ __forceinline__ void wave_bitonic(double (&v)[R]) {
  const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
  for (uint32_t k = 2; k <= 64u * R; k <<= 1) {
#pragma unroll
    for (uint32_t jd = k >> 1; jd > 0; jd >>= 1) {
      if (jd >= 64) {
        const uint32_t rj = jd / 64;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int p = r ^ (int)rj;
          if (p > r) {
            const bool up = ((r * 64u + lane) & k) == 0;
            const double lo = fmin(v[r], v[p]), hi = fmax(v[r], v[p]);
            v[r] = up ? lo : hi;
            v[p] = up ? hi : lo;
          }
        }
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const double o = __shfl_xor(v[r], (int)jd, 64);
          const bool up = ((r * 64u + lane) & k) == 0;
          const bool lower = (lane & jd) == 0;
          v[r] = (lower == up) ? fmin(v[r], o) : fmax(v[r], o);
        }
      }
    }
  }
}

template <int R>
__device__ __forceinline__ double wave_elem(const double (&v)[R], uint32_t e) {
  double x = v[0];
#pragma unroll
  for (int r = 1; r < R; ++r)
    if (e / 64u == (uint32_t)r) x = v[r];
  return __shfl(x, (int)(e % 64u), 64);
}

// rank of the order statistic that is kept as the next pass's pivot: the
// lower median, or the last value of the trimmed mean
__device__ __forceinline__ uint32_t pivot_rank(uint32_t m, int use_mean, int pct) {
  if (!use_mean) return (m - 1) / 2;
  uint32_t jm = x86_u32(pct / 100.0 * (double)m);
  if (jm < 1) jm = 1;
  return jm - 1;
}

// The order statistic from a full sort of the wavefront's values; *piv gets
// the value at pivot_rank.
template <int R>
__device__ __forceinline__ double wave_select(double (&v)[R], uint32_t m, double l, int use_mean,
                                              int pct, double* piv) {
  const uint32_t lane = threadIdx.x & 63u;
  wave_bitonic<R>(v);
  *piv = wave_elem<R>(v, pivot_rank(m, use_mean, pct));
  if (!use_mean) return order_stat_c(l, m, 0, pct, [&](uint32_t k) { return wave_elem<R>(v, k); });
  // trimmed mean: S of the lowest jm values, summed over the wavefront
  uint32_t jm = x86_u32(pct / 100.0 * (double)m);
  if (jm < 1) jm = 1;
  double s = 0.0;
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (r * 64u + lane < jm) s = s + (2.0 * l + v[r] * (l * l));
  s = wave_sum(s);
  return s / (double)jm * 0.5;
}

template <int R, class P>
__device__ __forceinline__ uint32_t wave_count_if(const double (&v)[R], P pred) {
  uint32_t n = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) n += (uint32_t)__popcll(__ballot(pred(v[r])));
  return n;
}

__device__ __forceinline__ double wave_max(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
  return x;
}
__device__ __forceinline__ double wave_min(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = fmin(x, __shfl_xor(x, o, 64));
  return x;
}

constexpr int PIVOT_STEPS = 16;  // distinct values searched past the pivot before a full sort

// The value of rank k among the wavefront's values (+inf padding beyond m
// never counts: k < m), from a pivot p with cl values below it and ce equal
// to it: the distinct values between p and the answer are visited one at a
// time (the largest below, or the smallest above, the last one).  false
// when more than PIVOT_STEPS are needed.
template <int R>
__device__ __forceinline__ bool rank_from_pivot(const double (&v)[R], uint32_t k, double p, uint32_t cl,
                                                uint32_t ce, double* out) {
  if (cl <= k && k < cl + ce) { *out = p; return true; }
  double t = p;
  if (k < cl) {
    uint32_t have = cl;  // the values below t hold ranks [0, have)
    for (int it = 0; it < PIVOT_STEPS; ++it) {
      double mx = -INFINITY;
#pragma unroll
      for (int r = 0; r < R; ++r) mx = fmax(mx, v[r] < t ? v[r] : -INFINITY);
      const double t2 = wave_max(mx);
      const uint32_t n2 = wave_count_if(v, [&](double x) { return x == t2; });
      if (have - n2 <= k) { *out = t2; return true; }
      have -= n2;
      t = t2;
    }
    return false;
  }
  uint32_t have = cl + ce;  // the values up to t hold ranks [0, have)
  for (int it = 0; it < PIVOT_STEPS; ++it) {
    double mn = INFINITY;
#pragma unroll
    for (int r = 0; r < R; ++r) mn = fmin(mn, v[r] > t ? v[r] : INFINITY);
    const double t2 = wave_min(mn);
    const uint32_t n2 = wave_count_if(v, [&](double x) { return x == t2; });
    if (k < have + n2) { *out = t2; return true; }
    have += n2;
    t = t2;
  }
  return false;
}

// The same order statistic from the previous pass's pivot p (the voxel's
// values move little from pass to pass, so the answer is a few distinct
// values away from it); false when it is not (a full sort then runs).
template <int R>
__device__ __forceinline__ bool wave_select_pivot(const double (&v)[R], uint32_t m, double l, int use_mean,
                                                  int pct, double p, double* out, double* piv) {
  const uint32_t lane = threadIdx.x & 63u;
  if (!(p == p)) return false;  // no pivot yet
  const uint32_t cl = wave_count_if(v, [&](double x) { return x < p; });
  const uint32_t ce = wave_count_if(v, [&](double x) { return x == p; });
  const double l2 = l * l;
  if (!use_mean) {
    const uint32_t klo = (m - 1) / 2, khi = m / 2;
    double elo, ehi;
    if (!rank_from_pivot(v, klo, p, cl, ce, &elo)) return false;
    if (khi == klo) {
      ehi = elo;
    } else if (!rank_from_pivot(v, khi, p, cl, ce, &ehi)) {
      return false;
    }
    *piv = elo;
    const double med = 2.0 * l + ehi * l2;  // S of rank m/2, as order_stat_c
    *out = m % 2 == 0 ? (med + (2.0 * l + elo * l2)) * 0.25 : med * 0.5;
    return true;
  }
  // trimmed mean: the sum of the lowest jm values from the sum below p,
  // adjusted by the distinct values between p and rank jm - 1
  uint32_t jm = x86_u32(pct / 100.0 * (double)m);
  if (jm < 1) jm = 1;
  double sl = 0.0;
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (v[r] < p) sl = sl + v[r];
  double sc = wave_sum(sl);
  (void)lane;
  double t = p;
  if (jm < cl) {
    uint32_t have = cl;
    int it = 0;
    while (have > jm) {
      if (++it > PIVOT_STEPS) return false;
      double mx = -INFINITY;
#pragma unroll
      for (int r = 0; r < R; ++r) mx = fmax(mx, v[r] < t ? v[r] : -INFINITY);
      t = wave_max(mx);
      const uint32_t n2 = wave_count_if(v, [&](double x) { return x == t; });
      const uint32_t take = min(n2, have - jm);  // values equal to t dropped from the top
      sc = sc - (double)take * t;
      have -= take;
    }
  } else if (jm <= cl + ce) {
    sc = sc + (double)(jm - cl) * p;
  } else {
    sc = sc + (double)ce * p;
    uint32_t have = cl + ce;
    int it = 0;
    while (have < jm) {
      if (++it > PIVOT_STEPS) return false;
      double mn = INFINITY;
#pragma unroll
      for (int r = 0; r < R; ++r) mn = fmin(mn, v[r] > t ? v[r] : INFINITY);
      t = wave_min(mn);
      const uint32_t n2 = wave_count_if(v, [&](double x) { return x == t; });
      const uint32_t take = min(n2, jm - have);
      sc = sc + (double)take * t;
      have += take;
    }
  }
  *piv = t;  // near rank jm - 1 (a pivot only has to be close)
  *out = (2.0 * l * (double)jm + l2 * sc) / (double)jm * 0.5;
  return true;
}

template <int R, class V>
__device__ __forceinline__ double wave_update(const V* __restrict__ src, uint32_t m, double l,
                                              int use_mean, int pct, double p, double* piv) {
  const uint32_t lane = threadIdx.x & 63u;
  double v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = r * 64u + lane;
    v[r] = e < m ? (double)src[e] : INFINITY;  // (FP32C values widen exactly)
  }
  double out;
  if (wave_select_pivot<R>(v, m, l, use_mean, pct, p, &out, piv)) return out;
  return wave_select<R>(v, m, l, use_mean, pct, piv);
}

// Per voxel, one wavefront (M_j <= 256; longer lists are listed at set-up
// and go to k_sel64b): the order statistic of the voxel's M_j values of c,
// from the previous pass's pivot when it is close, else by a sort in
// registers; the update written in place (each voxel reads only its own
// lambda) and the new pivot kept.  M_j = 0 gives 0 (src/MutoMLSD.cpp:
// 494-496).  MAP: lambda_ML to ml.  V: double (FP64 stream) or float
// (FP32C) values of c.
template <class V>
__global__ void __launch_bounds__(256)
k_sel64w(size_t nv, const uint32_t* __restrict__ mj, const uint64_t* __restrict__ col,
         const V* __restrict__ cval, int use_mean, int pct, double* __restrict__ lam,
         double* __restrict__ ml, double* __restrict__ pivot) {
  const size_t nw = (size_t)gridDim.x * (blockDim.x >> 6);
  for (size_t j = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; j < nv; j += nw) {
    const uint32_t m = mj[j];
    double out, piv = NAN;
    if (m == 0) {
      out = 0.0;
    } else if (m > 256) {
      continue;
    } else {
      const double l = lam[j];
      const V* src = cval + col[j];
      const double p = pivot[j];
      if (m <= 64)
        out = wave_update<1>(src, m, l, use_mean, pct, p, &piv);
      else if (m <= 128)
        out = wave_update<2>(src, m, l, use_mean, pct, p, &piv);
      else
        out = wave_update<4>(src, m, l, use_mean, pct, p, &piv);
    }
    if ((threadIdx.x & 63u) == 0) {
      if (ml) ml[j] = out; else lam[j] = out;
      pivot[j] = piv;
    }
  }
}

__global__ void k_fill_u8(size_t n, uint8_t* __restrict__ x, uint8_t v) {
  const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) x[j] = v;
}

// a voxel's first bracket: everything when M_j <= MED_K, else none (a miss,
// so the first pass collects the whole list); no last median
__global__ void k_brk_init(size_t nv, const uint32_t* __restrict__ mj, float2* __restrict__ brk,
                           float* __restrict__ bmed) {
  const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nv) {
    brk[j] = mj[j] <= MED_K ? make_float2(-INFINITY, INFINITY) : make_float2(INFINITY, -INFINITY);
    bmed[j] = NAN;
  }
}

// a value as a bracket bound: FP32, rounded outwards (exact for FP32C values)
__device__ __forceinline__ float bound_lo(float x) { return x; }
__device__ __forceinline__ float bound_hi(float x) { return x; }
__device__ __forceinline__ float bound_lo(double x) { return __double2float_rd(x); }
__device__ __forceinline__ float bound_hi(double x) { return __double2float_ru(x); }

// The next bracket: the values lo and hi MED_H ranks (or more) either side of
// the median value cm, widened towards where the median is heading (cm plus
// twice its last move, the same spread around it), since the medians drift
// the same way pass after pass while the image converges.  (Twice: 1.1e3
// misses per pass at most over 50 passes at 256^3 x 1e7 muons, against 1.5e3
// and rising with the last move once.)
__device__ __forceinline__ float2 next_bracket(float lo, float hi, float cm, float last) {
  if (!(last == last)) return make_float2(lo, hi);  // no last median
  const float pred = cm + 2.0f * (cm - last);
  return make_float2(fminf(lo, pred - (cm - lo)), fmaxf(hi, pred + (hi - cm)));
}

__global__ void k_fill_nan(size_t n, double* __restrict__ x) {
  const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) x[j] = NAN;
}

// Listed voxels with longer lists, one block each: bitonic sort in LDS
// (M_j <= 4096) or in the voxel's own range of a global scratch.  With
// `miss`, only the voxels it marks, and the next bracket to `brk` (the FP32C
// median by brackets).
__device__ void block_bitonic64(double* a, uint32_t n) {
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

template <class V>
__global__ void __launch_bounds__(1024)
k_sel64b(uint32_t n_list, const uint32_t* __restrict__ list, const uint64_t* __restrict__ off,
         const uint32_t* __restrict__ mj, const uint64_t* __restrict__ col,
         const V* __restrict__ cval, int use_mean, int pct, double* __restrict__ scratch,
         double* __restrict__ lam, double* __restrict__ ml, const uint8_t* __restrict__ miss,
         float2* __restrict__ brk, float* __restrict__ bmed) {
  __shared__ double a_s[4096];
  for (uint32_t b = blockIdx.x; b < n_list; b += gridDim.x) {
    const uint32_t j = list[b];
    if (miss && !miss[j]) continue;  // uniform over the block
    const uint32_t m = mj[j];
    uint32_t n = 1;
    while (n < m) n <<= 1;
    double* a = n <= 4096 ? a_s : scratch + off[b];
    const V* src = cval + col[j];
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) a[i] = i < m ? (double)src[i] : INFINITY;
    __syncthreads();
    block_bitonic64(a, n);
    if (threadIdx.x == 0) {
      const double out = order_stat_c(lam[j], m, use_mean, pct, [&](uint32_t k) { return a[k]; });
      if (ml) ml[j] = out; else lam[j] = out;
      if (brk) {  // as k_sel_miss
        const uint32_t r1 = m / 2, r0 = m % 2 == 0 ? r1 - 1 : r1;
        const float cm = (float)a[r1];
        brk[j] = next_bracket(bound_lo(a[r0 >= 2 * MED_H ? r0 - 2 * MED_H : 0u]),
                              bound_hi(a[min(r1 + 2 * MED_H, m - 1)]), cm, bmed[j]);
        bmed[j] = cm;
      }
    }
    __syncthreads();
  }
}

// ---------------- FP32C median by brackets --------------------------------
// The median of a voxel's c_ij moves little from one pass to the next, so
// each voxel keeps a bracket [lo, hi] of c values around its last median
// (MED_H ranks either side, or everything when M_j <= MED_K).  A pass over
// the stream, one task per brick (so that a brick's counts and write
// positions are its own), computes each c_ij exactly as the slot pass does
// and counts per voxel in LDS the values below the bracket and those inside
// it, writing up to MED_K inside values to the voxel's buffer.  When both
// middle ranks fall inside the bracket the median is selected from those
// values, the same FP32 values the slot pass would store, so the update is
// bit-identical to it; the other voxels (misses) and their bricks go
// through a second pass that collects their whole lists into their slot
// ranges, and a sort.


// FB = false: the bracket pass; FB = true: the bricks with misses, whose
// missed voxels' values all go to their slot ranges.  A workgroup of SUB
// chunk widths CW takes SUB chunks of its brick at a time (its LDS arrays
// are per brick, so two chunks cost no more of them).
template <uint32_t CW, uint32_t SUB, bool FB>
__global__ void __launch_bounds__(CW * SUB)
k_bwd32m(IterArgs a, MedArgs m, const double4* __restrict__ trk_g) {
  constexpr uint32_t BS = CW * SUB;
  extern __shared__ __align__(16) double lds_dyn[];
  const uint32_t vol = a.G.b[0] * a.G.b[1] * a.G.b[2];
  float2* brk_s = reinterpret_cast<float2*>(lds_dyn);
  uint32_t* cnt_s = reinterpret_cast<uint32_t*>(brk_s + vol);
  __shared__ double4 sg[BS];
  __shared__ uint32_t sv0[BS];
  __shared__ int stab[16];
  __shared__ uint32_t sh_[4];
  uint32_t& sh = sh_[0];
  const uint32_t t = threadIdx.x;
  const uint32_t sub = SUB == 1 ? 0u : t / CW;  // this thread's chunk of the SUB (whole wavefronts)
  const uint32_t tl = SUB == 1 ? t : t % CW;    // and its thread in that chunk
  fill_step_table(a, stab);
  const double U = a.ec.unit, U2 = U * U, U3 = U2 * U;
  for (;;) {
    if (t == 0) sh = atomicAdd(m.work, 1u);
    __syncthreads();
    const uint32_t ti = sh;
    if (ti >= m.n_btask) break;
    const Task tk = m.btask[ti];
    if (FB && !m.bmiss[tk.brick]) {
      __syncthreads();  // every thread has read sh before it is rewritten
      continue;
    }
    uint32_t o[3];
    brick_origin(a.G, tk.brick, o);
    for (uint32_t v = t; v < vol; v += BS) {
      const uint32_t rr = a.G.db0.div(v), lx = v - rr * a.G.b[0];
      const uint32_t lz = a.G.db1.div(rr), ly = rr - lz * a.G.b[1];
      const uint32_t x = o[0] + lx, y = o[1] + ly, z = o[2] + lz;
      float2 br = make_float2(INFINITY, -INFINITY);  // outside the grid: no values
      if (x < a.G.n[0] && y < a.G.n[1] && z < a.G.n[2]) {
        const size_t j = (size_t)x + (size_t)a.G.n[0] * ((size_t)y + (size_t)a.G.n[1] * z);
        if (FB)
          br = m.miss[j] ? make_float2(-INFINITY, INFINITY) : make_float2(INFINITY, -INFINITY);
        else
          br = m.brk[j];
      }
      brk_s[v] = br;
      cnt_s[v] = 0u;
    }
    for (uint32_t c0 = tk.c0; c0 < tk.c1; c0 += SUB) {
      Chunk ch{};  // no chunk (cnt = ne = 0) past the task's last
      if (c0 + sub < tk.c1) ch = a.chunks[c0 + sub];
      double4* sgb = sg + sub * CW;
      uint32_t* svb = sv0 + sub * CW;
      if (tl < ch.cnt) {
        const uint64_t q = ch.q0 + tl;
        const uint32_t tfx = a.seg_tfx[q];
        const double4 g = trk_g[a.seg_trk[q]];
        // as in k_bwd32s: A^T G' A in FP64, in units of U
        const double shf = (double)((int64_t)tfx - (int64_t)g.w) * a.inv_scale;
        const double gy = fma(shf, g.z, g.y);
        sgb[tl] = make_double4(fma(shf, gy + g.y, g.x) * U, gy * U2, g.z * U3, 0.0);
        svb[tl] = a.seg_v0[q];
      }
      const bool act = tl * GE < ch.ne;
      Group gr;
      if (act) gr.load(a, (ch.base / GE) + tl);
      __syncthreads();
      if (act) {
        gr.decode();
        uint32_t vx[GE];
        gr.voxels(stab, [&](uint32_t q) { return svb[q]; }, [&](int i, uint32_t v) { vx[i] = v; });
        const uint32_t e0 = tl * GE;
        // each half group's brackets read before its FP64 terms, so the LDS
        // reads overlap the arithmetic
        float2 brs[GE];
#pragma unroll
        for (int i = GE / 2; i < GE; ++i) brs[i] = brk_s[vx[i]];
        double T = gr.carry;
        uint32_t mk = gr.mask, sk = gr.s0, base_v = ENT_BASE;  // opaque copies, as in k_bwd32
        asm volatile("" : "+v"(mk), "+v"(sk), "+v"(base_v));
        double4 g = sgb[sk + (uint32_t)__popc(mk & 0xFFFEu)];
#pragma unroll
        for (int i = GE - 1; i >= 0; --i) {
          if (i == GE / 2 - 1) {
#pragma unroll
            for (int k = 0; k < GE / 2; ++k) brs[k] = brk_s[vx[k]];
          }
          // c_ij exactly as k_bwd32s stores it
          const double L = ent_len_v(gr.w[i], base_v);
          const uint32_t st = start_mask(mk, i);
          const double k = fma(L, fma(L, 1.0 / 3.0, T), T * T);
          const double cv = L * fma(k, g.z, fma(fma(2.0, T, L), g.y, g.x));
          const float cf = (float)cv;
          if (e0 + (uint32_t)i < ch.ne && ent_code(gr.w[i]) != ENT_PAD_CODE) {
            const uint32_t v = vx[i];
            const float2 br = brs[i];
            if (cf < br.x) {
              atomicAdd(&cnt_s[v], 1u);
            } else if (!(cf > br.y)) {
              const uint32_t pos = atomicAdd(&cnt_s[v], 0x10000u) >> 16;
              const uint32_t rr = a.G.db0.div(v), lx = v - rr * a.G.b[0];
              const uint32_t lz = a.G.db1.div(rr), ly = rr - lz * a.G.b[1];
              const size_t j = (size_t)(o[0] + lx) +
                               (size_t)a.G.n[0] * ((size_t)(o[1] + ly) + (size_t)a.G.n[1] * (o[2] + lz));
              if (FB)
                m.cval[m.col[j] + pos] = cf;  // pos < M_j
              else if (pos < MED_K)
                m.buf[j * MED_K + pos] = cf;
            }
          }
          if (i > 0 && st != 0u) g = sgb[sk + (uint32_t)__popc(mk & ((2u << i) - 2u)) - 1u];
          T = st != 0u ? 0.0 : T + L;
        }
      }
      __syncthreads();  // the tables are rewritten for the next chunks
    }
    if (!FB) {
      for (uint32_t v = t; v < vol; v += BS) {
        const uint32_t rr = a.G.db0.div(v), lx = v - rr * a.G.b[0];
        const uint32_t lz = a.G.db1.div(rr), ly = rr - lz * a.G.b[1];
        const uint32_t x = o[0] + lx, y = o[1] + ly, z = o[2] + lz;
        if (x >= a.G.n[0] || y >= a.G.n[1] || z >= a.G.n[2]) continue;
        m.bcnt[(size_t)x + (size_t)a.G.n[0] * ((size_t)y + (size_t)a.G.n[1] * z)] = cnt_s[v];
      }
    }
    __syncthreads();
  }
}

// the median (order_stat_c with update_with_mean 0) of two middle values
__device__ __forceinline__ double median_update(double l, uint32_t m, double c_lo, double c_hi) {
  const double l2 = l * l;
  const double med = 2.0 * l + c_hi * l2;
  if (m % 2 == 0) return (med + (2.0 * l + c_lo * l2)) * 0.25;
  return med * 0.5;
}

// Per voxel (one thread): the median from the bracket's values when both
// middle ranks (m/2 - 1 and m/2 for even m, m/2 for odd) fall inside it, with
// the next bracket MED_H ranks either side (kept where the values run out);
// else a miss, marked on the voxel and its brick.  M_j = 0 gives 0.
template <class V>
__device__ __forceinline__ bool sel_brk_voxel(const BrickGeo& G, size_t j, uint32_t mm, const MedArgs& m,
                                              double* __restrict__ lam, double* __restrict__ ml);
// V: float (FP32C, values in m.buf) or double (FP64 stream, m.buf64)
template <class V>
__global__ void __launch_bounds__(256)
k_sel_brk(BrickGeo G, size_t nv, const uint32_t* __restrict__ mj, MedArgs m,
          double* __restrict__ lam, double* __restrict__ ml) {
  const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool miss = j < nv && sel_brk_voxel<V>(G, j, mj[j], m, lam, ml);
  const unsigned long long bm = __ballot(miss);
  if ((threadIdx.x & 63u) == 0 && bm) atomicAdd(m.nmiss, (unsigned long long)__popcll(bm));
}
// one voxel of k_sel_brk; true on a miss
template <class V>
__device__ __forceinline__ bool sel_brk_voxel(const BrickGeo& G, size_t j, uint32_t mm, const MedArgs& m,
                                              double* __restrict__ lam, double* __restrict__ ml) {
  if (mm == 0) {
    if (ml) ml[j] = 0.0; else lam[j] = 0.0;
    m.miss[j] = 0;
    return false;
  }
  const uint32_t pk = m.bcnt[j], nlo = pk & 0xFFFFu, nin = pk >> 16;
  const uint32_t r1 = mm / 2, r0 = mm % 2 == 0 ? r1 - 1 : r1;
  if (!(nin <= MED_K && nlo <= r0 && r1 < nlo + nin)) {
    m.miss[j] = 1;
    const uint32_t tq = G.dn0.div((uint32_t)j), x = (uint32_t)j - tq * G.n[0];
    const uint32_t z = G.dn1.div(tq), y = tq - z * G.n[1];
    m.bmiss[G.db0.div(x) + G.nb[0] * (G.db1.div(y) + G.nb[1] * G.db2.div(z))] = 1;
    return true;
  }
  m.miss[j] = 0;
  // only the filled part of the buffer is read (about half of it)
  V w[MED_K];
  if constexpr (std::is_same<V, float>::value) {
    const float4* src = reinterpret_cast<const float4*>(m.buf + j * MED_K);
#pragma unroll
    for (uint32_t k = 0; k < MED_K / 4; ++k) {
      float4 q = make_float4(INFINITY, INFINITY, INFINITY, INFINITY);
      if (4 * k < nin) q = src[k];
      w[4 * k] = 4 * k < nin ? q.x : INFINITY;
      w[4 * k + 1] = 4 * k + 1 < nin ? q.y : INFINITY;
      w[4 * k + 2] = 4 * k + 2 < nin ? q.z : INFINITY;
      w[4 * k + 3] = 4 * k + 3 < nin ? q.w : INFINITY;
    }
  } else {
    const double2* src = reinterpret_cast<const double2*>(m.buf64 + j * MED_K);
#pragma unroll
    for (uint32_t k = 0; k < MED_K / 2; ++k) {
      double2 q = make_double2(INFINITY, INFINITY);
      if (2 * k < nin) q = src[k];
      w[2 * k] = 2 * k < nin ? q.x : INFINITY;
      w[2 * k + 1] = 2 * k + 1 < nin ? q.y : INFINITY;
    }
  }
  // bitonic sorting network, ascending
#pragma unroll
  for (uint32_t kk = 2; kk <= MED_K; kk <<= 1) {
#pragma unroll
    for (uint32_t jd = kk >> 1; jd > 0; jd >>= 1) {
#pragma unroll
      for (uint32_t i = 0; i < MED_K; ++i) {
        const uint32_t p = i ^ jd;
        if (p > i) {
          const V lo = fmin(w[i], w[p]), hi = fmax(w[i], w[p]);
          const bool up = (i & kk) == 0;
          w[i] = up ? lo : hi;
          w[p] = up ? hi : lo;
        }
      }
    }
  }
  const uint32_t ia = r0 - nlo, ib = r1 - nlo;
  const float2 br = m.brk[j];
  const uint32_t il = ia >= MED_H ? ia - MED_H : MED_K, ih = ib + MED_H < nin ? ib + MED_H : MED_K;
  V ca = 0, cb = 0;
  float nl = br.x, nh = br.y;
#pragma unroll
  for (uint32_t k = 0; k < MED_K; ++k) {
    if (k == ia) ca = w[k];
    if (k == ib) cb = w[k];
    if (k == il) nl = bound_lo(w[k]);
    if (k == ih) nh = bound_hi(w[k]);
  }
  const double out = median_update(lam[j], mm, (double)ca, (double)cb);
  if (ml) ml[j] = out; else lam[j] = out;
  if (mm > MED_K) {
    m.brk[j] = next_bracket(nl, nh, (float)cb, m.bmed[j]);
    m.bmed[j] = (float)cb;
  }
  return false;
}

// The missed voxels (M_j <= 256; longer ones go to k_sel64b): one wavefront
// per voxel sorts its whole list, as k_sel64w without a pivot, and sets the
// next bracket MED_H ranks either side of the median.
template <int R, class V>
__device__ __forceinline__ void miss_select(const V* __restrict__ src, uint32_t mm, double l,
                                            double* out, float2* nb, float* cm) {
  const uint32_t lane = threadIdx.x & 63u;
  double v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = r * 64u + lane;
    v[r] = e < mm ? (double)src[e] : INFINITY;
  }
  wave_bitonic<R>(v);
  const uint32_t r1 = mm / 2, r0 = mm % 2 == 0 ? r1 - 1 : r1;
  const double cb = wave_elem<R>(v, r1);
  *out = median_update(l, mm, wave_elem<R>(v, r0), cb);
  // a voxel that missed gets twice the ranks either side
  *nb = make_float2(bound_lo(wave_elem<R>(v, r0 >= 2 * MED_H ? r0 - 2 * MED_H : 0u)),
                    bound_hi(wave_elem<R>(v, min(r1 + 2 * MED_H, mm - 1))));
  *cm = (float)cb;
}
template <class V>
__global__ void __launch_bounds__(256)
k_sel_miss(size_t nv, const uint32_t* __restrict__ mj, MedArgs m, double* __restrict__ lam,
           double* __restrict__ ml) {
  // each wavefront looks at 64 voxels at a time and sorts the ones missed
  const size_t nw = (size_t)gridDim.x * (blockDim.x >> 6);
  const uint32_t lane = threadIdx.x & 63u;
  for (size_t base = (((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 64; base < nv; base += nw * 64) {
    const size_t jl = base + lane;
    unsigned long long bm = __ballot(jl < nv && m.miss[jl] && mj[jl] <= 256);
    while (bm) {
      const size_t j = base + (size_t)(__ffsll((long long)bm) - 1);
      bm &= bm - 1;
      const uint32_t mm = mj[j];
      const V* src = std::is_same<V, float>::value ? (const V*)(m.cval + m.col[j]) : (const V*)(m.cval64 + m.col[j]);
      const double l = lam[j];
      double out;
      float2 nb;
      float cm;
      if (mm <= 64)
        miss_select<1, V>(src, mm, l, &out, &nb, &cm);
      else if (mm <= 128)
        miss_select<2, V>(src, mm, l, &out, &nb, &cm);
      else
        miss_select<4, V>(src, mm, l, &out, &nb, &cm);
      if (lane == 0) {
        if (ml) ml[j] = out; else lam[j] = out;
        if (mm > MED_K) {
          m.brk[j] = next_bracket(nb.x, nb.y, cm, m.bmed[j]);
          m.bmed[j] = cm;
        }
      }
    }
  }
}

__global__ void k_zero_uncrossed64(size_t nv, const uint32_t* __restrict__ mj, double* __restrict__ lam) {
  const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nv && mj[j] == 0 && lam[j] != 0.0) lam[j] = 0.0;
}

// clears the FP64 stream iteration's cost, counters and work counters
__global__ void k_iter_begin64(double* __restrict__ cost_acc, unsigned long long* __restrict__ cnt,
                               unsigned int* __restrict__ work, unsigned int* __restrict__ qwork,
                               uint32_t nq) {
  if (threadIdx.x == 0) cost_acc[0] = 0.0;
  if (threadIdx.x < 3 || threadIdx.x == C_MISS) cnt[threadIdx.x] = 0ull;
  if (threadIdx.x < 4) work[threadIdx.x] = 0u;  // (2, 3: the median's bracket passes)
  if (qwork && threadIdx.x < nq) qwork[threadIdx.x] = 0u;
}

__global__ void k_copy64(size_t n, const double* __restrict__ a, double* __restrict__ b) {
  const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) b[j] = a[j];
}

// Voxels no track crosses become 0 (src/MutoMLSD.cpp:494-496); done once
// before a run's first mean-update pass, since no later pass touches them.
__global__ void __launch_bounds__(256)
k_zero_uncrossed(size_t nv, const uint32_t* __restrict__ mj, double* __restrict__ lam,
                 float* __restrict__ lam32) {
  const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nv && mj[j] == 0 && lam[j] != 0.0) {
    lam[j] = 0.0;
    lam32[j] = 0.0f;
  }
}

// lambda_j <- lambda_j + lambda_j^2 / (2 M_j) * sum_i c_ij from the FP64
// image, which stays the iteration's state (the float mirror only feeds the
// forward pass); clears the accumulator and the work counters for the next
// iteration.  MAP: lambda_ML goes to ml and the image is left as
// lambda_last for the voxel step.
__global__ void __launch_bounds__(256)
k_upd32(size_t nv, const uint32_t* __restrict__ mj, double* __restrict__ acc,
        double* __restrict__ lam, float* __restrict__ lam32,
        double* __restrict__ ml, unsigned int* __restrict__ work,
        const unsigned long long* __restrict__ cmax, int fixed, int fx_bits) {
  size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j == 0) { work[0] = 0; work[1] = 0; }
  if (j >= nv) return;
  const uint32_t m = mj[j];
  // no track crosses the voxel: k_zero_uncrossed made it 0 (MutoMLSD.cpp:
  // 494-496) before the run's first pass, and its accumulator, which no
  // entry adds to, stays 0 (the padding of a batch's combined grid is
  // mostly such voxels); the ML image is not read there
  if (m == 0) return;
  const double l = lam[j];
  // the backprojection's sums: int64 fixed point (ACC 1) or FP64
  const double c = fixed ? (double)reinterpret_cast<const long long*>(acc)[j] * (1.0 / fx_scale(cmax, fx_bits))
                         : acc[j];
  const double out = m == 0 ? 0.0 : l + l * l / (2.0 * (double)m) * c;
  acc[j] = 0.0;
  if (ml) {
    ml[j] = out;
  } else {
    lam[j] = out;
    lam32[j] = (float)out;
  }
}

// int64 fixed-point backprojection sums to FP64, in place
__global__ void k_fx_to_f64(size_t nv, double* __restrict__ acc,
                            const unsigned long long* __restrict__ cmax, int fx_bits) {
  size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nv) return;
  const long long q = reinterpret_cast<const long long*>(acc)[j];
  acc[j] = (double)q * (1.0 / fx_scale(cmax, fx_bits));
}

// clears the iteration's cost sum and per-iteration counters
__global__ void k_iter_begin(uint32_t n_scans, double* __restrict__ cost_acc,
                             unsigned long long* __restrict__ cnt,
                             unsigned long long* __restrict__ cmax, unsigned int* __restrict__ work,
                             unsigned int* __restrict__ qwork, uint32_t nq) {
  for (uint32_t k = threadIdx.x; k < n_scans; k += blockDim.x) cost_acc[k] = 0.0;
  if (threadIdx.x < 3 || threadIdx.x == C_MISS) cnt[threadIdx.x] = 0ull;
  if (threadIdx.x == 0) cmax[0] = 0ull;  // the track kernel sets this pass's |c| bound
  if (threadIdx.x < 4) work[threadIdx.x] = 0u;  // (2, 3: the median's bracket passes)
  if (qwork && threadIdx.x < nq) qwork[threadIdx.x] = 0u;  // order statistics: per-XCD queues
}

// largest M_j (a grid-stride loop per thread, one atomic per block)
__global__ void __launch_bounds__(256)
k_max_u32(size_t n, const uint32_t* __restrict__ a, unsigned int* __restrict__ out) {
  __shared__ uint32_t wm[4];
  uint32_t v = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    v = max(v, a[i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    v = max(max(wm[0], wm[1]), max(wm[2], wm[3]));
    if (v) atomicMax(out, v);
  }
}

// Entries per voxel from the brick stream: one block per task counts its
// entries per brick voxel in LDS (the voxel chain of the passes, padding
// entries skipped), then stores the counts (a task that covers its brick)
// or adds them (tasks sharing a brick; mj starts at zero).
template <uint32_t BS>
__global__ void __launch_bounds__(BS)
k_count32(IterArgs a, uint32_t* __restrict__ mj) {
  extern __shared__ __align__(16) double lds_dyn[];
  uint32_t* cnt_s = reinterpret_cast<uint32_t*>(lds_dyn);
  __shared__ uint32_t sv0[BS];
  __shared__ int stab[16];
  const uint32_t t = threadIdx.x;
  fill_step_table(a, stab);
  const uint32_t vol = a.G.b[0] * a.G.b[1] * a.G.b[2];
  const Task tk = a.tasks[blockIdx.x];
  for (uint32_t v = t; v < vol; v += BS) cnt_s[v] = 0u;
  __syncthreads();
  for (uint32_t c = tk.c0; c < tk.c1; ++c) {
    const Chunk ch = a.chunks[c];
    if (t < ch.cnt) sv0[t] = a.seg_v0[ch.q0 + t];
    const bool act = t * GE < ch.ne;
    Group gr;
    if (act) gr.load(a, (ch.base / GE) + t);
    __syncthreads();
    if (act) {
      gr.decode();
      const uint32_t e0 = t * GE;
      gr.voxels(stab, [&](uint32_t q) { return sv0[q]; }, [&](int i, uint32_t v) {
        if (e0 + (uint32_t)i < ch.ne && ent_code(gr.w[i]) != ENT_PAD_CODE) atomicAdd(&cnt_s[v], 1u);
      });
    }
    __syncthreads();
  }
  uint32_t o[3];
  brick_origin(a.G, tk.brick, o);
  for (uint32_t v = t; v < vol; v += BS) {
    const uint32_t rr = a.G.db0.div(v), lx = v - rr * a.G.b[0];
    const uint32_t lz = a.G.db1.div(rr), ly = rr - lz * a.G.b[1];
    const uint32_t x = o[0] + lx, y = o[1] + ly, z = o[2] + lz;
    if (x >= a.G.n[0] || y >= a.G.n[1] || z >= a.G.n[2]) continue;
    const size_t j = (size_t)x + (size_t)a.G.n[0] * ((size_t)y + (size_t)a.G.n[1] * z);
    if (tk.ntask == 1)
      mj[j] = cnt_s[v];
    else if (cnt_s[v])
      atomicAdd(&mj[j], cnt_s[v]);
  }
}

// appends the iteration's cost sums (one per scan) to the history
__global__ void k_cost_push(uint32_t n_scans, const double* __restrict__ cost_acc,
                            double* __restrict__ hist, unsigned int* __restrict__ n) {
  const unsigned int k = n[0];
  for (uint32_t q = threadIdx.x; q < n_scans; q += blockDim.x) hist[(size_t)k * n_scans + q] = cost_acc[q];
  __syncthreads();
  if (threadIdx.x == 0) n[0] = k + 1u;
}

__global__ void k_mirror32(size_t n, const double* __restrict__ a, float* __restrict__ b) {
  size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) b[j] = (float)a[j];
}

// MAP: the voxel step's result becomes the image and its float mirror
__global__ void k_commit(size_t n, const double* __restrict__ a, double* __restrict__ lam,
                         float* __restrict__ lam32) {
  size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) {
    const double v = a[j];
    lam[j] = v;
    lam32[j] = (float)v;
  }
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
  G.dn0 = FastDiv::make(G.n[0]);
  G.dn1 = FastDiv::make(G.n[1]);
  G.db0 = FastDiv::make(G.b[0]);
  G.db1 = FastDiv::make(G.b[1]);
  G.db2 = FastDiv::make(G.b[2]);
  return G;
}

static EntCode entcode(const muto_em_plan* pl) {
  EntCode ec;
  ec.unit = pl->f32.unit;
  ec.bx = (int)pl->f32.bs.bx;
  ec.bxy = (int)(pl->f32.bs.bx * pl->f32.bs.by);
  return ec;
}

static IterArgs iter_args(muto_em_plan* pl) {
  Fp32cData& f = pl->f32;
  IterArgs a;
  a.G = brick_geo(pl);
  a.tasks = f.tasks.p;
  a.chunks = f.chunks.p;
  a.n_tasks = f.n_tasks;
  a.work = pl->work.p;
  a.fx_bits = f.fx_bits;
  a.n_seg = f.n_seg;
  a.seg_tfx = f.seg_tfx.p;
  a.gdesc = f.gdesc.p;
  a.gcarry = f.gcarry.p;
  a.inv_scale = 1.0 / f.tail_scale;
  a.seg_trk = f.seg_trk.p;
  a.seg_v0 = f.seg_v0.p;
  a.ent = f.ent.p;
  a.gv0 = f.gv0.p;
  a.ec = entcode(pl);
  return a;
}

int fp32c_release(muto_em_plan* pl) {
  Fp32cData& f = pl->f32;
  if (f.graph) (void)hipGraphExecDestroy(f.graph);
  f.graph = nullptr;
  if (f.graph64) (void)hipGraphExecDestroy(f.graph64);
  f.graph64 = nullptr;
  return MUTO_OK;
}

static int bits_for(uint64_t v) {
  int b = 0;
  while ((1ull << b) <= v && b < 63) ++b;
  return b;
}

// calls f with the block size as a compile-time constant
template <class F>
static hipError_t by_block(uint32_t block, F&& f) {
  switch (block) {
    case 512: return f(std::integral_constant<uint32_t, 512>{});
    case 1024: return f(std::integral_constant<uint32_t, 1024>{});
    default: return f(std::integral_constant<uint32_t, 256>{});
  }
}

// Workgroups per CU of the forward pass (*f) and of the backprojection
// (*b), and the backprojection's chunks per workgroup (*sub): two when that
// keeps more waves resident (the brick array is held once per workgroup).
static hipError_t occupancy(int acc, uint32_t block, int* f, int* b, int* sub, size_t lds_f,
                            size_t lds_b) {
  return by_block(block, [&](auto B) {
    constexpr uint32_t BS = decltype(B)::value;
    hipError_t e = acc == 1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(f, k_fwd32<1, BS>, BS, lds_f)
                            : hipOccupancyMaxActiveBlocksPerMultiprocessor(f, k_fwd32<0, BS>, BS, lds_f);
    if (e != hipSuccess) return e;
    e = acc == 1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(b, k_bwd32<1, BS, 1>, BS, lds_b)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(b, k_bwd32<0, BS, 1>, BS, lds_b);
    if (e != hipSuccess) return e;
    *sub = 1;
    if constexpr (BS <= 512) {
      int b2 = 0;
      e = acc == 1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b2, k_bwd32<1, BS, 2>, 2 * BS, lds_b)
                   : hipOccupancyMaxActiveBlocksPerMultiprocessor(&b2, k_bwd32<0, BS, 2>, 2 * BS, lds_b);
      if (e != hipSuccess) return e;
      if (2 * b2 > *b) {
        *b = b2;
        *sub = 2;
      }
    }
    return hipSuccess;
  });
}

static hipError_t launch_fwd(const Fp32cData& f, const IterArgs& a, const float* lam, float4* part,
                             hipStream_t s) {
  const size_t lds = f.bs.vol() * sizeof(float);
  return by_block(f.bs.block, [&](auto B) {
    constexpr uint32_t BS = decltype(B)::value;
    if (f.acc_mode == 1)
      k_fwd32<1, BS><<<f.grid_fwd, BS, lds, s>>>(a, lam, part);
    else
      k_fwd32<0, BS><<<f.grid_fwd, BS, lds, s>>>(a, lam, part);
    return hipGetLastError();
  });
}

static hipError_t launch_bwd(const Fp32cData& f, const IterArgs& a, const float* g,
                             const unsigned long long* cmax, double* acc, hipStream_t s) {
  const size_t lds = 2 * f.bs.vol() * sizeof(float);
  return by_block(f.bs.block, [&](auto B) {
    constexpr uint32_t BS = decltype(B)::value;
    if constexpr (BS <= 512) {
      if (f.bwd_sub == 2) {
        if (f.acc_mode == 1)
          k_bwd32<1, BS, 2><<<f.grid_bwd, 2 * BS, lds, s>>>(a, g, cmax, acc);
        else
          k_bwd32<0, BS, 2><<<f.grid_bwd, 2 * BS, lds, s>>>(a, g, cmax, acc);
        return hipGetLastError();
      }
    }
    if (f.acc_mode == 1)
      k_bwd32<1, BS, 1><<<f.grid_bwd, BS, lds, s>>>(a, g, cmax, acc);
    else
      k_bwd32<0, BS, 1><<<f.grid_bwd, BS, lds, s>>>(a, g, cmax, acc);
    return hipGetLastError();
  });
}

// whether both passes' LDS (static + brick arrays) fit one workgroup, and
// the block size is one the passes are built for
static bool fp32c_fits(muto_ctx* ctx, const BrickShape& b) {
  if (b.block != 256 && b.block != 512 && b.block != 1024) return false;
  int max_lds = 0;
  if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) != hipSuccess)
    return false;
  const size_t fwd = fwd_static_lds(b.block) + (size_t)b.vol() * sizeof(float);
  const size_t bwd = bwd_static_lds(b.block, 1) + 2 * (size_t)b.vol() * sizeof(float);
  return std::max(fwd, bwd) <= (size_t)max_lds;
}

int fp32c_build(muto_em_plan* pl) {
  Fp32cData& f = pl->f32;
  if (f.built) return MUTO_OK;
  muto_ctx* ctx = pl->ctx;
  hipStream_t s = ctx->stream;
  {
    const char* ev = getenv("MUTO_FP32C_ACC");
    if (ev && !f.acc_mode_set) f.acc_mode = atoi(ev);
    if (f.acc_mode != 0 && f.acc_mode != 1) f.acc_mode = 1;
  }
  const BrickGeo G = brick_geo(pl);
  f.nb[0] = G.nb[0]; f.nb[1] = G.nb[1]; f.nb[2] = G.nb[2];
  const uint64_t nbk = (uint64_t)G.nb[0] * G.nb[1] * G.nb[2];
  // the backprojection keeps 8 bytes per brick voxel in LDS
  if (f.bs.vol() == 0 || !fp32c_fits(ctx, f.bs)) {
    ctx->set_error("FP32C brick shape does not fit the LDS of one workgroup");
    return MUTO_ERR_LIMIT;
  }
  const size_t nt = pl->n_trk;
  DBuf<uint32_t> nseg, order, rank;
  DBuf<uint64_t> tkey, sstart;
  MUTO_CHECK_HIP(ctx, nseg.alloc(nt + 1));
  MUTO_CHECK_HIP(ctx, order.alloc(nt));
  MUTO_CHECK_HIP(ctx, rank.alloc(nt));
  MUTO_CHECK_HIP(ctx, tkey.alloc(nt));
  MUTO_CHECK_HIP(ctx, sstart.alloc(nt + 1));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(nseg.p, 0, (nt + 1) * sizeof(uint32_t), s));
  // one wavefront (64 threads) per track
  k_seg_count<<<grid_for(nt, 4), 256, 0, s>>>(G, nt, pl->nv, pl->row_ptr.p, pl->vox.p,
                                              pl->trk_scan.p, nseg.p, tkey.p, order.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  MUTO_CHECK_HIP(ctx, sort_pairs_u64_u32(tkey.p, order.p, nt, 0, 64, s));
  k_invert<<<grid_for(nt, 256), 256, 0, s>>>(nt, order.p, rank.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  MUTO_CHECK_HIP(ctx, exclusive_scan_u32_to_u64(nseg.p, sstart.p, nt + 1, s));
  uint64_t n_seg = 0;
  MUTO_CHECK_HIP(ctx, copy_sync(ctx, &n_seg, sstart.p + nt, sizeof n_seg, hipMemcpyDeviceToHost));
  if (n_seg >= 0xFFFFFFFFull) { ctx->set_error("more than 2^32 segments"); return MUTO_ERR_LIMIT; }
  f.n_seg = n_seg;
  // occupancy-derived persistent grids
  const size_t lds_f = f.bs.vol() * sizeof(float), lds_b = 2 * f.bs.vol() * sizeof(float);
  int occ_f = 1, occ_b = 1, sub_b = 1;
  MUTO_CHECK_HIP(ctx, occupancy(f.acc_mode, f.bs.block, &occ_f, &occ_b, &sub_b, lds_f, lds_b));
  f.bwd_sub = sub_b;
  f.grid_fwd = std::max(1, occ_f) * ctx->n_cu;
  f.grid_bwd = std::max(1, occ_b) * ctx->n_cu;
  {
    DBuf<uint64_t> skey, st_start, bstart;
    DBuf<uint32_t> sval, st_len, st_trk, lenq;
    DBuf<double> st_tend;
    MUTO_CHECK_HIP(ctx, skey.alloc(n_seg));
    MUTO_CHECK_HIP(ctx, sval.alloc(n_seg));
    MUTO_CHECK_HIP(ctx, st_start.alloc(n_seg));
    MUTO_CHECK_HIP(ctx, st_len.alloc(n_seg));
    MUTO_CHECK_HIP(ctx, st_tend.alloc(n_seg));
    MUTO_CHECK_HIP(ctx, st_trk.alloc(n_seg));
    k_seg_fill<<<grid_for(nt, 4), 256, 0, s>>>(G, nt, pl->nv, pl->row_ptr.p, pl->vox.p,
                                               pl->tail.p, sstart.p, rank.p, skey.p, sval.p,
                                               st_start.p, st_len.p, st_tend.p, st_trk.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
    MUTO_CHECK_HIP(ctx, sort_pairs_u64_u32(skey.p, sval.p, n_seg, 0, 32 + bits_for(nbk), s));
    // lengths (brick-major) and brick starts
    MUTO_CHECK_HIP(ctx, lenq.alloc(n_seg + 1));
    MUTO_CHECK_HIP(ctx, bstart.alloc(nbk + 1));
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(bstart.p, 0xFF, (nbk + 1) * sizeof(uint64_t), s));
    if (n_seg) {
      k_seg_len<<<grid_for(n_seg, 256), 256, 0, s>>>(n_seg, sval.p, st_len.p, skey.p, lenq.p, bstart.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    }
    std::vector<uint32_t> hl(n_seg);
    std::vector<uint64_t> hb(nbk + 1);
    MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
    if (n_seg) MUTO_CHECK_HIP(ctx, copy_sync(ctx, hl.data(), lenq.p, n_seg * sizeof(uint32_t), hipMemcpyDeviceToHost));
    MUTO_CHECK_HIP(ctx, copy_sync(ctx, hb.data(), bstart.p, (nbk + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost));
    hb[nbk] = n_seg;
    for (uint64_t b = nbk; b-- > 0;)
      if (hb[b] == ~0ull) hb[b] = hb[b + 1];  // empty brick
    uint64_t n_ent = 0;
    uint32_t max_len = 0;  // padded
    for (uint32_t v : hl) { n_ent += v; max_len = std::max(max_len, seg_padded(v)); }
    const uint32_t BSZ = f.bs.block, ECAP = BSZ * GE;  // chunk capacity
    if (max_len > ECAP) {
      ctx->set_error("a brick segment is longer than the staging capacity");
      return MUTO_ERR_LIMIT;
    }
    // tasks: bricks cut into runs of about `target` entries (at least 16 per
    // brick voxel so the flush is amortised; small enough for the queue to
    // balance); tasks into chunks of <= block - 2 segments and <= ECAP
    // entries.
    // Every chunk starts on a 16-entry boundary: its entries are padded to
    // whole groups with zero-length entries.
    const uint64_t workers = (uint64_t)std::max(f.grid_fwd, f.grid_bwd);
    const uint64_t target = std::max<uint64_t>(16ull * f.bs.vol(), n_ent / (4 * workers) + 1);
    std::vector<Task> tasks;
    std::vector<uint64_t> tent;
    std::vector<Chunk> chunks;
    uint64_t n_ent_pad = 0;
    for (uint64_t b = 0; b < nbk; ++b) {
      const uint64_t s0 = hb[b], s1 = hb[b + 1];
      if (s1 == s0) continue;
      uint64_t eb = 0;
      for (uint64_t q = s0; q < s1; ++q) eb += seg_padded(hl[q]);
      const uint64_t k = std::max<uint64_t>(1, (eb + target - 1) / target);
      const uint64_t per = (eb + k - 1) / k;
      uint64_t q = s0;
      std::vector<Task> bt;
      while (q < s1) {
        Task tk;
        tk.brick = (uint32_t)b;
        tk.c0 = (uint32_t)chunks.size();
        uint64_t te = 0;
        const uint64_t tq0 = q;
        while (q < s1 && (te < per || bt.size() + 1 == k) && q - tq0 + BSZ <= MAX_TASK_SEG) {
          Chunk ch{};
          ch.q0 = q;
          uint32_t ce = 0, cn = 0;
          while (q < s1 && cn < chunk_seg_cap(BSZ) && ce + seg_padded(hl[q]) <= ECAP) {
            ce += seg_padded(hl[q]);
            ++cn;
            ++q;
          }
          ch.cnt = cn;
          ch.ne = ce;
          ch.base = n_ent_pad;
          n_ent_pad += (ce + GE - 1) / GE * GE;
          chunks.push_back(ch);
          te += ce;
        }
        tk.c1 = (uint32_t)chunks.size();
        bt.push_back(tk);
        tent.push_back(te);
      }
      for (Task& tk : bt) {
        tk.ntask = (uint32_t)bt.size();
        tasks.push_back(tk);
      }
    }
    if (chunks.size() >= 0xFFFFFFFFull) { ctx->set_error("too many chunks"); return MUTO_ERR_LIMIT; }
    std::vector<size_t> idx(tasks.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](size_t x, size_t y) { return tent[x] > tent[y]; });
    std::vector<Task> sorted(tasks.size());
    for (size_t i = 0; i < idx.size(); ++i) sorted[i] = tasks[idx[i]];
    f.n_tasks = (uint32_t)sorted.size();
    f.n_chunks = (uint32_t)chunks.size();
    MUTO_CHECK_HIP(ctx, f.tasks.alloc(sorted.size()));
    MUTO_CHECK_HIP(ctx, f.chunks.alloc(chunks.size()));
    if (!sorted.empty())
      MUTO_CHECK_HIP(ctx, copy_sync(ctx, f.tasks.p, sorted.data(), sorted.size() * sizeof(Task), hipMemcpyHostToDevice));
    if (!chunks.empty())
      MUTO_CHECK_HIP(ctx, copy_sync(ctx, f.chunks.p, chunks.data(), chunks.size() * sizeof(Chunk), hipMemcpyHostToDevice));
    // brick-major segment arrays in final order; tails as fixed point with
    // the finest power-of-two step that keeps the longest path below 2^32
    {
      DBuf<unsigned long long> mp;
      MUTO_CHECK_HIP(ctx, mp.alloc(1));
      MUTO_CHECK_HIP(ctx, hipMemsetAsync(mp.p, 0, sizeof(unsigned long long), s));
      k_max_path<<<grid_for(nt, 256), 256, 0, s>>>(nt, pl->row_ptr.p, pl->len.p, pl->tail.p, mp.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
      unsigned long long bitsv = 0;
      MUTO_CHECK_HIP(ctx, copy_sync(ctx, &bitsv, mp.p, sizeof bitsv, hipMemcpyDeviceToHost));
      double maxp;
      memcpy(&maxp, &bitsv, sizeof maxp);
      const double bound = maxp * 1.0001 + 1.0;
      int bits = (int)floor(log2(4294967295.0 / bound));
      if (bits > 40) bits = 40;
      f.tail_scale = ldexp(1.0, bits);
    }
    MUTO_CHECK_HIP(ctx, f.seg_trk.alloc(n_seg));
    MUTO_CHECK_HIP(ctx, f.seg_tfx.alloc(n_seg));
    if (f.with64) MUTO_CHECK_HIP(ctx, f.seg_tail64.alloc(n_seg));
    DBuf<uint64_t> seg_off, seg_off_p;
    MUTO_CHECK_HIP(ctx, seg_off.alloc(n_seg + 1));
    MUTO_CHECK_HIP(ctx, seg_off_p.alloc(n_seg + 1));
    if (n_seg) {
      k_seg_gather<<<grid_for(n_seg, 256), 256, 0, s>>>(n_seg, sval.p, st_tend.p, st_trk.p, rank.p,
                                                        f.tail_scale, f.seg_trk.p, f.seg_tfx.p,
                                                        f.with64 ? f.seg_tail64.p : nullptr);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    }
    // segment offsets from the padded lengths
    if (n_seg) {
      k_pad_len<<<grid_for(n_seg, 256), 256, 0, s>>>(n_seg, lenq.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    }
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(lenq.p + n_seg, 0, sizeof(uint32_t), s));
    MUTO_CHECK_HIP(ctx, exclusive_scan_u32_to_u64(lenq.p, seg_off.p, n_seg + 1, s));
    f.n_ent = n_ent;
    f.n_ent_pad = n_ent_pad;
    if (f.n_chunks) {
      k_chunk_offsets<<<grid_for(f.n_chunks, 256), 256, 0, s>>>(f.n_chunks, f.chunks.p, seg_off.p,
                                                                seg_off_p.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    }
    // entries with their padding (zero-length entries of voxel 0)
    const uint64_t n_grp = n_ent_pad / GE;
    MUTO_CHECK_HIP(ctx, f.ent.alloc(n_ent_pad + GE));
    MUTO_CHECK_HIP(ctx, f.seg_v0.alloc(n_seg + 1));
    MUTO_CHECK_HIP(ctx, f.gdesc.alloc(n_grp + 1));
    MUTO_CHECK_HIP(ctx, f.gcarry.alloc(n_grp + 1));
    MUTO_CHECK_HIP(ctx, f.gv0.alloc(n_grp + 1));
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(f.ent.p, 0, (n_ent_pad + GE) * sizeof(uint32_t), s));
    if (f.with64) {  // padding entries: zero length
      MUTO_CHECK_HIP(ctx, f.len64.alloc(n_ent_pad + GE));
      MUTO_CHECK_HIP(ctx, f.gtail64.alloc(n_grp + 1));
      MUTO_CHECK_HIP(ctx, hipMemsetAsync(f.len64.p, 0, (n_ent_pad + GE) * sizeof(double), s));
      MUTO_CHECK_HIP(ctx, hipMemsetAsync(f.gtail64.p, 0, (n_grp + 1) * sizeof(double), s));
    }
    {
      // lengths as FP32 in units of U, the power of two above twice the
      // voxel diagonal (an entry of a scattered track may hold both legs in
      // one voxel)
      const double dv = sqrt(pl->g.d[0] * pl->g.d[0] + pl->g.d[1] * pl->g.d[1] + pl->g.d[2] * pl->g.d[2]);
      f.unit = ldexp(1.0, (int)floor(log2(2.0 * dv)) + 1);
      const double inv_unit = 1.0 / f.unit;
      DBuf<unsigned long long> ncl;
      MUTO_CHECK_HIP(ctx, ncl.alloc(1));
      MUTO_CHECK_HIP(ctx, hipMemsetAsync(ncl.p, 0, sizeof(unsigned long long), s));
      if (f.n_chunks) {
        MUTO_CHECK_HIP(ctx, by_block(f.bs.block, [&](auto B) {
          constexpr uint32_t BS = decltype(B)::value;
          k_ent_fill<BS><<<f.n_chunks, BS, 0, s>>>(G, f.chunks.p, sval.p, st_start.p, st_len.p, seg_off_p.p,
                                                   pl->vox.p, pl->len.p, inv_unit, f.ent.p,
                                                   f.seg_v0.p, ncl.p, pl->tail.p,
                                                   f.with64 ? f.len64.p : nullptr,
                                                   f.with64 ? f.gtail64.p : nullptr);
          return hipGetLastError();
        }));
      }
      unsigned long long nc = 0;
      MUTO_CHECK_HIP(ctx, copy_sync(ctx, &nc, ncl.p, sizeof nc, hipMemcpyDeviceToHost));
      if (nc) {
        ctx->set_error("FP32C: " + std::to_string(nc) +
                       " path lengths outside 0..2 voxel diagonals (the packed length range)");
        return MUTO_ERR_LIMIT;
      }
    }
    if (f.n_chunks) {
      MUTO_CHECK_HIP(ctx, by_block(f.bs.block, [&](auto B) {
        constexpr uint32_t BS = decltype(B)::value;
        k_group_desc<BS><<<f.n_chunks, BS, 0, s>>>(f.chunks.p, seg_off_p.p, f.ent.p, f.seg_v0.p,
                                                   entcode(pl), f.gdesc.p, f.gcarry.p, f.gv0.p);
        return hipGetLastError();
      }));
    }
    // per-rank segment lists and per-track data
    DBuf<uint32_t> nseg_rank;
    DBuf<uint64_t> tptr64;
    MUTO_CHECK_HIP(ctx, nseg_rank.alloc(nt + 1));
    MUTO_CHECK_HIP(ctx, tptr64.alloc(nt + 1));
    MUTO_CHECK_HIP(ctx, f.trk_d.alloc(4 * nt));
    if (pl->n_scans > 1) MUTO_CHECK_HIP(ctx, f.trk_scan.alloc(nt));
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(nseg_rank.p, 0, (nt + 1) * sizeof(uint32_t), s));
    k_trk_d<<<grid_for(nt, 256), 256, 0, s>>>(nt, order.p, pl->trk4.p, pl->pr.p, nseg.p,
                                              f.trk_d.p, nseg_rank.p, pl->trk_scan.p,
                                              pl->n_scans > 1 ? f.trk_scan.p : nullptr);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
    MUTO_CHECK_HIP(ctx, exclusive_scan_u32_to_u64(nseg_rank.p, tptr64.p, nt + 1, s));
    std::vector<uint64_t> h64(nt + 1);
    MUTO_CHECK_HIP(ctx, copy_sync(ctx, h64.data(), tptr64.p, (nt + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost));
    std::vector<uint32_t> h32(h64.begin(), h64.end());
    MUTO_CHECK_HIP(ctx, f.tseg_ptr.alloc(nt + 1));
    MUTO_CHECK_HIP(ctx, copy_sync(ctx, f.tseg_ptr.p, h32.data(), (nt + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
    MUTO_CHECK_HIP(ctx, f.tseg_pos.alloc(n_seg));
    if (n_seg) {
      k_tseg_fill<<<grid_for(n_seg, 256), 256, 0, s>>>(n_seg, sval.p, st_trk.p, sstart.p, rank.p,
                                                       f.tseg_ptr.p, f.tseg_pos.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    }
    MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  }
  MUTO_CHECK_HIP(ctx, f.seg_part.alloc(4 * n_seg));
  MUTO_CHECK_HIP(ctx, f.trk_g.alloc(4 * nt));
  if (f.with64) {
    MUTO_CHECK_HIP(ctx, f.seg_part64.alloc(4 * n_seg));
    MUTO_CHECK_HIP(ctx, f.trk_g64.alloc(4 * nt));
  }
  MUTO_CHECK_HIP(ctx, f.cmax.alloc(1));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(f.cmax.p, 0, sizeof(unsigned long long), s));
  {
    // fixed-point bits of the backprojection sums: a voxel's M_j entries,
    // each at most 2^fx_bits, must sum inside int64
    if (!pl->mj_ready) {
      // M_j from the stream (mj holds zeros until first counted)
      if (f.n_tasks) {
        const IterArgs a = iter_args(pl);
        const size_t lds = (size_t)f.bs.vol() * sizeof(uint32_t);
        MUTO_CHECK_HIP(ctx, by_block(f.bs.block, [&](auto B) {
          constexpr uint32_t BS = decltype(B)::value;
          k_count32<BS><<<f.n_tasks, BS, lds, s>>>(a, pl->mj.p);
          return hipGetLastError();
        }));
      }
      pl->mj_ready = true;
    }
    DBuf<unsigned int> mx;
    MUTO_CHECK_HIP(ctx, mx.alloc(1));
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(mx.p, 0, sizeof(unsigned int), s));
    k_max_u32<<<1024, 256, 0, s>>>(pl->nv, pl->mj.p, mx.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
    unsigned int max_mj = 0;
    MUTO_CHECK_HIP(ctx, copy_sync(ctx, &max_mj, mx.p, sizeof max_mj, hipMemcpyDeviceToHost));
    f.max_mj = max_mj;
    f.fx_bits = std::min(FX_BITS, 62 - bits_for(std::max<uint64_t>(max_mj, 1)));
  }
  {
    const double dv = sqrt(pl->g.d[0] * pl->g.d[0] + pl->g.d[1] * pl->g.d[1] + pl->g.d[2] * pl->g.d[2]);
    const double bx = f.bs.bx * pl->g.d[0], by = f.bs.by * pl->g.d[1], bz = f.bs.bz * pl->g.d[2];
    f.lmax = 2.0 * dv;  // a merged entry holds both legs
    f.ubound = 2.0 * sqrt(bx * bx + by * by + bz * bz) + f.lmax;
  }
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(pl->work.p, 0, 4 * sizeof(unsigned int), s));
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  f.built = true;
  pl->counters.n_segments = (int64_t)n_seg;
  return MUTO_OK;
}

static bool mean_all(const muto_em_plan* pl);
static Args64 args64(muto_em_plan* pl);
static int fp64s_slots(muto_em_plan* pl);

// whether this plan's order statistic runs by brackets: the median (not
// the trimmed mean), the option on, M_j within the 16-bit counts, and the
// arrays built (med_setup)
static bool fp32c_brackets(const muto_em_plan* pl) {
  const Fp32cData& f = pl->f32;
  return !pl->cfg.update_with_mean && f.med_brk && f.med_ready && f.bbuf.p && f.grid_med > 0;
}
static bool fp64s_brackets(const muto_em_plan* pl) {
  const Fp32cData& f = pl->f32;
  return !pl->cfg.update_with_mean && f.med_brk && f.med_ready && f.bbuf64.p && f.grid_med64 > 0;
}

// the FP64 stream's passes run 256- or 512-thread blocks
template <class F>
static hipError_t by_block64(uint32_t block, F&& f) {
  if (block == 512) return f(std::integral_constant<uint32_t, 512>{});
  return f(std::integral_constant<uint32_t, 256>{});
}

static MedArgs med_args(muto_em_plan* pl, unsigned int* work) {
  Fp32cData& f = pl->f32;
  MedArgs m;
  m.btask = f.btask.p;
  m.n_btask = f.n_btask;
  m.work = work;
  m.brk = f.brk.p;
  m.bmed = f.bmed.p;
  m.bcnt = f.bcnt.p;
  m.buf = f.bbuf.p;
  m.buf64 = f.bbuf64.p;
  m.miss = f.bmiss_v.p;
  m.bmiss = f.bmiss_b.p;
  m.col = f.col64.p;
  m.cval = reinterpret_cast<float*>(f.cval.p);
  m.cval64 = f.cval.p;
  m.nmiss = pl->dcount.p + C_MISS;
  return m;
}

// One task per brick (its chunks in order), largest first; the bracket
// arrays, every voxel starting without a bracket unless M_j <= MED_K.  Needs
// the slot ranges (fp64s_slots).  Skipped (the slot pass then runs) when a
// voxel has 65536 entries or more.
static int med_setup(muto_em_plan* pl, bool fp64) {
  Fp32cData& f = pl->f32;
  if (!f.med_brk || pl->cfg.update_with_mean || f.max_mj >= 65536u) return MUTO_OK;
  if (fp64 ? (f.bbuf64.p != nullptr) : (f.bbuf.p != nullptr)) return MUTO_OK;
  muto_ctx* ctx = pl->ctx;
  hipStream_t s = ctx->stream;
  const size_t nv = pl->nv;
  // the pass's LDS: bracket and counts per brick voxel, and its tables
  // (FP32C: two chunks per workgroup below 1024-thread blocks)
  const size_t lds = (size_t)f.bs.vol() * (sizeof(float2) + sizeof(uint32_t));
  const size_t threads = !fp64 && f.bs.block <= 512 ? 2 * (size_t)f.bs.block : (size_t)f.bs.block;
  int max_lds = 0;
  MUTO_CHECK_HIP(ctx, hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
  if (lds + threads * (sizeof(double4) + sizeof(uint32_t)) + 128 > (size_t)max_lds) return MUTO_OK;
  int ob = 0;
  if (fp64) {
    MUTO_CHECK_HIP(ctx, by_block64(f.bs.block, [&](auto B) {
      constexpr uint32_t BS = decltype(B)::value;
      return hipOccupancyMaxActiveBlocksPerMultiprocessor(&ob, k_bwd64s<2, BS>, BS, lds);
    }));
  } else {
    MUTO_CHECK_HIP(ctx, by_block(f.bs.block, [&](auto B) {
      constexpr uint32_t CW = decltype(B)::value;
      constexpr uint32_t SUB = CW <= 512 ? 2 : 1;
      return hipOccupancyMaxActiveBlocksPerMultiprocessor(&ob, k_bwd32m<CW, SUB, false>, CW * SUB, lds);
    }));
  }
  if (ob < 1) return MUTO_OK;
  if (!f.med_ready) {
    // one task per brick, largest first, and the brackets (shared by both
    // precisions: any bracket gives the same medians)
    std::vector<Task> tk(f.n_tasks);
    if (f.n_tasks)
      MUTO_CHECK_HIP(ctx, copy_sync(ctx, tk.data(), f.tasks.p, f.n_tasks * sizeof(Task), hipMemcpyDeviceToHost));
    std::vector<Chunk> ch(f.n_chunks);
    if (f.n_chunks)
      MUTO_CHECK_HIP(ctx, copy_sync(ctx, ch.data(), f.chunks.p, f.n_chunks * sizeof(Chunk), hipMemcpyDeviceToHost));
    std::sort(tk.begin(), tk.end(), [](const Task& x, const Task& y) {
      return x.brick != y.brick ? x.brick < y.brick : x.c0 < y.c0;
    });
    std::vector<Task> bt;
    std::vector<uint64_t> ne;
    for (const Task& t : tk) {
      if (!bt.empty() && bt.back().brick == t.brick) {
        bt.back().c0 = std::min(bt.back().c0, t.c0);
        bt.back().c1 = std::max(bt.back().c1, t.c1);
      } else {
        bt.push_back(Task{t.brick, 1u, t.c0, t.c1});
      }
    }
    for (const Task& t : bt) {
      uint64_t n = 0;
      for (uint32_t c = t.c0; c < t.c1; ++c) n += ch[c].ne;
      ne.push_back(n);
    }
    std::vector<size_t> order(bt.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return ne[x] > ne[y]; });
    std::vector<Task> sorted(bt.size());
    for (size_t i = 0; i < order.size(); ++i) sorted[i] = bt[order[i]];
    const uint64_t nbk = (uint64_t)f.nb[0] * f.nb[1] * f.nb[2];
    MUTO_CHECK_HIP(ctx, f.btask.alloc(std::max<size_t>(sorted.size(), 1)));
    if (!sorted.empty())
      MUTO_CHECK_HIP(ctx, copy_sync(ctx, f.btask.p, sorted.data(), sorted.size() * sizeof(Task), hipMemcpyHostToDevice));
    f.n_btask = (uint32_t)sorted.size();
    MUTO_CHECK_HIP(ctx, f.brk.alloc(nv));
    MUTO_CHECK_HIP(ctx, f.bmed.alloc(nv));
    MUTO_CHECK_HIP(ctx, f.bcnt.alloc(nv));
    MUTO_CHECK_HIP(ctx, f.bmiss_v.alloc(nv));
    MUTO_CHECK_HIP(ctx, f.bmiss_b.alloc(std::max<uint64_t>(nbk, 1)));
    k_brk_init<<<grid_for(nv, 256), 256, 0, s>>>(nv, pl->mj.p, f.brk.p, f.bmed.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
    f.med_ready = true;
  }
  if (fp64) {
    MUTO_CHECK_HIP(ctx, f.bbuf64.alloc(nv * MED_K));
    f.grid_med64 = ob * ctx->n_cu;
  } else {
    MUTO_CHECK_HIP(ctx, f.bbuf.alloc(nv * MED_K));
    f.grid_med = ob * ctx->n_cu;
  }
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  return MUTO_OK;
}

// One FP32C iteration enqueued on the plan's stream with no host
// synchronisation and no memset, so it can be captured into a graph: a
// kernel clears the per-iteration sums and counters, the last kernel appends
// the iteration's cost sum to the device cost history.  With events (5),
// kernel boundaries are recorded for per-kernel timing.  The mean update
// backprojects into per-voxel sums (k_bwd32, k_upd32); the median and the
// trimmed mean write each c_ij to its voxel's slot (k_bwd32s) and select
// per voxel (k_sel64w / k_sel64b on the float values), as the FP64 stream
// does.
static int fp32c_enqueue(muto_em_plan* pl, hipEvent_t* ev) {
  Fp32cData& f = pl->f32;
  muto_ctx* ctx = pl->ctx;
  hipStream_t s = ctx->stream;
  const IterArgs a = iter_args(pl);
  const ErrM E = error_matrix(pl->cfg);
  const bool map = pl->is_map, mean = mean_all(pl);
  k_iter_begin<<<1, 256, 0, s>>>(pl->n_scans, pl->cost_acc.p, pl->dcount.p, f.cmax.p, pl->work.p,
                                 mean ? nullptr : f.qwork64.p, f.nq64);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  if (ev) MUTO_CHECK_HIP(ctx, hipEventRecord(ev[0], s));
  if (f.n_tasks)
    MUTO_CHECK_HIP(ctx, launch_fwd(f, a, pl->lam32.p, reinterpret_cast<float4*>(f.seg_part.p), s));
  if (ev) MUTO_CHECK_HIP(ctx, hipEventRecord(ev[1], s));
  k_track32<<<grid_for(pl->n_trk, 256), 256, 0, s>>>(pl->n_trk, f.tseg_ptr.p, f.tseg_pos.p,
                                                      reinterpret_cast<const float4*>(f.seg_part.p),
                                                      f.trk_d.p, E, f.tail_scale, f.lmax, f.ubound, 0,
                                                      f.trk_g.p, mean ? nullptr : reinterpret_cast<double4*>(f.trk_g64.p),
                                                      f.cmax.p, pl->n_scans > 1 ? f.trk_scan.p : nullptr,
                                                      pl->cost_acc.p, pl->dcount.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  if (ev) MUTO_CHECK_HIP(ctx, hipEventRecord(ev[2], s));
  float* cval32 = reinterpret_cast<float*>(f.cval.p);  // order statistics: the slots' values
  const bool brk = !mean && fp32c_brackets(pl);
  const size_t lds_med = (size_t)f.bs.vol() * (sizeof(float2) + sizeof(uint32_t));
  const double4* g64 = reinterpret_cast<const double4*>(f.trk_g64.p);
  if (f.n_tasks && mean) MUTO_CHECK_HIP(ctx, launch_bwd(f, a, f.trk_g.p, f.cmax.p, pl->acc64.p, s));
  if (brk) {
    // counts below / inside each voxel's bracket and the values inside
    const uint64_t nbk = (uint64_t)f.nb[0] * f.nb[1] * f.nb[2];
    k_fill_u8<<<grid_for(nbk, 256), 256, 0, s>>>(nbk, f.bmiss_b.p, 0);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
    const MedArgs m = med_args(pl, pl->work.p + 2);
    MUTO_CHECK_HIP(ctx, by_block(f.bs.block, [&](auto B) {
      constexpr uint32_t CW = decltype(B)::value;
      constexpr uint32_t SUB = CW <= 512 ? 2 : 1;
      k_bwd32m<CW, SUB, false><<<f.grid_med, CW * SUB, lds_med, s>>>(a, m, g64);
      return hipGetLastError();
    }));
  }
  if (f.n_tasks && !mean && !brk) {
    const Args64 b = args64(pl);
    MUTO_CHECK_HIP(ctx, by_block(f.bs.block, [&](auto B) {
      constexpr uint32_t BS = decltype(B)::value;
      k_bwd32s<BS><<<f.grid_bwd32s, BS, 0, s>>>(a, b, reinterpret_cast<const double4*>(f.trk_g64.p), cval32);
      return hipGetLastError();
    }));
  }
  if (ev) MUTO_CHECK_HIP(ctx, hipEventRecord(ev[3], s));
  if (mean) {
    k_upd32<<<grid_for(pl->nv, 256), 256, 0, s>>>(pl->nv, pl->mj.p, pl->acc64.p, pl->lam64.p,
                                                  pl->lam32.p, map ? pl->ml64.p : nullptr,
                                                  pl->work.p, f.cmax.p, f.acc_mode == 1 ? 1 : 0,
                                                  f.fx_bits);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
  } else if (brk) {
    // medians inside their brackets; then the misses: their bricks' pass
    // collects their whole lists into their slot ranges, and a sort
    double* ml = map ? pl->ml64.p : nullptr;
    MedArgs m = med_args(pl, pl->work.p + 3);
    k_sel_brk<float><<<grid_for(pl->nv, 256), 256, 0, s>>>(a.G, pl->nv, pl->mj.p, m, pl->lam64.p, ml);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
    MUTO_CHECK_HIP(ctx, by_block(f.bs.block, [&](auto B) {
      constexpr uint32_t CW = decltype(B)::value;
      constexpr uint32_t SUB = CW <= 512 ? 2 : 1;
      k_bwd32m<CW, SUB, true><<<f.grid_med, CW * SUB, lds_med, s>>>(a, m, g64);
      return hipGetLastError();
    }));
    const unsigned g = (unsigned)std::min<size_t>(grid_for(pl->nv, 256), (size_t)ctx->n_cu * 16);
    k_sel_miss<float><<<g, 256, 0, s>>>(pl->nv, pl->mj.p, m, pl->lam64.p, ml);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
    if (f.n_big64) {
      k_sel64b<float><<<std::min<uint32_t>(f.n_big64, 4096), 1024, 0, s>>>(
          f.n_big64, f.big64.p, f.big64_off.p, pl->mj.p, f.col64.p, cval32, 0, 100, f.big64_scr.p,
          pl->lam64.p, ml, f.bmiss_v.p, f.brk.p, f.bmed.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    }
    if (!map) {  // the forward pass reads the float mirror
      k_mirror32<<<grid_for(pl->nv, 256), 256, 0, s>>>(pl->nv, pl->lam64.p, pl->lam32.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    }
  } else {
    const muto_mlsd_cfg& c = pl->cfg;
    double* ml = map ? pl->ml64.p : nullptr;
    const unsigned g = (unsigned)std::min<size_t>(grid_for(pl->nv, 4), (size_t)ctx->n_cu * 64);
    k_sel64w<float><<<g, 256, 0, s>>>(pl->nv, pl->mj.p, f.col64.p, cval32, c.update_with_mean,
                                      c.mean_percentile, pl->lam64.p, ml, f.pivot64.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
    if (f.n_big64) {
      k_sel64b<float><<<std::min<uint32_t>(f.n_big64, 4096), 1024, 0, s>>>(
          f.n_big64, f.big64.p, f.big64_off.p, pl->mj.p, f.col64.p, cval32, c.update_with_mean,
          c.mean_percentile, f.big64_scr.p, pl->lam64.p, ml, nullptr, nullptr, nullptr);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    }
    if (!map) {  // the forward pass reads the float mirror
      k_mirror32<<<grid_for(pl->nv, 256), 256, 0, s>>>(pl->nv, pl->lam64.p, pl->lam32.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    }
  }
  if (map) {
    if (int rc = map_step(pl, pl->lam64.p, pl->ml64.p, pl->tmp64.p)) return rc;
    k_commit<<<grid_for(pl->nv, 256), 256, 0, s>>>(pl->nv, pl->tmp64.p, pl->lam64.p, pl->lam32.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
  }
  k_cost_push<<<1, 256, 0, s>>>(pl->n_scans, pl->cost_acc.p, f.hist.p, f.hist_n.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  if (ev) MUTO_CHECK_HIP(ctx, hipEventRecord(ev[4], s));
  return MUTO_OK;
}

// The iteration as a graph, captured once per plan (again when the cost
// history grows or the MAP flag changes).
static int fp32c_graph(muto_em_plan* pl) {
  Fp32cData& f = pl->f32;
  muto_ctx* ctx = pl->ctx;
  if (f.graph && f.graph_map == pl->is_map && f.graph_mean == mean_all(pl) && f.graph_brk == fp32c_brackets(pl))
    return MUTO_OK;
  if (f.graph) (void)hipGraphExecDestroy(f.graph);
  f.graph = nullptr;
  hipGraph_t g = nullptr;
  MUTO_CHECK_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  const int rc = fp32c_enqueue(pl, nullptr);
  const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  if (rc != MUTO_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  MUTO_CHECK_HIP(ctx, e);
  const hipError_t ei = hipGraphInstantiate(&f.graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  MUTO_CHECK_HIP(ctx, ei);
  f.graph_map = pl->is_map;
  f.graph_mean = mean_all(pl);
  f.graph_brk = fp32c_brackets(pl);
  return MUTO_OK;
}

// `iters` FP32C iterations; cost sums per iteration (n_scans per iteration,
// scan fastest) into sums[], *done the
// iterations run.  Early stop (MutoMLSD.cpp:203-210) needs each cost before
// the next iteration starts: then every iteration is synchronised,
// otherwise all are enqueued back to back and the costs read once.
// Per-kernel timing (pl->timing) runs without the graph, with events.
int fp32c_run(muto_em_plan* pl, int iters, double* sums, int* done) {
  Fp32cData& f = pl->f32;
  muto_ctx* ctx = pl->ctx;
  hipStream_t s = ctx->stream;
  *done = 0;
  if (pl->is_map && !pl->tmp64.p) MUTO_CHECK_HIP(ctx, pl->tmp64.alloc(pl->nv));
  const size_t ns = pl->n_scans;
  if (f.hist_cap < (size_t)iters * ns) {
    if (f.graph) (void)hipGraphExecDestroy(f.graph);
    f.graph = nullptr;
    f.hist.release();
    MUTO_CHECK_HIP(ctx, f.hist.alloc((size_t)iters * ns));
    f.hist_cap = (size_t)iters * ns;
  }
  if (!f.hist_n.p) MUTO_CHECK_HIP(ctx, f.hist_n.alloc(1));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(f.hist_n.p, 0, sizeof(unsigned int), s));
  if (!pl->is_map && iters > 0) {
    k_zero_uncrossed<<<grid_for(pl->nv, 256), 256, 0, s>>>(pl->nv, pl->mj.p, pl->lam64.p, pl->lam32.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
  }
  if (!mean_all(pl)) {
    // median / trimmed mean: the FP64 stream's voxel-major slots and per-XCD
    // queues (built once per stream), float values in its c_ij array; every
    // run starts without pivots
    if (int rc = fp64s_slots(pl)) return rc;
    if (!f.trk_g64.p) MUTO_CHECK_HIP(ctx, f.trk_g64.alloc(4 * pl->n_trk));
    if (f.grid_bwd32s == 0) {
      int ob = 1;
      MUTO_CHECK_HIP(ctx, by_block(f.bs.block, [&](auto B) {
        constexpr uint32_t BS = decltype(B)::value;
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&ob, k_bwd32s<BS>, BS, 0);
      }));
      f.grid_bwd32s = std::max(1, ob) * ctx->n_cu;
    }
    k_fill_nan<<<grid_for(pl->nv, 256), 256, 0, s>>>(pl->nv, f.pivot64.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
    if (int rc = med_setup(pl, false)) return rc;
  }
  const bool graph = pl->use_graph && !pl->timing;
  if (graph)
    if (int rc = fp32c_graph(pl)) return rc;
  const bool each = pl->cfg.early_stop || pl->timing;
  double tsum[5] = {0, 0, 0, 0, 0};
  double prev = INFINITY;
  for (int it = 0; it < iters; ++it) {
    if (graph)
      MUTO_CHECK_HIP(ctx, hipGraphLaunch(f.graph, s));
    else if (int rc = fp32c_enqueue(pl, pl->timing ? pl->ev : nullptr))
      return rc;
    if (!each) continue;
    MUTO_CHECK_HIP(ctx, hipMemcpyAsync(&sums[it * ns], f.hist.p + it * ns, ns * sizeof(double),
                                       hipMemcpyDeviceToHost, s));
    MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
    *done = it + 1;
    if (pl->timing) {
      float t[4];
      for (int k = 0; k < 4; ++k) MUTO_CHECK_HIP(ctx, hipEventElapsedTime(&t[k], pl->ev[k], pl->ev[k + 1]));
      for (int k = 0; k < 4; ++k) tsum[k] += t[k];
      tsum[4] += t[0] + t[1] + t[2] + t[3];
    }
    const double c = sums[it * ns] / (double)pl->n_trk;  // early stop: single scan
    if (pl->cfg.early_stop && prev < c) break;
    prev = c;
  }
  if (!each && iters > 0) {
    MUTO_CHECK_HIP(ctx, hipMemcpyAsync(sums, f.hist.p, iters * ns * sizeof(double), hipMemcpyDeviceToHost, s));
    MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
    *done = iters;
  }
  if (pl->timing && *done)
    for (int k = 0; k < 5; ++k) pl->t_ms[k] = tsum[k] / *done;
  return MUTO_OK;
}

// ---------------------------------------------------------------------
// FP64 stream (MUTO_OPT_FP64_KERNEL 2): host side
// ---------------------------------------------------------------------
static Args64 args64(muto_em_plan* pl) {
  Fp32cData& f = pl->f32;
  Args64 b;
  b.len64 = f.len64.p;
  b.gtail64 = f.gtail64.p;
  b.seg_tail64 = f.seg_tail64.p;
  b.slot = f.slot64.p;
  b.cval = f.cval.p;
  b.qtask = f.qtask64.p;
  b.qoff = f.qoff64.p;
  b.qwork = f.qwork64.p;
  b.nq = f.nq64;
  return b;
}

// 256- or 512-thread blocks, FP64 brick and tables within one workgroup's LDS
static bool fp64s_fits(muto_ctx* ctx, const BrickShape& b) {
  if (b.block != 256 && b.block != 512) return false;
  int max_lds = 0;
  if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) != hipSuccess)
    return false;
  const size_t fwd = fwd64_static_lds(b.block) + (size_t)b.vol() * sizeof(double);
  const size_t bwd = bwd64_static_lds(b.block) + (size_t)b.vol() * sizeof(double);
  return std::max(fwd, bwd) <= (size_t)max_lds;
}

// the brick stream with its FP64 side arrays (rebuilt with them when the
// stream was built for FP32C alone)
int fp64s_build(muto_em_plan* pl) {
  Fp32cData& f = pl->f32;
  muto_ctx* ctx = pl->ctx;
  if (f.built && f.with64 && f.grid_fwd64 > 0) return MUTO_OK;
  if (!fp64s_fits(ctx, f.bs)) {
    ctx->set_error("FP64 stream: the brick shape needs 256- or 512-thread blocks and its FP64 brick "
                   "array and tables within one workgroup's LDS");
    return MUTO_ERR_LIMIT;
  }
  if (f.built && !f.with64) f.reset();
  f.with64 = true;
  if (int rc = fp32c_build(pl)) return rc;
  const size_t lds = (size_t)f.bs.vol() * sizeof(double);
  int of = 1, ob = 1;
  MUTO_CHECK_HIP(ctx, by_block64(f.bs.block, [&](auto B) {
    constexpr uint32_t BS = decltype(B)::value;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&of, k_fwd64s<BS>, BS, lds);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&ob, k_bwd64s<1, BS>, BS, lds);
  }));
  f.grid_fwd64 = std::max(1, of) * ctx->n_cu;
  f.grid_bwd64 = std::max(1, ob) * ctx->n_cu;
  MUTO_CHECK_HIP(ctx, f.cost_part64.alloc(grid_for(pl->n_trk, 256)));
  return MUTO_OK;
}

// voxel-major slots of the per-entry c_ij (median, trimmed mean), and the
// voxels whose lists are too long for one wavefront
static int fp64s_slots(muto_em_plan* pl) {
  Fp32cData& f = pl->f32;
  if (f.slots64) return MUTO_OK;
  muto_ctx* ctx = pl->ctx;
  hipStream_t s = ctx->stream;
  if (f.n_ent >= 0xFFFFFFFFull) {
    ctx->set_error("FP64 stream order statistics: more than 2^32 - 1 entries");
    return MUTO_ERR_LIMIT;
  }
  const size_t nv = pl->nv;
  std::vector<uint32_t> h(nv);
  MUTO_CHECK_HIP(ctx, copy_sync(ctx, h.data(), pl->mj.p, nv * sizeof(uint32_t), hipMemcpyDeviceToHost));
  // first slot of each voxel, the voxels taken brick by brick (so that
  // the entries of one brick, which one backprojection task writes, land in
  // one contiguous range); col[nv] is the total
  std::vector<uint64_t> col(nv + 1);
  {
    const uint32_t n0 = pl->g.n[0], n1 = pl->g.n[1], n2 = pl->g.n[2];
    const uint32_t b0 = f.bs.bx, b1 = f.bs.by, b2 = f.bs.bz;
    uint64_t run = 0;
    for (uint32_t oz = 0; oz < n2; oz += b2)
      for (uint32_t oy = 0; oy < n1; oy += b1)
        for (uint32_t ox = 0; ox < n0; ox += b0)
          for (uint32_t z = oz; z < std::min(oz + b2, n2); ++z)
            for (uint32_t y = oy; y < std::min(oy + b1, n1); ++y)
              for (uint32_t x = ox; x < std::min(ox + b0, n0); ++x) {
                const size_t j = (size_t)x + (size_t)n0 * ((size_t)y + (size_t)n1 * z);
                col[j] = run;
                run += h[j];
              }
    col[nv] = run;
  }
  MUTO_CHECK_HIP(ctx, f.col64.alloc(nv + 1));
  MUTO_CHECK_HIP(ctx, copy_sync(ctx, f.col64.p, col.data(), (nv + 1) * sizeof(uint64_t), hipMemcpyHostToDevice));
  MUTO_CHECK_HIP(ctx, f.slot64.alloc(f.n_ent_pad + GE));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(f.slot64.p, 0xFF, (f.n_ent_pad + GE) * sizeof(uint32_t), s));
  MUTO_CHECK_HIP(ctx, f.cval.alloc(std::max<uint64_t>(col[nv], 1)));
  MUTO_CHECK_HIP(ctx, f.pivot64.alloc(nv));
  {
    DBuf<uint32_t> fill;
    MUTO_CHECK_HIP(ctx, fill.alloc(nv));
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(fill.p, 0, nv * sizeof(uint32_t), s));
    if (f.n_tasks) {
      const IterArgs a = iter_args(pl);
      MUTO_CHECK_HIP(ctx, by_block(f.bs.block, [&](auto B) {
        constexpr uint32_t BS = decltype(B)::value;
        k_slot64s<BS><<<f.n_tasks, BS, 0, s>>>(a, f.col64.p, fill.p, f.slot64.p);
        return hipGetLastError();
      }));
    }
    MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  }
  std::vector<uint32_t> ids;
  std::vector<uint64_t> off;
  uint64_t tot = 0;
  for (size_t j = 0; j < nv; ++j) {
    if (h[j] <= 256) continue;
    uint64_t n = 1;
    while (n < h[j]) n <<= 1;
    ids.push_back((uint32_t)j);
    off.push_back(n > 4096 ? tot : 0);
    if (n > 4096) tot += n;
  }
  {
    // per-XCD task queues for the order statistics' backprojection: each
    // brick's chunks cut into tasks of QCHUNKS chunks, the bricks dealt to
    // NQ queues (largest first, to the least loaded), a queue's tasks brick
    // by brick
    constexpr uint32_t NQ = 8, QCHUNKS = 1;
    std::vector<Task> tk(f.n_tasks);
    if (f.n_tasks)
      MUTO_CHECK_HIP(ctx, copy_sync(ctx, tk.data(), f.tasks.p, f.n_tasks * sizeof(Task), hipMemcpyDeviceToHost));
    std::vector<Chunk> ch(f.n_chunks);
    if (f.n_chunks)
      MUTO_CHECK_HIP(ctx, copy_sync(ctx, ch.data(), f.chunks.p, f.n_chunks * sizeof(Chunk), hipMemcpyDeviceToHost));
    struct BR { uint32_t brick, c0, c1; uint64_t ne; };
    std::vector<BR> br;
    {
      std::sort(tk.begin(), tk.end(), [](const Task& x, const Task& y) {
        return x.brick != y.brick ? x.brick < y.brick : x.c0 < y.c0;
      });
      for (const Task& t : tk) {
        if (!br.empty() && br.back().brick == t.brick) {
          br.back().c0 = std::min(br.back().c0, t.c0);
          br.back().c1 = std::max(br.back().c1, t.c1);
        } else {
          br.push_back({t.brick, t.c0, t.c1, 0});
        }
      }
      for (BR& x : br)
        for (uint32_t c = x.c0; c < x.c1; ++c) x.ne += ch[c].ne;
    }
    std::vector<size_t> order(br.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return br[x].ne > br[y].ne; });
    std::vector<std::vector<Task>> q(NQ);
    std::vector<uint64_t> load(NQ, 0);
    for (size_t k : order) {
      const uint32_t qi = (uint32_t)(std::min_element(load.begin(), load.end()) - load.begin());
      load[qi] += br[k].ne;
      for (uint32_t c = br[k].c0; c < br[k].c1; c += QCHUNKS)
        q[qi].push_back(Task{br[k].brick, 1u, c, std::min(c + QCHUNKS, br[k].c1)});
    }
    std::vector<Task> all;
    std::vector<uint32_t> off(NQ + 1, 0);
    for (uint32_t qi = 0; qi < NQ; ++qi) {
      off[qi] = (uint32_t)all.size();
      all.insert(all.end(), q[qi].begin(), q[qi].end());
    }
    off[NQ] = (uint32_t)all.size();
    f.nq64 = NQ;
    MUTO_CHECK_HIP(ctx, f.qtask64.alloc(all.size()));
    MUTO_CHECK_HIP(ctx, f.qoff64.alloc(NQ + 1));
    MUTO_CHECK_HIP(ctx, f.qwork64.alloc(NQ));
    if (!all.empty())
      MUTO_CHECK_HIP(ctx, copy_sync(ctx, f.qtask64.p, all.data(), all.size() * sizeof(Task), hipMemcpyHostToDevice));
    MUTO_CHECK_HIP(ctx, copy_sync(ctx, f.qoff64.p, off.data(), (NQ + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(f.qwork64.p, 0, NQ * sizeof(unsigned int), s));
  }
  f.n_big64 = (uint32_t)ids.size();
  MUTO_CHECK_HIP(ctx, f.big64.alloc(ids.size()));
  MUTO_CHECK_HIP(ctx, f.big64_off.alloc(off.size()));
  MUTO_CHECK_HIP(ctx, f.big64_scr.alloc(tot));
  if (!ids.empty()) {
    MUTO_CHECK_HIP(ctx, copy_sync(ctx, f.big64.p, ids.data(), ids.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    MUTO_CHECK_HIP(ctx, copy_sync(ctx, f.big64_off.p, off.data(), off.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
  }
  f.slots64 = true;
  return MUTO_OK;
}

// mean of all S (update_with_mean, mean_percentile 100) or an order statistic
static bool mean_all(const muto_em_plan* pl) {
  return pl->cfg.update_with_mean && pl->cfg.mean_percentile == 100;
}

// One FP64 stream iteration on the plan's stream, no host synchronisation
// and no memset (it is captured as a graph).  ev: kernel boundaries for
// per-phase timing (5 events), or null.
static int fp64s_enqueue(muto_em_plan* pl, hipEvent_t* ev) {
  Fp32cData& f = pl->f32;
  muto_ctx* ctx = pl->ctx;
  hipStream_t s = ctx->stream;
  const IterArgs a = iter_args(pl);
  const Args64 b = args64(pl);
  const ErrM E = error_matrix(pl->cfg);
  const bool map = pl->is_map, mean = mean_all(pl);
  const size_t lds = (size_t)f.bs.vol() * sizeof(double);
  k_iter_begin64<<<1, 64, 0, s>>>(pl->cost_acc.p, pl->dcount.p, pl->work.p, f.qwork64.p, b.nq);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  if (ev) MUTO_CHECK_HIP(ctx, hipEventRecord(ev[0], s));
  if (f.n_tasks) {
    MUTO_CHECK_HIP(ctx, by_block64(f.bs.block, [&](auto B) {
      constexpr uint32_t BS = decltype(B)::value;
      k_fwd64s<BS><<<f.grid_fwd64, BS, lds, s>>>(a, b, pl->lam64.p,
                                                 reinterpret_cast<double4*>(f.seg_part64.p));
      return hipGetLastError();
    }));
  }
  if (ev) MUTO_CHECK_HIP(ctx, hipEventRecord(ev[1], s));
  const unsigned nblk = grid_for(pl->n_trk, 256);
  k_track64s<<<nblk, 256, 0, s>>>(pl->n_trk, f.tseg_ptr.p, f.tseg_pos.p,
                                  reinterpret_cast<const double4*>(f.seg_part64.p), f.trk_d.p, E, 0,
                                  reinterpret_cast<double4*>(f.trk_g64.p), f.cost_part64.p, pl->dcount.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  k_sum_parts_f64<<<1, 256, 0, s>>>(f.cost_part64.p, nblk, pl->cost_acc.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  if (ev) MUTO_CHECK_HIP(ctx, hipEventRecord(ev[2], s));
  const bool brk = !mean && fp64s_brackets(pl);
  const size_t lds_med = (size_t)f.bs.vol() * (sizeof(float2) + sizeof(uint32_t));
  const double4* g64 = reinterpret_cast<const double4*>(f.trk_g64.p);
  if (brk) {
    // as the FP32C median: counts below / inside each bracket, values inside
    const uint64_t nbk = (uint64_t)f.nb[0] * f.nb[1] * f.nb[2];
    k_fill_u8<<<grid_for(nbk, 256), 256, 0, s>>>(nbk, f.bmiss_b.p, 0);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
    const MedArgs md = med_args(pl, pl->work.p + 2);
    MUTO_CHECK_HIP(ctx, by_block64(f.bs.block, [&](auto B) {
      constexpr uint32_t BS = decltype(B)::value;
      k_bwd64s<2, BS><<<f.grid_med64, BS, lds_med, s>>>(a, b, md, g64, nullptr);
      return hipGetLastError();
    }));
  } else if (f.n_tasks) {
    MUTO_CHECK_HIP(ctx, by_block64(f.bs.block, [&](auto B) {
      constexpr uint32_t BS = decltype(B)::value;
      if (mean)
        k_bwd64s<0, BS><<<f.grid_bwd64, BS, lds, s>>>(a, b, MedArgs{}, g64, pl->acc64.p);
      else
        k_bwd64s<1, BS><<<f.grid_bwd64, BS, lds, s>>>(a, b, MedArgs{}, g64, nullptr);
      return hipGetLastError();
    }));
  }
  if (ev) MUTO_CHECK_HIP(ctx, hipEventRecord(ev[3], s));
  double* ml = map ? pl->ml64.p : nullptr;
  if (brk) {
    const MedArgs md = med_args(pl, pl->work.p + 3);
    k_sel_brk<double><<<grid_for(pl->nv, 256), 256, 0, s>>>(a.G, pl->nv, pl->mj.p, md, pl->lam64.p, ml);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
    MUTO_CHECK_HIP(ctx, by_block64(f.bs.block, [&](auto B) {
      constexpr uint32_t BS = decltype(B)::value;
      k_bwd64s<3, BS><<<f.grid_med64, BS, lds_med, s>>>(a, b, md, g64, nullptr);
      return hipGetLastError();
    }));
    const unsigned g = (unsigned)std::min<size_t>(grid_for(pl->nv, 256), (size_t)ctx->n_cu * 16);
    k_sel_miss<double><<<g, 256, 0, s>>>(pl->nv, pl->mj.p, md, pl->lam64.p, ml);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
    if (f.n_big64) {
      k_sel64b<double><<<std::min<uint32_t>(f.n_big64, 4096), 1024, 0, s>>>(
          f.n_big64, f.big64.p, f.big64_off.p, pl->mj.p, f.col64.p, f.cval.p, 0, 100, f.big64_scr.p,
          pl->lam64.p, ml, f.bmiss_v.p, f.brk.p, f.bmed.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    }
  } else if (mean) {
    k_upd64s<<<grid_for(pl->nv, 256), 256, 0, s>>>(pl->nv, pl->mj.p, pl->acc64.p, pl->lam64.p, ml);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
  } else {
    const muto_mlsd_cfg& c = pl->cfg;
    const unsigned g = (unsigned)std::min<size_t>(grid_for(pl->nv, 4), (size_t)ctx->n_cu * 64);
    k_sel64w<double><<<g, 256, 0, s>>>(pl->nv, pl->mj.p, f.col64.p, f.cval.p, c.update_with_mean,
                               c.mean_percentile, pl->lam64.p, ml, f.pivot64.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
    if (f.n_big64) {
      k_sel64b<double><<<std::min<uint32_t>(f.n_big64, 4096), 1024, 0, s>>>(
          f.n_big64, f.big64.p, f.big64_off.p, pl->mj.p, f.col64.p, f.cval.p, c.update_with_mean,
          c.mean_percentile, f.big64_scr.p, pl->lam64.p, ml, nullptr, nullptr, nullptr);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    }
  }
  if (map) {
    if (int rc = map_step(pl, pl->lam64.p, pl->ml64.p, pl->tmp64.p)) return rc;
    k_copy64<<<grid_for(pl->nv, 256), 256, 0, s>>>(pl->nv, pl->tmp64.p, pl->lam64.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
  }
  k_cost_push<<<1, 256, 0, s>>>(1, pl->cost_acc.p, f.hist.p, f.hist_n.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  if (ev) MUTO_CHECK_HIP(ctx, hipEventRecord(ev[4], s));
  return MUTO_OK;
}

static int fp64s_graph(muto_em_plan* pl) {
  Fp32cData& f = pl->f32;
  muto_ctx* ctx = pl->ctx;
  const int mode = (pl->is_map ? 2 : 0) | (mean_all(pl) ? 1 : 0) | (fp64s_brackets(pl) ? 4 : 0);
  if (f.graph64 && f.graph64_mode == mode) return MUTO_OK;
  if (f.graph64) (void)hipGraphExecDestroy(f.graph64);
  f.graph64 = nullptr;
  hipGraph_t g = nullptr;
  MUTO_CHECK_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  const int rc = fp64s_enqueue(pl, nullptr);
  const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  if (rc != MUTO_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  MUTO_CHECK_HIP(ctx, e);
  const hipError_t ei = hipGraphInstantiate(&f.graph64, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  MUTO_CHECK_HIP(ctx, ei);
  f.graph64_mode = mode;
  return MUTO_OK;
}

// `iters` FP64 stream iterations (any update rule, MLSD or MAP): cost sums
// into sums[], *done the iterations run; early stop and per-phase timing as
// in fp32c_run.
int fp64s_run(muto_em_plan* pl, int iters, double* sums, int* done) {
  Fp32cData& f = pl->f32;
  muto_ctx* ctx = pl->ctx;
  hipStream_t s = ctx->stream;
  *done = 0;
  if (int rc = fp64s_build(pl)) return rc;
  if (!mean_all(pl)) {
    if (int rc = fp64s_slots(pl)) return rc;
    if (int rc = med_setup(pl, true)) return rc;
  }
  if (pl->is_map && !pl->tmp64.p) MUTO_CHECK_HIP(ctx, pl->tmp64.alloc(pl->nv));
  if (f.hist_cap < (size_t)iters) {
    if (f.graph) (void)hipGraphExecDestroy(f.graph);
    f.graph = nullptr;
    if (f.graph64) (void)hipGraphExecDestroy(f.graph64);
    f.graph64 = nullptr;
    f.hist.release();
    MUTO_CHECK_HIP(ctx, f.hist.alloc((size_t)iters));
    f.hist_cap = (size_t)iters;
  }
  if (!f.hist_n.p) MUTO_CHECK_HIP(ctx, f.hist_n.alloc(1));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(f.hist_n.p, 0, sizeof(unsigned int), s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(pl->acc64.p, 0, pl->nv * sizeof(double), s));
  if (!pl->is_map && iters > 0 && mean_all(pl)) {
    k_zero_uncrossed64<<<grid_for(pl->nv, 256), 256, 0, s>>>(pl->nv, pl->mj.p, pl->lam64.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
  }
  if (!mean_all(pl)) {  // every run starts without pivots: the same work, run to run
    k_fill_nan<<<grid_for(pl->nv, 256), 256, 0, s>>>(pl->nv, f.pivot64.p);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
  }
  const bool graph = pl->use_graph && !pl->timing;
  if (graph)
    if (int rc = fp64s_graph(pl)) return rc;
  const bool each = pl->cfg.early_stop || pl->timing;
  double tsum[5] = {0, 0, 0, 0, 0};
  double prev = INFINITY;
  for (int it = 0; it < iters; ++it) {
    if (graph)
      MUTO_CHECK_HIP(ctx, hipGraphLaunch(f.graph64, s));
    else if (int rc = fp64s_enqueue(pl, pl->timing ? pl->ev : nullptr))
      return rc;
    if (!each) continue;
    MUTO_CHECK_HIP(ctx, hipMemcpyAsync(&sums[it], f.hist.p + it, sizeof(double), hipMemcpyDeviceToHost, s));
    MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
    *done = it + 1;
    if (pl->timing) {
      float t[4];
      for (int k = 0; k < 4; ++k) MUTO_CHECK_HIP(ctx, hipEventElapsedTime(&t[k], pl->ev[k], pl->ev[k + 1]));
      for (int k = 0; k < 4; ++k) tsum[k] += t[k];
      tsum[4] += t[0] + t[1] + t[2] + t[3];
    }
    const double c = sums[it] / (double)pl->n_trk;
    if (pl->cfg.early_stop && prev < c) break;
    prev = c;
  }
  if (!each && iters > 0) {
    MUTO_CHECK_HIP(ctx, hipMemcpyAsync(sums, f.hist.p, iters * sizeof(double), hipMemcpyDeviceToHost, s));
    MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
    *done = iters;
  }
  if (pl->timing && *done)
    for (int k = 0; k < 5; ++k) pl->t_ms[k] = tsum[k] / *done;
  return MUTO_OK;
}

// MAP objective operator, FP32C: the forward pass of the EM iteration on
// `lam` (FP64 on the device; its float mirror is rewritten) and the track
// kernel with the gradient's convention (G = 0 where cost_i = 0); G' per
// track and the |c| bound are copied out for the backward call, the cost
// sum returned.
int fp32c_map_forward(muto_em_plan* pl, const double* lam, float* g_out,
                      unsigned long long* cmax_out, double* cost_sum) {
  Fp32cData& f = pl->f32;
  muto_ctx* ctx = pl->ctx;
  hipStream_t s = ctx->stream;
  if (!pl->lam32.p) MUTO_CHECK_HIP(ctx, pl->lam32.alloc(pl->nv));
  k_mirror32<<<grid_for(pl->nv, 256), 256, 0, s>>>(pl->nv, lam, pl->lam32.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  pl->cur_prec = pl->cur_prec < 0 ? -1 : MUTO_FP64;  // the EM's float mirror is stale now
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(pl->work.p, 0, 4 * sizeof(unsigned int), s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(f.cmax.p, 0, sizeof(unsigned long long), s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(pl->cost_acc.p, 0, sizeof(double), s));
  const IterArgs a = iter_args(pl);
  const ErrM E = error_matrix(pl->cfg);
  if (f.n_tasks)
    MUTO_CHECK_HIP(ctx, launch_fwd(f, a, pl->lam32.p, reinterpret_cast<float4*>(f.seg_part.p), s));
  k_track32<<<grid_for(pl->n_trk, 256), 256, 0, s>>>(pl->n_trk, f.tseg_ptr.p, f.tseg_pos.p,
                                                      reinterpret_cast<const float4*>(f.seg_part.p),
                                                      f.trk_d.p, E, f.tail_scale, f.lmax, f.ubound, 1,
                                                      f.trk_g.p, nullptr, f.cmax.p,
                                                      pl->n_scans > 1 ? f.trk_scan.p : nullptr,
                                                      pl->cost_acc.p, pl->dcount.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(g_out, f.trk_g.p, 4 * pl->n_trk * sizeof(float), hipMemcpyDeviceToDevice, s));
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(cmax_out, f.cmax.p, sizeof(unsigned long long), hipMemcpyDeviceToDevice, s));
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(cost_sum, pl->cost_acc.p, sizeof(double), hipMemcpyDeviceToHost, s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(pl->work.p, 0, 4 * sizeof(unsigned int), s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(f.cmax.p, 0, sizeof(unsigned long long), s));
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  return MUTO_OK;
}

// ... and the backprojection of that G' into acc (sum_i <W_ij, G_i>, FP64).
int fp32c_map_backward(muto_em_plan* pl, const float* g, const unsigned long long* cmax,
                       double* acc) {
  Fp32cData& f = pl->f32;
  muto_ctx* ctx = pl->ctx;
  hipStream_t s = ctx->stream;
  const IterArgs a = iter_args(pl);
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(acc, 0, pl->nv * sizeof(double), s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(pl->work.p, 0, 4 * sizeof(unsigned int), s));
  if (f.n_tasks) MUTO_CHECK_HIP(ctx, launch_bwd(f, a, g, cmax, acc, s));
  if (f.acc_mode == 1) {
    k_fx_to_f64<<<grid_for(pl->nv, 256), 256, 0, s>>>(pl->nv, acc, cmax, f.fx_bits);
    MUTO_CHECK_HIP(ctx, hipGetLastError());
  }
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(pl->work.p, 0, 4 * sizeof(unsigned int), s));
  return MUTO_OK;
}

// The MAP objective's passes on the FP64 stream: the forward pass with the
// caller's lambda, then the track kernel in its gradient mode, p^2 G per
// track in rank order into g_out (zero where cost_i is 0) and the cost sum
// to the host...
int fp64s_map_forward(muto_em_plan* pl, const double* lam, double* g_out, double* cost_sum) {
  Fp32cData& f = pl->f32;
  muto_ctx* ctx = pl->ctx;
  hipStream_t s = ctx->stream;
  if (int rc = fp64s_build(pl)) return rc;
  const IterArgs a = iter_args(pl);
  const Args64 b = args64(pl);
  const ErrM E = error_matrix(pl->cfg);
  const size_t lds = (size_t)f.bs.vol() * sizeof(double);
  k_iter_begin64<<<1, 64, 0, s>>>(pl->cost_acc.p, pl->dcount.p, pl->work.p, f.qwork64.p, b.nq);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  if (f.n_tasks) {
    MUTO_CHECK_HIP(ctx, by_block64(f.bs.block, [&](auto B) {
      constexpr uint32_t BS = decltype(B)::value;
      k_fwd64s<BS><<<f.grid_fwd64, BS, lds, s>>>(a, b, lam, reinterpret_cast<double4*>(f.seg_part64.p));
      return hipGetLastError();
    }));
  }
  const unsigned nblk = grid_for(pl->n_trk, 256);
  k_track64s<<<nblk, 256, 0, s>>>(pl->n_trk, f.tseg_ptr.p, f.tseg_pos.p,
                                  reinterpret_cast<const double4*>(f.seg_part64.p), f.trk_d.p, E, 1,
                                  reinterpret_cast<double4*>(g_out), f.cost_part64.p, pl->dcount.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  k_sum_parts_f64<<<1, 256, 0, s>>>(f.cost_part64.p, nblk, pl->cost_acc.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(cost_sum, pl->cost_acc.p, sizeof(double), hipMemcpyDeviceToHost, s));
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  return MUTO_OK;
}

// ... and the backprojection of that p^2 G into acc (sum_i <W_ij, p^2 G_i>
// with the exact FP64 W_ij, FP64 sums per brick in LDS; acc is cleared
// first: bricks that no task covers are not written)
int fp64s_map_backward(muto_em_plan* pl, const double* g, double* acc) {
  Fp32cData& f = pl->f32;
  muto_ctx* ctx = pl->ctx;
  hipStream_t s = ctx->stream;
  const IterArgs a = iter_args(pl);
  const Args64 b = args64(pl);
  const size_t lds = (size_t)f.bs.vol() * sizeof(double);
  k_iter_begin64<<<1, 64, 0, s>>>(pl->cost_acc.p, pl->dcount.p, pl->work.p, f.qwork64.p, b.nq);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(acc, 0, pl->nv * sizeof(double), s));
  if (f.n_tasks) {
    MUTO_CHECK_HIP(ctx, by_block64(f.bs.block, [&](auto B) {
      constexpr uint32_t BS = decltype(B)::value;
      k_bwd64s<0, BS><<<f.grid_bwd64, BS, lds, s>>>(a, b, MedArgs{}, reinterpret_cast<const double4*>(g), acc);
      return hipGetLastError();
    }));
  }
  return MUTO_OK;
}

// the float mirror of the plan's FP64 image
int fp32c_mirror(muto_em_plan* pl) {
  muto_ctx* ctx = pl->ctx;
  if (!pl->lam32.p) MUTO_CHECK_HIP(ctx, pl->lam32.alloc(pl->nv));
  k_mirror32<<<grid_for(pl->nv, 256), 256, 0, ctx->stream>>>(pl->nv, pl->lam64.p, pl->lam32.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  return MUTO_OK;
}

// Brick-shape autotuner: for each candidate shape, rebuilds the stream and
// times `iters` FP32C iterations (HIP events, after one untimed iteration)
// from the plan's current image, which is restored after every trial; keeps
// the fastest shape built.  ms[k] receives each candidate's ms per iteration
// (shapes as bx | by << 8 | bz << 16 in shapes[k]).
int fp32c_autotune(muto_em_plan* pl, int iters, int n_cand, const uint32_t* cand, double* ms) {
  muto_ctx* ctx = pl->ctx;
  hipStream_t s = ctx->stream;
  if (iters < 1) iters = 1;
  DBuf<double> save;
  MUTO_CHECK_HIP(ctx, save.alloc(pl->nv));
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(save.p, pl->lam64.p, pl->nv * sizeof(double), hipMemcpyDeviceToDevice, s));
  const BrickShape keep = pl->f32.bs;
  const int timing = pl->timing;
  double best = INFINITY;
  int best_k = -1;
  std::vector<double> sums((size_t)(iters + 1) * pl->n_scans);
  for (int k = 0; k < n_cand; ++k) {
    const BrickShape b = BrickShape::from_code(cand[k]);
    ms[k] = INFINITY;
    if (b.bx == 0 || b.by == 0 || b.bz == 0 || !fp32c_fits(ctx, b)) continue;
    pl->f32.reset();
    pl->f32.bs = b;
    if (int rc = fp32c_build(pl)) return rc;
    MUTO_CHECK_HIP(ctx, hipMemcpyAsync(pl->lam64.p, save.p, pl->nv * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (int rc = fp32c_mirror(pl)) return rc;
    int done = 0;
    pl->timing = 0;
    if (int rc = fp32c_run(pl, 1, sums.data(), &done)) return rc;  // warm (graph capture)
    pl->timing = 1;
    if (int rc = fp32c_run(pl, iters, sums.data(), &done)) return rc;
    ms[k] = pl->t_ms[4];
    if (ms[k] < best) {
      best = ms[k];
      best_k = k;
    }
  }
  pl->timing = timing;
  pl->f32.reset();
  if (best_k >= 0) {
    pl->f32.bs = BrickShape::from_code(cand[best_k]);
  } else {
    pl->f32.bs = keep;
  }
  if (int rc = fp32c_build(pl)) return rc;
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(pl->lam64.p, save.p, pl->nv * sizeof(double), hipMemcpyDeviceToDevice, s));
  if (int rc = fp32c_mirror(pl)) return rc;
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  return MUTO_OK;
}

}  // namespace muto
