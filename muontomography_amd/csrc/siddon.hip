// Batched Siddon tracing into CSR: count pass, exclusive scan, fill pass.
// One lane per segment: the alpha merge is a sequential recurrence
// (repeated addition, src/MutoSiddon.cpp:108-138) and is kept bit-exact.
#include "geom_dev.hpp"

namespace muto {

namespace {

__global__ void __launch_bounds__(256)
k_siddon_count(GridD g, const double* __restrict__ p1,
               const double* __restrict__ p2, size_t n,
               uint64_t* __restrict__ counts) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a[3] = {p1[3 * i], p1[3 * i + 1], p1[3 * i + 2]};
  double b[3] = {p2[3 * i], p2[3 * i + 1], p2[3 * i + 2]};
  counts[i] = siddon_walk(g, a, b, [](uint32_t, uint32_t, uint32_t, double) {});
}

// Fill pass: each lane walks its own segment, up to FILL_STAGE recorded
// voxels at a time, into an LDS stage (slot k of thread t at
// k * FILL_BS + t); then the wavefront writes every lane's staged run
// together, FILL_STAGE lanes per source lane, so that each store instruction
// writes 64 / FILL_STAGE runs of consecutive entries instead of 64 scattered
// words (the lanes' lists lie far apart).  A window ends where the lane's
// output position reaches a multiple of FILL_STAGE, and its entries sit at
// their output position modulo FILL_STAGE: every window after a lane's
// first covers whole aligned runs (64 B of each u32 array, 128 B of the
// lengths), so the caches see whole lines instead of two partial ones per
// run (2e6 in-legs at 200^3: 3.7 -> 2.4 ms).
// PACK: the voxel is staged as x | y << 11 | z << 22 (12 B per staged entry
// instead of 20, so more wavefronts fit the LDS).  A recorded index outside
// 11/11/10 bits sets *unpackable and the host runs the unpacked form again,
// so the lists are exact for every input.
// Every lane of a wavefront takes part in every flush (the loop runs while
// any lane's walk is live).
constexpr int FILL_STAGE = 16;
constexpr uint32_t FILL_BS = 64;
template <bool PACK>
__global__ void __launch_bounds__(FILL_BS)
k_siddon_fill(GridD g, const double* __restrict__ p1,
              const double* __restrict__ p2, size_t n,
              const uint64_t* __restrict__ row_ptr, uint32_t* __restrict__ vx,
              uint32_t* __restrict__ vy, uint32_t* __restrict__ vz,
              double* __restrict__ len, uint32_t* __restrict__ unpackable) {
  constexpr int NW = PACK ? 1 : 3;  // staged voxel words per entry
  __shared__ uint32_t st_v[NW * FILL_STAGE * FILL_BS];
  __shared__ double st_l[FILL_STAGE * FILL_BS];
  const uint32_t tid = threadIdx.x;
  const int lane = (int)(tid & 63u);
  const uint32_t wbase = tid & ~63u;
  const size_t i = (size_t)blockIdx.x * blockDim.x + tid;
  SiddonCursor c;
  bool live = false;
  uint64_t o = 0, e = 0;
  uint32_t wide = 0;  // PACK: x | y >> 1 | z of every recorded voxel, in bits 10 and up when one does not fit
  if (i < n) {
    const double a[3] = {p1[3 * i], p1[3 * i + 1], p1[3 * i + 2]};
    const double b[3] = {p2[3 * i], p2[3 * i + 1], p2[3 * i + 2]};
    o = row_ptr[i];
    e = row_ptr[i + 1];
    live = c.begin(g, a, b);
  }
  while (__any(live)) {
    // the window: output positions [o - sh, o - sh + FILL_STAGE), this
    // lane's entries from slot sh on
    const int sh = (int)(o % (uint64_t)FILL_STAGE);
    int cnt = 0;
    if (live) {
      c.run((uint32_t)(FILL_STAGE - sh), [&](uint32_t x, uint32_t y, uint32_t z, double L) {
        const uint32_t sl = (uint32_t)(sh + cnt) * FILL_BS + tid;
        if (PACK) {
          wide |= ((x | y) >> 1) | z;
          st_v[sl] = x | (y << 11) | (z << 22);
        } else {
          st_v[sl] = x;
          st_v[FILL_STAGE * FILL_BS + sl] = y;
          st_v[2 * FILL_STAGE * FILL_BS + sl] = z;
        }
        st_l[sl] = L;
        ++cnt;
      });
      live = c.live;
    }
    // slots [sh, hi) hold entries, never past the end of the lane's list
    const uint64_t rem = e > o ? e - o : 0;
    const int hi = sh + ((uint64_t)cnt < rem ? cnt : (int)rem);
    const uint64_t ob = o - (uint64_t)sh;
    // the wavefront's flush, FILL_STAGE rounds of R = 64 / FILL_STAGE source
    // lanes: source lane src's slot k is written by lane
    // (src % R) * FILL_STAGE + k of round src / R
#pragma unroll
    for (int r = 0; r < FILL_STAGE; ++r) {
      const int src = r * (64 / FILL_STAGE) + lane / FILL_STAGE, k = lane % FILL_STAGE;
      const int kse = __shfl(sh | (hi << 8), src, 64), ks = kse & 255, ke = kse >> 8;
      const uint64_t os = (uint64_t)(uint32_t)__shfl((int)(uint32_t)ob, src, 64) |
                          ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(ob >> 32), src, 64) << 32);
      if (k >= ks && k < ke) {
        const uint32_t sl = (uint32_t)k * FILL_BS + wbase + (uint32_t)src;
        uint32_t X, Y, Z;
        if (PACK) {
          const uint32_t w = st_v[sl];
          X = w & 2047u;
          Y = (w >> 11) & 2047u;
          Z = w >> 22;
        } else {
          X = st_v[sl];
          Y = st_v[FILL_STAGE * FILL_BS + sl];
          Z = st_v[2 * FILL_STAGE * FILL_BS + sl];
        }
        vx[os + k] = X;
        if (vy) vy[os + k] = Y;
        if (vz) vz[os + k] = Z;
        if (len) len[os + k] = st_l[sl];
      }
    }
    o += (uint64_t)cnt;
  }
  if (PACK && (wide >> 10) != 0u) atomicOr(unpackable, 1u);
}

// Per-segment count and digest of the list: FNV-1a over 64-bit words, three
// per entry (x | y << 32, z, length bits), in list order.
__global__ void __launch_bounds__(256)
k_siddon_digest(GridD g, const double* __restrict__ p1,
                const double* __restrict__ p2, size_t n,
                uint32_t* __restrict__ count, uint64_t* __restrict__ digest) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a[3] = {p1[3 * i], p1[3 * i + 1], p1[3 * i + 2]};
  double b[3] = {p2[3 * i], p2[3 * i + 1], p2[3 * i + 2]};
  uint64_t h = 0xcbf29ce484222325ull;
  const uint64_t prime = 0x100000001b3ull;
  count[i] = siddon_walk(g, a, b, [&](uint32_t x, uint32_t y, uint32_t z, double L) {
    h = (h ^ ((uint64_t)x | ((uint64_t)y << 32))) * prime;
    h = (h ^ (uint64_t)z) * prime;
    h = (h ^ (uint64_t)__double_as_longlong(L)) * prime;
  });
  digest[i] = h;
}

}  // namespace

hipError_t siddon_count(const GridD& g, const double* p1, const double* p2,
                        size_t n, uint64_t* counts, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_siddon_count<<<grid_for(n, 256), 256, 0, s>>>(g, p1, p2, n, counts);
  return hipGetLastError();
}

hipError_t siddon_fill(const GridD& g, const double* p1, const double* p2,
                       size_t n, const uint64_t* row_ptr, uint32_t* vx,
                       uint32_t* vy, uint32_t* vz, double* len,
                       uint32_t* unpackable, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (unpackable)
    k_siddon_fill<true><<<grid_for(n, FILL_BS), FILL_BS, 0, s>>>(g, p1, p2, n, row_ptr, vx, vy, vz,
                                                                 len, unpackable);
  else
    k_siddon_fill<false><<<grid_for(n, FILL_BS), FILL_BS, 0, s>>>(g, p1, p2, n, row_ptr, vx, vy, vz,
                                                                  len, nullptr);
  return hipGetLastError();
}

}  // namespace muto

using namespace muto;

extern "C" int muto_siddon_csr(muto_ctx* ctx, const muto_grid* grid, size_t n,
                               const double* p1, const double* p2,
                               uint64_t* row_ptr, uint32_t* vx, uint32_t* vy,
                               uint32_t* vz, double* len) {
  if (!ctx) return MUTO_ERR_ARG;
  if (!grid || !row_ptr || (n && (!p1 || !p2))) {
    ctx->set_error("muto_siddon_csr: null argument");
    return MUTO_ERR_ARG;
  }
  if (grid->nx == 0 || grid->ny == 0 || grid->nz == 0) {
    ctx->set_error("muto_siddon_csr: empty grid");
    return MUTO_ERR_ARG;
  }
  const GridD g = to_dev(*grid);
  hipStream_t s = ctx->stream;
  DBuf<double> dp1, dp2;
  DBuf<uint64_t> dcnt, drp;
  MUTO_CHECK_HIP(ctx, dp1.alloc(3 * n));
  MUTO_CHECK_HIP(ctx, dp2.alloc(3 * n));
  MUTO_CHECK_HIP(ctx, dcnt.alloc(n + 1));
  MUTO_CHECK_HIP(ctx, drp.alloc(n + 1));
  if (n) {
    MUTO_CHECK_HIP(ctx, hipMemcpyAsync(dp1.p, p1, 3 * n * sizeof(double),
                                       hipMemcpyHostToDevice, s));
    MUTO_CHECK_HIP(ctx, hipMemcpyAsync(dp2.p, p2, 3 * n * sizeof(double),
                                       hipMemcpyHostToDevice, s));
  }
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(dcnt.p, 0, (n + 1) * sizeof(uint64_t), s));
  MUTO_CHECK_HIP(ctx, siddon_count(g, dp1.p, dp2.p, n, dcnt.p, s));
  MUTO_CHECK_HIP(ctx, exclusive_scan_u64(dcnt.p, drp.p, n + 1, s));
  if (!vx) {
    MUTO_CHECK_HIP(ctx, hipMemcpyAsync(row_ptr, drp.p, (n + 1) * sizeof(uint64_t),
                                       hipMemcpyDeviceToHost, s));
    MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
    return MUTO_OK;
  }
  uint64_t nnz = 0;
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(&nnz, drp.p + n, sizeof(uint64_t),
                                     hipMemcpyDeviceToHost, s));
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  if (row_ptr[n] != nnz) {
    ctx->set_error("muto_siddon_csr: row_ptr does not match the count pass");
    return MUTO_ERR_ARG;
  }
  DBuf<uint32_t> dx, dy, dz;
  DBuf<double> dl;
  MUTO_CHECK_HIP(ctx, dx.alloc(nnz));
  if (vy) MUTO_CHECK_HIP(ctx, dy.alloc(nnz));
  if (vz) MUTO_CHECK_HIP(ctx, dz.alloc(nnz));
  if (len) MUTO_CHECK_HIP(ctx, dl.alloc(nnz));
  // packed staging when every in-grid index fits 11/11/10 bits (indices
  // outside that are caught by the kernel and filled again unpacked)
  const bool pack = g.n[0] <= 2047 && g.n[1] <= 2047 && g.n[2] <= 1023;
  DBuf<uint32_t> flag;
  if (pack) {
    MUTO_CHECK_HIP(ctx, flag.alloc(1));
    MUTO_CHECK_HIP(ctx, hipMemsetAsync(flag.p, 0, sizeof(uint32_t), s));
  }
  ctx->mark_begin();
  MUTO_CHECK_HIP(ctx, siddon_fill(g, dp1.p, dp2.p, n, drp.p, dx.p,
                                  vy ? dy.p : nullptr, vz ? dz.p : nullptr,
                                  len ? dl.p : nullptr, pack ? flag.p : nullptr, s));
  ctx->mark_end();
  if (pack) {
    uint32_t wide = 0;
    MUTO_CHECK_HIP(ctx, hipMemcpyAsync(&wide, flag.p, sizeof wide, hipMemcpyDeviceToHost, s));
    MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
    if (wide) {
      ctx->mark_begin();
      MUTO_CHECK_HIP(ctx, siddon_fill(g, dp1.p, dp2.p, n, drp.p, dx.p,
                                      vy ? dy.p : nullptr, vz ? dz.p : nullptr,
                                      len ? dl.p : nullptr, nullptr, s));
      ctx->mark_end();
    }
  }
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(vx, dx.p, nnz * sizeof(uint32_t),
                                     hipMemcpyDeviceToHost, s));
  if (vy)
    MUTO_CHECK_HIP(ctx, hipMemcpyAsync(vy, dy.p, nnz * sizeof(uint32_t),
                                       hipMemcpyDeviceToHost, s));
  if (vz)
    MUTO_CHECK_HIP(ctx, hipMemcpyAsync(vz, dz.p, nnz * sizeof(uint32_t),
                                       hipMemcpyDeviceToHost, s));
  if (len)
    MUTO_CHECK_HIP(ctx, hipMemcpyAsync(len, dl.p, nnz * sizeof(double),
                                       hipMemcpyDeviceToHost, s));
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  return MUTO_OK;
}

extern "C" int muto_siddon_digest(muto_ctx* ctx, const muto_grid* grid, size_t n,
                                  const double* p1, const double* p2,
                                  uint32_t* count, uint64_t* digest) {
  if (!ctx) return MUTO_ERR_ARG;
  if (!grid || !count || !digest || (n && (!p1 || !p2))) {
    ctx->set_error("muto_siddon_digest: null argument");
    return MUTO_ERR_ARG;
  }
  if (grid->nx == 0 || grid->ny == 0 || grid->nz == 0) {
    ctx->set_error("muto_siddon_digest: empty grid");
    return MUTO_ERR_ARG;
  }
  if (n == 0) return MUTO_OK;
  const GridD g = to_dev(*grid);
  hipStream_t s = ctx->stream;
  DBuf<double> dp1, dp2;
  DBuf<uint32_t> dc;
  DBuf<uint64_t> dh;
  MUTO_CHECK_HIP(ctx, dp1.alloc(3 * n));
  MUTO_CHECK_HIP(ctx, dp2.alloc(3 * n));
  MUTO_CHECK_HIP(ctx, dc.alloc(n));
  MUTO_CHECK_HIP(ctx, dh.alloc(n));
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(dp1.p, p1, 3 * n * sizeof(double),
                                     hipMemcpyHostToDevice, s));
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(dp2.p, p2, 3 * n * sizeof(double),
                                     hipMemcpyHostToDevice, s));
  ctx->mark_begin();
  k_siddon_digest<<<grid_for(n, 256), 256, 0, s>>>(g, dp1.p, dp2.p, n, dc.p, dh.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  ctx->mark_end();
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(count, dc.p, n * sizeof(uint32_t),
                                     hipMemcpyDeviceToHost, s));
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(digest, dh.p, n * sizeof(uint64_t),
                                     hipMemcpyDeviceToHost, s));
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  return MUTO_OK;
}
