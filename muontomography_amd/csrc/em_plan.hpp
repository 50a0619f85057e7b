// EM plan: device-resident track-voxel lists and per-iteration buffers.
#pragma once

#include "muto_internal.hpp"

namespace muto {

// Brick (LDS tile) shape of the FP32C path.
// Brick (LDS tile) shape and the block size of the passes that process it.
// Option / tuner code: bx | by << 8 | bz << 16 | log2(block / 256) << 24.
struct BrickShape {
  uint32_t bx = 16, by = 16, bz = 16;
  uint32_t block = 256;
  uint32_t vol() const { return bx * by * bz; }
  uint32_t code() const {
    return bx | by << 8 | bz << 16 | (block == 1024 ? 2u : block == 512 ? 1u : 0u) << 24;
  }
  static BrickShape from_code(uint32_t c) {
    BrickShape b;
    b.bx = c & 0xFFu;
    b.by = (c >> 8) & 0xFFu;
    b.bz = (c >> 16) & 0xFFu;
    const uint32_t lb = (c >> 24) & 0xFFu;
    b.block = lb > 2 ? 0u : 256u << lb;  // 0: not a block size the passes take
    return b;
  }
  bool operator!=(const BrickShape& o) const {
    return bx != o.bx || by != o.by || bz != o.bz || block != o.block;
  }
};

// Chunk: up to `block` - 2 consecutive brick-major segments [q0, q0 + cnt) whose
// ne entries start at entry `base` (a multiple of 16; the entries are
// padded to whole 16-entry groups); one block processes it as one unit.
struct Chunk {
  uint64_t q0;
  uint64_t base;
  uint32_t cnt;
  uint32_t ne;
};

// A unit of work of the persistent FP32C kernels: chunks [c0, c1) that all
// lie in brick `brick`, which `ntask` tasks share (1: the task covers the
// whole brick and stores its sums; else the tasks add their fixed-point sums
// and the last one to finish converts them).
struct Task {
  uint32_t brick;
  uint32_t ntask;
  uint32_t c0, c1;
};

struct Fp32cData {
  bool built = false;
  BrickShape bs;
  uint32_t nb[3] = {0, 0, 0};
  uint64_t n_seg = 0, n_ent = 0, n_ent_pad = 0;
  DBuf<uint32_t> seg_trk;   // [n_seg] track rank
  DBuf<uint32_t> seg_tfx;   // [n_seg] tail after the segment, fixed point
  double tail_scale = 1.0;  // fixed-point scale of seg_tfx (a power of two)
  DBuf<uint32_t> ent;       // [n_ent_pad] packed entry: length code | voxel step << 26
  DBuf<uint16_t> seg_v0;    // [n_seg] first voxel of the segment inside its brick
  DBuf<uint16_t> gv0;       // [groups] voxel of the group's first entry
  double unit = 1.0;        // entry length unit U in mm (a power of two)
  DBuf<uint32_t> gdesc;     // [groups] segment-start mask | first segment << 16
  DBuf<float> gcarry;       // [groups] length of the last segment after the group
  DBuf<uint32_t> tseg_ptr;  // [n_trk+1] per rank: segment list offsets
  DBuf<uint32_t> tseg_pos;  // [n_seg] brick-major positions, rank order
  DBuf<double> trk_d;       // [4*n_trk] rank order: D00, D01, D11, p_r
  DBuf<float> seg_part;     // [4*n_seg] forward partial sums (float4, brick-major)
  DBuf<float> trk_g;        // [4*n_trk] p^2 G' (g00, g01, g11, basis t_c bits)
  DBuf<unsigned long long> cmax;  // [1] bound on |c_ij| (FP64 bits), per iteration
  DBuf<uint16_t> trk_scan;  // [n_trk] rank order: scan of each track (batches)
  double lmax = 0.0, ubound = 0.0;  // entry length bound; segment extent bound
  DBuf<Task> tasks;
  int fx_bits = 46;  // backprojection sums: int64 fixed point, |c| bound scaled to 2^fx_bits
  DBuf<Chunk> chunks;
  uint32_t n_tasks = 0, n_chunks = 0;
  int grid_fwd = 0, grid_bwd = 0, block = 256;
  int bwd_sub = 1;  // chunks per backprojection workgroup (1 or 2)
  int acc_mode = 1;  // accumulation: 1 FP64 / int64 fixed point, 0 FP32 two-sum
  bool acc_mode_set = false;  // set through the option call (else environment)
  hipGraphExec_t graph = nullptr;  // one iteration, captured
  bool graph_map = false;          // the MAP flag the graph was captured with
  bool graph_mean = true;          // ... and the update rule (mean, or an order statistic)
  int grid_bwd32s = 0;             // persistent grid of the order statistics' backprojection
  // FP64 stream (the brick stream with FP64 lengths and tails), built with
  // the stream when with64 is set
  bool with64 = false;
  DBuf<double> len64;       // [n_ent_pad] exact FP64 length of each entry (0: padding)
  DBuf<double> gtail64;     // [groups] exact tail of the group's entry 15
  DBuf<double> seg_tail64;  // [n_seg] exact tail of the segment's last entry
  DBuf<double> seg_part64;  // [4*n_seg] forward partial sums (FP64, global tail basis)
  DBuf<double> trk_g64;     // [4*n_trk] p^2 G (g00, g01, g11, -) per rank
  // its order statistics: per-entry c_ij in voxel-major slots
  bool slots64 = false;
  DBuf<uint32_t> slot64;    // [n_ent_pad] slot of each entry (~0u: padding)
  DBuf<uint64_t> col64;     // [nv+1] first slot of each voxel
  DBuf<double> cval;        // [nnz] c_ij in slots
  DBuf<double> pivot64;     // [nv] last pass's order statistic per voxel (NaN: none)
  DBuf<Task> qtask64;       // per-XCD task queues of its backprojection
  DBuf<uint32_t> qoff64;    // [nq64 + 1]
  DBuf<unsigned int> qwork64;  // [nq64]
  uint32_t nq64 = 1;
  uint32_t n_big64 = 0;     // voxels above the one-block sort
  DBuf<uint32_t> big64;     // [n_big64] their indices
  DBuf<uint64_t> big64_off; // [n_big64] their offsets in big64_scr
  DBuf<double> big64_scr;
  DBuf<double> cost_part64; // per-block cost sums of the FP64 stream's track kernel
  int grid_fwd64 = 0, grid_bwd64 = 0;  // persistent grids of its passes
  hipGraphExec_t graph64 = nullptr;  // one FP64 stream iteration, captured
  int graph64_mode = -1;             // the update rule / MAP flag it was captured with
  // FP32C median by brackets (MUTO_OPT_MEDIAN_BRACKETS): one task per
  // brick; per voxel a bracket of c values around its last median, the
  // counts below and inside it, up to MED_K values inside, and miss marks
  // (voxels, bricks) for the voxels whose median left the bracket
  bool med_brk = true;      // the option
  bool med_ready = false;   // the arrays below are built for this stream
  uint32_t max_mj = 0;      // largest M_j (brackets count in 16 bits)
  DBuf<Task> btask;
  uint32_t n_btask = 0;
  int grid_med = 0;
  DBuf<float2> brk;
  DBuf<float> bmed;
  DBuf<uint32_t> bcnt;
  DBuf<float> bbuf;         // FP32C values inside the brackets
  DBuf<double> bbuf64;      // FP64 stream values inside the brackets
  int grid_med64 = 0;
  DBuf<uint8_t> bmiss_v, bmiss_b;
  bool graph_brk = false;   // the bracket mode the graph was captured with
  DBuf<double> hist;               // [hist_cap] cost sum per iteration of a run
  DBuf<unsigned int> hist_n;       // [1] entries in hist
  size_t hist_cap = 0;

  // drops the built stream (the brick shape and accumulation mode stay)
  void reset() {
    if (graph) (void)hipGraphExecDestroy(graph);
    graph = nullptr;
    if (graph64) (void)hipGraphExecDestroy(graph64);
    graph64 = nullptr;
    graph64_mode = -1;
    len64.release();
    gtail64.release();
    seg_tail64.release();
    seg_part64.release();
    trk_g64.release();
    slots64 = false;
    grid_bwd32s = 0;
    slot64.release();
    col64.release();
    cval.release();
    pivot64.release();
    qtask64.release();
    qoff64.release();
    qwork64.release();
    nq64 = 1;
    n_big64 = 0;
    big64.release();
    big64_off.release();
    big64_scr.release();
    cost_part64.release();
    built = false;
    n_seg = n_ent = n_ent_pad = 0;
    n_tasks = n_chunks = 0;
    seg_trk.release();
    seg_tfx.release();
    ent.release();
    seg_v0.release();
    gv0.release();
    gdesc.release();
    gcarry.release();
    tseg_ptr.release();
    tseg_pos.release();
    trk_d.release();
    seg_part.release();
    trk_g.release();
    cmax.release();
    trk_scan.release();
    tasks.release();
    chunks.release();
    hist.release();
    hist_n.release();
    hist_cap = 0;
    med_ready = false;
    btask.release();
    n_btask = 0;
    grid_med = 0;
    brk.release();
    bmed.release();
    bcnt.release();
    bbuf.release();
    bbuf64.release();
    grid_med64 = 0;
    bmiss_v.release();
    bmiss_b.release();
  }
};

// The recomputing MAP backward's record of one run of a track's path
// entries inside one tile: the Siddon walk's state before the step that
// recorded the run's first voxel, and the path length before that entry.
struct MapRec {
  double ax[3];       // plane crossings before the step
  double alpha_last;
  double s;           // path length before the entry
  uint32_t x, y, z;   // the step's voxel coordinates
  uint32_t tl;        // track << 1 | leg (1: the out-leg of a kinked path)
};
static_assert(sizeof(MapRec) == 56, "tile record");

// records [q0, q1) of tile `tile`, which `ntask` tasks share
struct MapTask {
  uint32_t tile, q0, q1, ntask;
};

}  // namespace muto

struct muto_em_plan {
  muto_ctx* ctx = nullptr;
  muto_mlsd_cfg cfg{};
  bool is_map = false;
  muto::GridD g{};
  size_t nv = 0;
  size_t n_trk = 0;
  uint64_t nnz = 0;
  muto_counters counters{};

  // track-major (selection order) inputs of the FP64 path
  muto::DBuf<double> trk4;      // Sx, Sy, Dx, Dy
  muto::DBuf<double> pr;        // p_r
  muto::DBuf<uint64_t> row_ptr; // n_trk+1
  muto::DBuf<uint32_t> vox;     // linear voxel index
  muto::DBuf<double> len, tail;
  muto::DBuf<uint32_t> mj;      // entries per voxel
  muto::DBuf<double> geo;       // [9*n_trk] pin, pout, PoCA point (path recompute)
  muto::DBuf<double> cost_part; // per-block cost sums of the FP64 iteration
  uint32_t n_scans = 1;         // > 1: a batch of scans laid side by side in one grid
  muto::DBuf<uint16_t> trk_scan;  // [n_trk] scan of each track (batches only)
  bool has_lists = true;        // row_ptr/vox/len/tail hold the paths
  uint64_t list_budget = UINT64_MAX;  // bytes allowed for the lists (and, without them, the tile records)
  // the recomputing MAP backward's tile records, sorted by tile, made at
  // its first call: mt_state 0 not made yet, 1 made, -1 not kept (over the
  // budget, or no memory: the backward then adds one atomic per entry)
  int mt_state = 0;
  muto::DBuf<muto::MapRec> mt_rec;
  muto::DBuf<muto::MapTask> mt_task;
  muto::DBuf<unsigned int> mt_work;  // [1] task counter
  muto::DBuf<uint32_t> mt_orig;      // [n_rec] each sorted record's place in track order
  muto::DBuf<uint64_t> mt_roff;      // [n_trk+1] each track's records in track order
  muto::DBuf<double> mt_plen;        // [n_trk] path lengths (walk_path's)
  muto::DBuf<double> mt_part;        // [3 n_rec] the forward's moments per record, track order
  uint64_t mt_n_rec = 0;
  uint32_t mt_n_task = 0;
  int mt_grid = 0, mt_grid_fwd = 0;
  // mj counted: by the FP32C stream build when that comes first (LDS counts
  // per brick), else by one atomic per entry (ensure_mj)
  bool mj_ready = false;

  // voxel-major slots for median / trimmed mean
  bool slots_built = false;
  muto::DBuf<uint64_t> col_ptr;
  muto::DBuf<uint64_t> slot;
  muto::DBuf<double> S;
  uint64_t max_mj = 0;
  uint32_t n_big = 0;               // voxels with M_j above the LDS sort's capacity
  muto::DBuf<uint32_t> big_vox;     // [n_big] their indices
  muto::DBuf<uint64_t> big_off;     // [n_big] their offsets in sel_scratch
  muto::DBuf<double> sel_scratch;

  // images and accumulators
  muto::DBuf<double> lam64, ml64, tmp64, acc64;
  muto::DBuf<float> lam32;
  muto::DBuf<double> cost_acc;     // [1]
  muto::DBuf<unsigned long long> dcount;  // device counters [8]
  muto::DBuf<unsigned int> work;   // work-queue counters [4]
  int cur_prec = -1;               // which image is current
  int fp64_kernel = 0;             // FP64 iteration: 0 lane per track, 1 wavefront per track,
                                   // 2 the FP64 brick stream
  int use_graph = 1;               // FP32C: replay the captured iteration
  int timing = 0;                  // per-kernel events (no graph)

  muto::Fp32cData f32;

  hipEvent_t ev[8] = {};
  double t_ms[5] = {0, 0, 0, 0, 0};
};

namespace muto {
// counts entries per voxel into plan->mj unless already done
int ensure_mj(muto_em_plan* pl);
}  // namespace muto
