/*
 * muto.h -- C ABI of the MI355X (gfx950) reconstruction hot path for muon
 * scattering tomography: Siddon voxel tracing, PoCA, MLSD/EM and MAP.
 *
 * Every entry point is a plain C function: plain pointers and sizes, no C++
 * or torch types.  Every function returns MUTO_OK (0) or a negative status;
 * muto_last_error() returns the message of the last failure on a context.
 * No exception crosses this boundary.
 *
 * Buffers are caller owned.  Functions taking "host" pointers copy in/out;
 * the plan API (muto_em_plan_*) keeps its inputs resident in device memory
 * so iterations run without host traffic.
 *
 * A context is not thread-safe: use one per host thread, the same rule as
 * the reference's per-instance reconstruction objects.
 *
 * Reference interfaces replaced (paths relative to LasDes/MuonTomography):
 *   VoxelGrid, RayData, VoxelData ........ include/MutoTypes.h:25-55
 *   MutoSiddon::getVoxelPath ............. src/MutoSiddon.cpp:9-141
 *   MutoPoCA::reconstruct ................ src/MutoPoCA.cpp:51-133
 *   MutoPoCA::getPointOfClosestApproach .. src/MutoPoCA.cpp:142-163
 *   MutoMLSD::reconstruct ................ src/MutoMLSD.cpp:91-220
 *   MutoMAP::reconstruct ................. src/MutoMAP.cpp:39-205
 *   MutoVReconstruction::reconstruct ..... include/MutoVReconstruction.h:11-16
 */
#ifndef MUTO_H
#define MUTO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MUTO_ABI_VERSION 1

/* status codes */
#define MUTO_OK 0
#define MUTO_ERR_ARG (-1)     /* invalid argument                      */
#define MUTO_ERR_HIP (-2)     /* HIP runtime failure                   */
#define MUTO_ERR_NOMEM (-3)   /* device or host allocation failed      */
#define MUTO_ERR_NODEV (-4)   /* no usable gfx950 device               */
#define MUTO_ERR_STATE (-5)   /* call out of order / plan not ready    */
#define MUTO_ERR_LIMIT (-6)   /* a size exceeds what the kernels take  */

/* Where a caller's buffer lives, for the entry points that take a `mem`
 * argument or the MUTO_MAP_DEVICE_BUFFERS flag.  Device buffers must be on
 * the context's GPU and fully written before the call (for example, the
 * caller's stream synchronised); each call has finished reading and
 * writing them when it returns. */
#define MUTO_MEM_HOST 0    /* host memory, copied in or out              */
#define MUTO_MEM_DEVICE 1  /* device memory, read or written in place    */

/* arithmetic of the EM / MAP iterations */
#define MUTO_FP64 0   /* reference order, FP64 throughout                    */
#define MUTO_FP32C 1  /* FP64 Siddon, 2x2 algebra and sums; FP32 entry terms;
                         the mean's sums are FP64 registers and int64
                         fixed-point LDS atomics by default, FP32 with
                         two-sum compensation under MUTO_OPT_ACC_MODE 0;
                         the median and trimmed mean select among the
                         FP32 c_ij of each voxel                        */
#define MUTO_FP64_STREAM 2  /* FP64 on the brick stream (MUTO_OPT_FP64_KERNEL 2):
                               the reference's W_ij bit for bit, the sums in
                               another association; every update rule     */

/* Voxel grid: same field order, types and 64-byte layout as VoxelGrid
 * (include/MutoTypes.h:25-37).  Linear voxel index j = x + y*nx + z*nx*ny
 * (include/MutoMLSD.h:76), x fastest. */
typedef struct muto_grid {
  uint32_t nx, ny, nz;
  uint32_t _pad;
  double x_min, y_min, z_min;
  double dx, dy, dz;
} muto_grid;

/* One muon: same fields and 112-byte layout as RayData (MutoTypes.h:48-55):
 * energies, in/out reference points and in/out directions. */
typedef struct muto_ray {
  double ein, eout;
  double pin[3], pout[3], din[3], dout[3];
} muto_ray;

/* Keys of the "poca" JSON object (src/MutoPoCA.cpp:24-47). */
typedef struct muto_poca_cfg {
  muto_grid grid;
  int32_t length_weighted, straight_line, has_energy, log_results;
  int64_t total_num_rays; /* 0 = all */
  double angle_unit_mag, eps;
} muto_poca_cfg;

/* Keys of the "mlsd" / "map" JSON objects (src/MutoMLSD.cpp:47-84,
 * src/MutoMAP.cpp:12-19).  MAP-only fields are ignored by MLSD. */
typedef struct muto_mlsd_cfg {
  muto_grid grid;
  int32_t update_with_mean, straight_line, has_energy, early_stop;
  int32_t log_results, mean_percentile, num_iteration, initialize_method;
  int64_t total_num_rays; /* 0 = all */
  double initialize_value, error_xy, error_dz_outer, error_dz_inner;
  double angle_unit_mag, eps;
  double beta, p, delta;               /* MAP prior */
  int32_t add_object_prior, all_neighbours;
  int32_t precision;                   /* MUTO_FP64 or MUTO_FP32C */
  int32_t _pad;
} muto_mlsd_cfg;

/* Counters and warning tallies.  The reference prints these
 * (src/MutoPoCA.cpp:176-183, src/MutoMLSD.cpp:535-543) or warns inline
 * (src/MutoMLSD.cpp:456-463, src/MutoMAP.cpp:356-358); here they are
 * returned. */
typedef struct muto_counters {
  int64_t n_total;          /* rays processed                                 */
  int64_t n_not_parallel;   /* PoCA lines that were not parallel              */
  int64_t n_scat_in_grid;   /* PoCA accepted / MLSD selected tracks           */
  int64_t n_empty_path;     /* selected tracks whose merged path is empty     */
  int64_t n_out_of_grid;    /* path entries with a voxel index outside grid   */
  int64_t n_singular;       /* covariances not invertible or det <= 0 (last it)*/
  int64_t n_cubic_warn;     /* MAP: |p3| < eps in the cubic solver (last it)  */
  int64_t n_newton_cap;     /* MAP: Newton iteration cap reached (last it)    */
  int64_t nnz;              /* CSR entries (track-voxel pairs)                */
  int64_t n_segments;       /* FP32C: brick segments                          */
  int64_t iterations_done;
  int64_t n_median_miss;    /* FP32C median by brackets: voxels whose median
                               left their bracket (last it)                   */
  int64_t n_map_tile_records; /* MAP operator with the paths walked again: runs
                               of entries recorded per 16^3 tile (0: none kept,
                               a lane per track, one atomic per entry)        */
  int64_t _reserved[3];
} muto_counters;

/* ------------------------------------------------------------------ */
/* context                                                             */
/* ------------------------------------------------------------------ */
typedef struct muto_ctx muto_ctx;

/* Open device `device` (must be gfx950).  Replaces nothing in the
 * reference (it is single-process CPU code); holds the HIP stream. */
int muto_ctx_create(int device, muto_ctx** out);
int muto_ctx_destroy(muto_ctx* ctx);
const char* muto_last_error(const muto_ctx* ctx);
int muto_abi_version(void);
/* GPU time (ms, HIP events on the context's stream) of the kernels of the
 * last muto_siddon_csr (fill pass), muto_siddon_digest, muto_poca_points,
 * muto_poca_reconstruct, muto_pocut_reconstruct or muto_slr_reconstruct call
 * on this context, without the host copies (the images: with the calls'
 * few waits for sizes, and on a context's first such call the allocation
 * of its workspaces).  No
 * counterpart in the reference, which times whole runs with a stopwatch
 * (MuToTest.cpp:25,80-81). */
int muto_ctx_kernel_ms(muto_ctx* ctx, double* ms);

/* ------------------------------------------------------------------ */
/* Siddon voxel tracing                                                */
/* ------------------------------------------------------------------ */
/* Batched MutoSiddon::getVoxelPath (src/MutoSiddon.cpp:9-141) for n
 * segments p1[i]->p2[i] (host arrays, 3 doubles per point).
 * Two-call protocol: with vx == NULL only row_ptr (n+1 entries, exclusive
 * prefix of per-segment counts) is written; then call again with arrays of
 * row_ptr[n] entries.  vx/vy/vz are the voxel (x,y,z) of each entry
 * (VoxelData, MutoTypes.h:40-45), len its path length; vy, vz, len may be
 * NULL to skip them.  Bit-exact with the reference's FP64 arithmetic. */
int muto_siddon_csr(muto_ctx* ctx, const muto_grid* grid, size_t n,
                    const double* p1, const double* p2, uint64_t* row_ptr,
                    uint32_t* vx, uint32_t* vy, uint32_t* vz, double* len);

/* Per-segment entry count and digest of the lists muto_siddon_csr would
 * return, without storing them: count[n], digest[n] (host).  The digest is
 * FNV-1a over 64-bit words (offset 0xcbf29ce484222325, prime 0x100000001b3),
 * three words per entry in list order: x | y << 32, z, and the bit pattern of
 * the FP64 length.  No counterpart in the reference: it verifies
 * MutoSiddon::getVoxelPath (src/MutoSiddon.cpp:9-141) on segment sets whose
 * lists do not fit host memory. */
int muto_siddon_digest(muto_ctx* ctx, const muto_grid* grid, size_t n,
                       const double* p1, const double* p2, uint32_t* count,
                       uint64_t* digest);

/* ------------------------------------------------------------------ */
/* PoCA                                                                */
/* ------------------------------------------------------------------ */
/* Per-ray PoCA point (MutoPoCA.cpp:142-163) and voxel assignment.
 * mode 0: the PoCA accept test (MutoPoCA.cpp:71-82, x86 double->u32 cast);
 * mode 1: the MLSD selection test (MutoMLSD.cpp:298-301).
 * zDetMin/zDetMax are taken from rays[0] as the reference does.
 * Outputs (host, may be NULL): p_scat[3n], vox[3n] (x,y,z casts),
 * accept[n] (1 = accepted). */
int muto_poca_points(muto_ctx* ctx, const muto_grid* grid, size_t n,
                     const muto_ray* rays, double eps, int mode,
                     double* p_scat, uint32_t* vox, uint8_t* accept,
                     muto_counters* counters);

/* MutoPoCA::reconstruct (MutoPoCA.cpp:51-133).  img: host, nx*ny*nz FP64.
 * The weight sums are made per 16^3-voxel tile in LDS from records of
 * where each leg's walk enters a tile; the context keeps those device
 * workspaces (about 40 B per tile visit and 120 B per ray: some 8 GB at
 * 1e7 muons on 200^3 voxels) for its later PoCA, PoCut and SLR calls and
 * frees them in muto_ctx_destroy. */
int muto_poca_reconstruct(muto_ctx* ctx, const muto_poca_cfg* cfg, size_t n,
                          const muto_ray* rays, double* img,
                          muto_counters* counters);

/* MutoPoCut::reconstruct (src/MutoPoCut.cpp:19-121): the PoCA weights of
 * the two-leg path (cfg->straight_line is ignored: PoCut forces it off,
 * :5,14), and per voxel the sum of the lowest (u32)(M * cut_percentile) of
 * its M scattering values (sorted ascending, summed in order) over the
 * weight.  A threshold above M is clamped to M (the reference reads past
 * the list). */
int muto_pocut_reconstruct(muto_ctx* ctx, const muto_poca_cfg* cfg,
                           double cut_percentile, size_t n,
                           const muto_ray* rays, double* img,
                           muto_counters* counters);

/* MutoSLR::reconstruct (src/MutoSLR.cpp:47-103): each track's theta^2 / 2
 * spread along the straight chord between the grid's top and bottom planes
 * (cfg->straight_line is not used).  Tiled as muto_poca_reconstruct; an
 * entry's value is computed as f L^2 with the track's f = theta^2 / 2 /
 * ltot^2 (the reference: theta^2 / 2 (L / ltot)^2) and the weight as
 * (sum L) / dz per voxel (the reference: sum of L / dz), a few ulps per
 * term apart. */
int muto_slr_reconstruct(muto_ctx* ctx, const muto_poca_cfg* cfg, size_t n,
                         const muto_ray* rays, double* img,
                         muto_counters* counters);

/* ------------------------------------------------------------------ */
/* Muon records                                                        */
/* ------------------------------------------------------------------ */
/* MutoFile::getMuonData's per-record work (src/MutoFile.cpp:19-82,
 * 248-304) for n_rec records already read from files: each record is
 * e_in, (x, y) of the hit on each of n_layers tracker planes at heights
 * z[n_layers], e_out, as float (precision 4) or double (8), host memory.
 * Records holding a NaN are dropped (:72); each side (planes 0..n_middle-1
 * above, the rest below) is fitted by the principal axis of its centred
 * hits, sign d.z > 0, with pin at z[n_middle-1] and pout at z[n_middle]
 * (:270-276, 292-298).  rays (host, room for n_rec) receives the kept
 * records in order, *n_rays their number.  2 <= n_middle,
 * n_layers - n_middle >= 2, n_layers <= MUTO_MAX_LAYERS. */
#define MUTO_MAX_LAYERS 16
int muto_rays_from_records(muto_ctx* ctx, size_t n_rec, int n_layers, int n_middle,
                           const double* z, const void* records, int precision,
                           muto_ray* rays, size_t* n_rays);

/* ------------------------------------------------------------------ */
/* MLSD / MAP one-shot                                                 */
/* ------------------------------------------------------------------ */
/* MutoMLSD::reconstruct (MutoMLSD.cpp:91-220).  init_img: the initial
 * image (initialize_method 0; the caller loads the file), or NULL for
 * methods 1/2.  cost_per_iter (may be NULL) receives num_iteration values.
 * img_out: nx*ny*nz FP64. */
int muto_mlsd_reconstruct(muto_ctx* ctx, const muto_mlsd_cfg* cfg, size_t n,
                          const muto_ray* rays, const double* init_img,
                          double* img_out, double* cost_per_iter,
                          muto_counters* counters);

/* MutoMAP::reconstruct (MutoMAP.cpp:39-205). */
int muto_map_reconstruct(muto_ctx* ctx, const muto_mlsd_cfg* cfg, size_t n,
                         const muto_ray* rays, const double* init_img,
                         double* img_out, double* cost_per_iter,
                         muto_counters* counters);

/* ------------------------------------------------------------------ */
/* EM plan: inputs resident in HBM, iterations on demand               */
/* ------------------------------------------------------------------ */
/* The plan performs the reference's preparation once (MutoMLSD.cpp:99-166:
 * selection, angles, displacements, two-leg Siddon paths, tails) and keeps
 * the track-voxel lists in device memory.  For precision MUTO_FP32C it also
 * builds the brick-major segment stream.  Iterations then run on device. */
typedef struct muto_em_plan muto_em_plan;

int muto_em_plan_create(muto_ctx* ctx, const muto_mlsd_cfg* cfg, size_t n,
                        const muto_ray* rays, muto_em_plan** out);
/* Same, but from already prepared per-track data (the reference's
 * gatherInformation output, MutoMLSD.cpp:268-326) and the merged path CSR
 * (getMuonPathAndTail, MutoMLSD.cpp:382-415): row_ptr[n_trk+1], vox (linear
 * j), len.  trk4[4*n_trk] = (Sx, Sy, Dx, Dy); p_r[n_trk] or NULL (p_r = 1).
 * Used to run the iteration on inputs computed elsewhere. */
int muto_em_plan_create_prepared(muto_ctx* ctx, const muto_mlsd_cfg* cfg,
                                 size_t n_trk, const double* trk4,
                                 const double* p_r, const uint64_t* row_ptr,
                                 const uint32_t* vox, const double* len,
                                 muto_em_plan** out);
int muto_em_plan_destroy(muto_em_plan* plan);
/* Set the image (host FP64, nx*ny*nz) the next iteration starts from. */
int muto_em_plan_set_image(muto_em_plan* plan, const double* img);
int muto_em_plan_get_image(muto_em_plan* plan, double* img);
/* The same, with img in host (MUTO_MEM_HOST) or device (MUTO_MEM_DEVICE)
 * memory: the reference's image hand-over (MutoMLSD.cpp:226-245 reads it,
 * MutoVReconstruction::reconstruct returns it, include/MutoVReconstruction.h:
 * 11-16) without a trip through host memory. */
int muto_em_plan_set_image_ex(muto_em_plan* plan, const double* img, int mem);
int muto_em_plan_get_image_ex(muto_em_plan* plan, double* img, int mem);
/* Run `iters` iterations in `precision` (MUTO_FP64 / MUTO_FP32C) with the
 * plan's update rule (median / mean / trimmed mean, MLSD or MAP).
 * cost (may be NULL) receives one value per iteration.  Stops early as the
 * reference does when cfg.early_stop is set. */
int muto_em_plan_run(muto_em_plan* plan, int iters, int precision,
                     double* cost);
/* MAP iterations on the same plan (MutoMAP.cpp:131-204): each iteration
 * runs the ML step on the current image, then the per-voxel surrogate
 * step with cfg.beta / p / delta / add_object_prior / all_neighbours. */
int muto_map_plan_run(muto_em_plan* plan, int iters, int precision,
                      double* cost);
/* As muto_em_plan_create, with a budget for the track-major path lists
 * (20 bytes per track-voxel entry: u32 voxel, f64 length, f64 tail).  Above
 * it the plan keeps no lists: muto_em_plan_run refuses it, and the MAP
 * objective operator below walks every path again by Siddon (FP64).  Its
 * passes then run per 16^3 tile from tile records (84 bytes per run of a
 * track's entries inside one tile, made at its first call), the backward's
 * sums in LDS, while those fit the same budget; else they walk one track
 * per lane and the backward adds one atomic per entry
 * (muto_counters.n_map_tile_records says which).  No
 * counterpart in the reference, which always stores its paths
 * (MutoMLSD.cpp:382-415). */
int muto_em_plan_create_ex(muto_ctx* ctx, const muto_mlsd_cfg* cfg, size_t n,
                           const muto_ray* rays, uint64_t list_budget_bytes,
                           muto_em_plan** out);

/* ---- MAP objective operator (forward / backward with saved context) ----
 *   Phi(lambda) = 1/2 sum_i cost_i(lambda) + beta R(lambda)
 * cost_i: the EM iteration's per-track cost (MutoMLSD.cpp:456-468; 0 for a
 * singular or non-positive Sigma_i); R: the MAP prior of MutoMAP.cpp:207-237
 * (psi(t) = (t^2 + delta^2)^(p/2) over the 6 or 26 clamp-to-edge
 * neighbours with weights 1, 1/sqrt2, 1/sqrt3, plus psi(l_j) with
 * add_object_prior; l = 1e7 lambda), summed over all voxels.  The reference
 * never evaluates Phi or its gradient; in mean mode its MAP step
 * (MutoMAP.cpp:162-184) minimises a per-voxel surrogate tangent to Phi, so
 * this gradient is the reference's MAP gradient there.
 * muto_map_forward evaluates Phi at lambda (n_vox doubles, host memory or,
 * with MUTO_MAP_DEVICE_BUFFERS, device memory) on the
 * plan's tracks and returns a saved context; muto_map_backward writes
 * grad_lambda = grad_phi dPhi/dlambda (host, n_vox doubles) and may be
 * called several times.  precision: MUTO_FP64 (track-major lists, FP64;
 * with MUTO_OPT_FP64_KERNEL 2 the FP64 brick stream), MUTO_FP64_STREAM (the
 * FP64 brick stream: the lists' exact FP64 lengths and tails, sums per
 * segment and per brick) or MUTO_FP32C (brick stream, FP32 entry terms).
 * flags: MUTO_MAP_RECOMPUTE_SIDDON walks the paths again instead of reading
 * the stored lists or a brick stream (implied for plans without lists;
 * refused with MUTO_FP64_STREAM): the walk and the forward sums are FP64;
 * with MUTO_FP32C the backward's entry terms are FP32.  A saved context
 * must be destroyed before its plan. */
typedef struct muto_map_saved muto_map_saved;
#define MUTO_MAP_RECOMPUTE_SIDDON 1
/* flag of muto_map_forward: lambda, and grad_lambda of every
 * muto_map_backward on the saved context, are device pointers (see
 * MUTO_MEM_DEVICE); the gradient is accumulated in the caller's array */
#define MUTO_MAP_DEVICE_BUFFERS 2
int muto_map_forward(muto_em_plan* plan, const double* lambda, int precision, int flags,
                     double* phi, muto_map_saved** saved);
int muto_map_backward(muto_map_saved* saved, double grad_phi, double* grad_lambda);
int muto_map_saved_destroy(muto_map_saved* saved);

/* Phi and grad_lambda = dPhi/dlambda in one call that allocates nothing:
 * no counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  FP64, from the plan's stored lists or, with
 * MUTO_MAP_RECOMPUTE_SIDDON (implied for plans without lists), with the
 * paths walked again; the same Phi (bit for bit) and gradient (up to the
 * order of FP64 atomic additions) as muto_map_forward + muto_map_backward
 * with grad_phi = 1.  lambda[n_vox], phi[1] and grad_lambda[n_vox] are
 * device pointers on the context's GPU (MUTO_MEM_DEVICE rules).  Its only
 * memory beyond those and the plan is a static per-GPU scratch of 1 MiB of
 * partial sums, so plans of more than 2^24 tracks or 2^24 voxels are
 * refused (MUTO_ERR_LIMIT).  Calls on one GPU share that scratch: calls
 * that overlap on one GPU (several host threads) race. */
int muto_map_value_grad_noalloc(muto_em_plan* plan, const double* lambda, int flags,
                                double* phi, double* grad_lambda);

/* ---- many independent scans (one launch per pass for all of them) ----
 * A batch combines plans (each with its own grid, voxel size and tracks;
 * sharing the EM configuration, mean update) into one problem: their grids
 * side by side in one brick-aligned grid, their track lists concatenated,
 * one size-sorted task queue for the persistent FP32C kernels and one
 * captured graph per iteration.  muto_batch_run runs `iters` mean-update
 * iterations of every scan in lock step, from and back to the scans' plan
 * images; costs (may be NULL): iters x n values, scan fastest.  The batch
 * takes the brick shape and block size of plans[0] (MUTO_OPT_BRICK).  The
 * plans must outlive the batch.  No counterpart in the reference (one scan
 * per process). */
typedef struct muto_batch muto_batch;
int muto_batch_create(muto_ctx* ctx, size_t n, muto_em_plan* const* plans, muto_batch** out);
int muto_batch_run(muto_batch* batch, int iters, double* costs);
int muto_batch_stats(const muto_batch* batch, double* wall_ms, muto_counters* counters);
int muto_batch_destroy(muto_batch* batch);

/* Plan options (value 0/1 unless stated).
 *   MUTO_OPT_GRAPH    FP32C iterations replay one captured HIP graph (default 1)
 *   MUTO_OPT_TIMING   record per-kernel times for muto_em_plan_timing; runs
 *                     without the graph (default 0)
 *   MUTO_OPT_ACC_MODE FP32C accumulation: 1 FP64 registers + int64 fixed-point
 *                     LDS sums (default), 0 FP32 with two-sum compensation
 *   MUTO_OPT_FP64_KERNEL FP64 iteration kernel: 0 one lane per track, the
 *                     reference's operation order (default; per-entry values
 *                     bit-identical to the CPU path); 1 one wavefront per
 *                     track, coalesced row loads and a butterfly sum of the
 *                     covariance (per-track values within FP64 rounding);
 *                     2 the FP64 brick stream: the FP32C stream's layout with
 *                     the exact FP64 lengths and tails, per-segment FP64
 *                     partial sums, LDS-staged backprojection (mean) or
 *                     per-entry c_ij in voxel-major slots and a per-voxel
 *                     sort (median, trimmed mean); every W_ij is the
 *                     reference's, only the association of the sums differs
 *   MUTO_OPT_MEDIAN_BRACKETS FP32C median: 1 (default) each voxel keeps a
 *                     bracket of c values around its last median; one pass
 *                     counts the values below it and keeps those inside,
 *                     the median is picked from those, and only the voxels
 *                     whose median left the bracket collect their whole
 *                     lists (the same values and the same pick as 0, so
 *                     the images are identical); 0 every c_ij is written
 *                     to its voxel's slot and each voxel's list is searched
 * No counterpart in the reference (which has no device state). */
#define MUTO_OPT_GRAPH 1
#define MUTO_OPT_TIMING 2
#define MUTO_OPT_ACC_MODE 3
#define MUTO_OPT_FP64_KERNEL 5
#define MUTO_OPT_MEDIAN_BRACKETS 6
#define MUTO_OPT_BRICK 4  /* FP32C brick (LDS tile) shape and block size:
                             bx | by << 8 | bz << 16 | log2(block / 256) << 24,
                             at most 16384 voxels, block 256, 512 or 1024
                             threads; default 16 x 16 x 16, 256 */
int muto_em_plan_set_option(muto_em_plan* plan, int option, int value);

/* FP32C autotuner: builds the brick stream for each candidate tile shape
 * (shapes[k] coded as for MUTO_OPT_BRICK; NULL/0: a built-in list of 4096-
 * to 16384-voxel bricks), times `iters` iterations of each from the plan's image
 * under the plan's update rule (after one untimed iteration, which for the
 * median also sets its brackets; the image is restored after every trial)
 * and keeps the fastest.  ms[k] (may be NULL)
 * receives each shape's ms per iteration, *best (may be NULL) the chosen
 * shape.  No counterpart in the reference. */
int muto_em_plan_autotune(muto_em_plan* plan, int iters, int n_shapes, const uint32_t* shapes,
                          double* ms, uint32_t* best);

int muto_em_plan_counters(const muto_em_plan* plan, muto_counters* out);
/* Per-voxel entry counts M_j (host, nx*ny*nz uint32). */
int muto_em_plan_voxel_counts(muto_em_plan* plan, uint32_t* mj);
/* Copy the plan's prepared data to host arrays (any may be NULL):
 * row_ptr[n_trk+1], vox/len/tail[nnz] (merged two-leg path, linear voxel
 * index, length, tail; MutoMLSD.cpp:382-415), trk4[4*n_trk] (Sx, Sy, Dx,
 * Dy; MutoMLSD.cpp:350-376), p_r[n_trk].  Sizes from muto_em_plan_counters
 * (n_scat_in_grid = n_trk, nnz). */
int muto_em_plan_export(muto_em_plan* plan, uint64_t* row_ptr, uint32_t* vox,
                        double* len, double* tail, double* trk4, double* p_r);
/* Average kernel time (ms) of the last run, per phase, measured with HIP
 * events on the plan's stream: t[0] forward, t[1] track, t[2] backward /
 * S-values, t[3] voxel update, t[4] whole iteration. */
int muto_em_plan_timing(const muto_em_plan* plan, double* t_ms);

#ifdef __cplusplus
}
#endif
#endif /* MUTO_H */
