/*
 * muto_oracle.h -- CPU FP64 restatement of the MuonTomography reconstruction
 * hot path.  TEST INFRASTRUCTURE ONLY: this is the checker the HIP kernels
 * are compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never calls it.
 *
 * It follows the reference's operation order so that results can be
 * compared bit for bit where the reference is deterministic:
 *   src/MutoSiddon.cpp:9-141, src/MutoPoCA.cpp:51-174,
 *   src/MutoMLSD.cpp:91-517, src/MutoMAP.cpp:39-407.
 * Build flags: -O3, no -march, -ffp-contract=off (the reference is built as
 * CMake Release with no -march, so x86-64 baseline: no FMA).
 */
#ifndef MUTO_ORACLE_H
#define MUTO_ORACLE_H

#include "../include/muto.h"

#ifdef __cplusplus
extern "C" {
#endif

/* static_cast<unsigned>(double) as GCC emits it on x86-64. */
uint32_t orc_x86_u32(double v);

/* One Siddon segment; returns the entry count and writes at most `cap`. */
size_t orc_siddon_path(const muto_grid* g, const double p1[3],
                       const double p2[3], uint32_t* vx, uint32_t* vy,
                       uint32_t* vz, double* vl, size_t cap);
/* Per-segment count and digest (FNV-1a over the 64-bit words x | y << 32,
 * z, length bits of each entry, in order) of the Siddon lists. */
int orc_siddon_digest(const muto_grid* g, size_t n, const double* p1,
                      const double* p2, uint32_t* count, uint64_t* digest);
/* Batched, CSR output (two-call protocol like muto_siddon_csr). */
int orc_siddon_csr(const muto_grid* g, size_t n, const double* p1,
                   const double* p2, uint64_t* row_ptr, uint32_t* vx,
                   uint32_t* vy, uint32_t* vz, double* len);

/* PoCA point; returns 1 if the lines were not parallel. */
int orc_poca_point(const muto_ray* r, double eps, double out[3]);
double orc_scat_angle_sq(const double d1[3], const double d2[3],
                         double angmag);
int orc_poca_points(const muto_grid* g, size_t n, const muto_ray* rays,
                    double eps, int mode, double* p_scat, uint32_t* vox,
                    uint8_t* accept, muto_counters* c);
int orc_poca_reconstruct(const muto_poca_cfg* cfg, size_t n,
                         const muto_ray* rays, double* img, muto_counters* c);
/* MutoPoCut::reconstruct (src/MutoPoCut.cpp:19-121): two-leg weights as
 * PoCA, per voxel the lowest (u32)(M * cut_percentile) scattering values
 * (sorted, summed in order) over the weight. */
int orc_pocut_reconstruct(const muto_poca_cfg* cfg, double cut_percentile,
                          size_t n, const muto_ray* rays, double* img,
                          muto_counters* c);
/* Records of n_layers (x, y) hits between e_in and e_out, precision 4
 * (float) or 8 (double), to rays: NaN records dropped, each side fitted by
 * its principal axis (src/MutoFile.cpp:19-82, 248-304). */
void orc_fit_line(const double* x, const double* y, const double* z, int n,
                  double zref, double point[3], double dir[3]);
int orc_rays_from_records(size_t n_rec, int n_layers, int n_middle,
                          const double* z, const void* records, int precision,
                          muto_ray* rays, size_t* n_rays);
/* MutoSLR::reconstruct (src/MutoSLR.cpp:47-103). */
int orc_slr_reconstruct(const muto_poca_cfg* cfg, size_t n,
                        const muto_ray* rays, double* img, muto_counters* c);

/* MLSD preparation (gatherInformation + getMuonPathAndTail).  Outputs are
 * malloc'ed; free with orc_em_free. */
typedef struct orc_em {
  muto_grid grid;
  size_t n_sel;        /* selected tracks                          */
  uint32_t* sel;       /* index of each selected ray               */
  double* trk4;        /* Sx, Sy, Dx, Dy per selected track        */
  double* p;           /* p_r per selected track                   */
  double* pscat;       /* PoCA point per selected track (3 each)   */
  uint64_t* row_ptr;   /* n_sel+1                                  */
  uint32_t* vox;       /* linear voxel index per entry             */
  double* len;         /* path length per entry                    */
  double* tail;        /* tail length per entry                    */
  /* voxel-major S slots (allocateSValueMatrix, MutoMLSD.cpp:432-436) */
  uint64_t* col_ptr;   /* n_vox+1                                  */
  uint64_t* slot;      /* slot of each entry in voxel-major order  */
  double* S;           /* S values, voxel-major                    */
  int64_t n_empty;
  int64_t n_out_of_grid;
} orc_em;

int orc_em_prepare(const muto_mlsd_cfg* cfg, size_t n, const muto_ray* rays,
                   orc_em* em, muto_counters* c);
/* Build from externally prepared data (copies). */
int orc_em_from_prepared(const muto_grid* g, size_t n_trk, const double* trk4,
                         const double* p, const uint64_t* row_ptr,
                         const uint32_t* vox, const double* len, orc_em* em);
void orc_em_free(orc_em* em);
/* One ML/EM iteration (MutoMLSD.cpp:179-213, 442-517): img_in -> img_out.
 * Returns the mean cost.  n_singular counts warnings. */
double orc_em_iterate(const muto_mlsd_cfg* cfg, orc_em* em,
                      const double* img_in, double* img_out,
                      int64_t* n_singular);
/* c_ij = p^2 <W_ij, G_i> per entry (compact form); for invariant tests. */
void orc_em_cij(const muto_mlsd_cfg* cfg, const orc_em* em,
                const double* img, double* cij, double* cost_i);
/* Per-voxel entry counts. */
void orc_em_mj(const orc_em* em, uint32_t* mj);

int orc_mlsd_reconstruct(const muto_mlsd_cfg* cfg, size_t n,
                         const muto_ray* rays, const double* init_img,
                         double* img_out, double* cost_per_iter,
                         muto_counters* c);

/* MAP (MutoMAP.cpp) */
uint32_t orc_cubic_solve(double* roots, double p3, double p2, double p1,
                         double p0, double tol, int* warn);
/* One MAP voxel step on a whole image (MutoMAP.cpp:162-184):
 * img_last, img_ml -> img_out, with M_j. */
void orc_map_voxel_step(const muto_mlsd_cfg* cfg, const uint32_t* mj,
                        const double* img_last, const double* img_ml,
                        double* img_out, int64_t* n_cubic_warn);
double orc_map_surrogate(const muto_mlsd_cfg* cfg, uint32_t mj, double ljn,
                         double ljml, double lj, const double* w,
                         const double* lm, int nn);
/* Neighbour values (scaled by 1e7) and weights of voxel j. Returns count. */
int orc_map_neighbours(const muto_mlsd_cfg* cfg, const double* img,
                       uint32_t j, double* w, double* lm);
int orc_map_reconstruct(const muto_mlsd_cfg* cfg, size_t n,
                        const muto_ray* rays, const double* init_img,
                        double* img_out, double* cost_per_iter,
                        muto_counters* c);

/* MAP objective Phi(lambda) (penalised negative log-likelihood) and its
 * gradient with respect to lambda (voxel scattering density). */
double orc_map_phi(const muto_mlsd_cfg* cfg, const orc_em* em,
                   const double* lambda);
void orc_map_grad(const muto_mlsd_cfg* cfg, const orc_em* em,
                  const double* lambda, double* grad);
/* The reference's own MAP gradient: the derivative of its surrogate
 * (MutoMAP.cpp:261-281) at l = l_n, by central differences of the
 * surrogate, mean mode; returned per voxel in lambda units. */
void orc_map_surrogate_grad(const muto_mlsd_cfg* cfg, orc_em* em,
                            const double* lambda, double* grad);

#ifdef __cplusplus
}
#endif
#endif
