/*
 * muto_gen.h -- seeded synthetic cosmic-muon generator (host code).
 *
 * No counterpart in the reference: the reference reads recorded tracker
 * hits (MutoFile::getMuonData, src/MutoFile.cpp:19-82) and fits the in/out
 * lines (src/MutoFile.cpp:248-304).  This generator produces the same
 * RayData records (muto_ray) from a seed, so every workload is
 * reproducible without data files:
 *   - zenith theta with intensity per solid angle ~ cos^2(theta), truncated
 *     at theta_max; azimuth uniform;
 *   - momentum log-normal (median p_median GeV/c, sigma p_sigma_ln),
 *     clipped to [p_min, p_max];
 *   - crossing point uniform over the volume's mid-plane square, i.e. entry
 *     point uniform on the top tracker plane restricted to lines that reach
 *     the volume;
 *   - multiple scattering with the Highland theta0 per material run along
 *     the track (consecutive voxel layers of one material), with correlated
 *     (angle, offset) kicks in the x-z and y-z planes;
 *   - three tracker layers above (z_top + {1000, 500, 0}) and below
 *     (z_bot - {0, 500, 1000}) with z_top = grid top + tracker_gap and
 *     z_bot = grid bottom - tracker_gap; hits smeared by smear_sigma;
 *   - least-squares line fits; pin on z_top, pout on z_bot as the reference
 *     places them (src/MutoFile.cpp:275-276,297-298); unit directions with
 *     d.z > 0.
 * Deterministic for a seed, independent of the thread count.
 */
#ifndef MUTO_GEN_H
#define MUTO_GEN_H

#include "muto.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Voxel-aligned box [x0,x1) x [y0,y1) x [z0,z1) of radiation length x0_mm. */
typedef struct muto_gen_block {
  uint32_t x0, y0, z0, x1, y1, z1;
  double x0_mm;
} muto_gen_block;

typedef struct muto_gen_cfg {
  muto_grid grid;
  uint64_t seed;
  double tracker_gap;   /* mm between grid face and nearest tracker plane */
  double layer_spacing; /* mm between tracker planes                      */
  double smear_sigma;   /* mm, Gaussian hit smear                         */
  double p_median, p_sigma_ln, p_min, p_max; /* GeV/c                     */
  double theta_max_deg;
  double air_x0;        /* mm                                             */
  int32_t n_blocks;
  int32_t min_voxels;   /* accept only straight chords crossing at least  */
  int32_t max_voxels;   /* / at most this many voxels (0 = no bound)      */
  int32_t _pad;
  const muto_gen_block* blocks;
} muto_gen_cfg;

/* Defaults: gap 500, spacing 500, smear 0.5, p 3 / 0.8 / [0.5, 100],
 * 60 degrees, air 303900 mm, no blocks, no voxel-count bounds. */
void muto_gen_default(muto_gen_cfg* cfg, const muto_grid* grid,
                      uint64_t seed);

/* Seeded block layout: n cubes with sides uniform in [min_mm, max_mm]
 * (rounded to whole voxels, at least one), placed inside the grid; block k
 * takes radiation length x0_choices[k % n_choices]. */
int muto_gen_layout(const muto_grid* grid, uint64_t seed, int n,
                    double min_mm, double max_mm, const double* x0_choices,
                    int n_choices, muto_gen_block* out);

/* Produce n accepted muons into out[n].  *n_tried (may be NULL) receives the
 * number of candidates drawn (acceptance = n / n_tried). */
int muto_generate(const muto_gen_cfg* cfg, int64_t n, muto_ray* out,
                  int64_t* n_tried);

/* Reference-format radiation length lambda_true = (13.6/3000)^2 / X0
 * (rad^2/mm at p0 = 3 GeV/c, src/MutoMLSD.cpp:275) per voxel for the
 * layout: img[nx*ny*nz], x fastest. */
int muto_gen_truth(const muto_gen_cfg* cfg, double* img);

#ifdef __cplusplus
}
#endif
#endif
