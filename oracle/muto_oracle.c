/*
 * muto_oracle.c -- CPU FP64 restatement of the MuonTomography hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Each function cites the reference lines it restates (paths relative to
 * LasDes/MuonTomography).  Operation order and association follow the
 * reference so that deterministic results match bit for bit; Eigen 3-vector
 * dot products are (x0*y0 + x1*y1) + x2*y2 and 2x2 products are two-term
 * sums.  The reference's per-voxel std::vector storage is replaced by a
 * voxel-major slot array with the same fill order.
 *
 * Pinning (what checks this restatement, since the reference cannot be
 * compiled here -- Eigen is absent): tests/test_oracle.py checks Siddon
 * against exact rational brute force and closed forms, PoCA against
 * constructed skew/intersecting/parallel lines, the MLSD selection window,
 * angles, displacements and two-leg merge against kinked tracks in closed
 * form (tests/mlsd_cases.py), the W/Sigma/S algebra
 * against the compact gradient identity and finite differences, the median
 * / mean / trimmed updates against numpy on small sets, the cubic solver
 * against numpy.roots, the MAP surrogate against its tangency to Phi, and
 * Phi's gradient against central differences.  The exact bits of Eigen's
 * 2x2 determinant/inverse evaluation are restated from the Eigen source
 * (determinant m00*m11 - m10*m01, inverse adj*(1/det), threshold 1e-12):
 * parity with real Eigen at that level is unpinned.
 */
#include "muto_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* std::min / std::max / std::minmax semantics (ties and signed zeros) */
static inline double smax(double a, double b) { return (a < b) ? b : a; }
static inline double smin(double a, double b) { return (b < a) ? b : a; }

uint32_t orc_x86_u32(double v) {
  /* GCC x86-64: cvttsd2si (64-bit signed, truncating), keep the low 32
   * bits.  NaN and |v| >= 2^63 give the "integer indefinite"
   * 0x8000000000000000, whose low 32 bits are 0. */
  if (!(v >= -9223372036854775808.0 && v < 9223372036854775808.0)) return 0u;
  return (uint32_t)(uint64_t)(int64_t)v;
}

/* ------------------------------------------------------------------ */
/* Siddon: src/MutoSiddon.cpp:9-141                                    */
/* ------------------------------------------------------------------ */
/* Digest of a list: FNV-1a over 64-bit words, three words per entry
 * (x | y << 32, z, the bit pattern of the length), in list order. */
#define ORC_DIGEST_BASIS 0xcbf29ce484222325ULL
#define ORC_DIGEST_PRIME 0x100000001b3ULL
static inline uint64_t digest_word(uint64_t h, uint64_t w) {
  return (h ^ w) * ORC_DIGEST_PRIME;
}

static size_t siddon_walk(const muto_grid* g, const double p1[3],
                          const double p2[3], uint32_t* vx, uint32_t* vy,
                          uint32_t* vz, double* vl, size_t cap,
                          uint64_t* digest);

size_t orc_siddon_path(const muto_grid* g, const double p1[3],
                       const double p2[3], uint32_t* vx, uint32_t* vy,
                       uint32_t* vz, double* vl, size_t cap) {
  return siddon_walk(g, p1, p2, vx, vy, vz, vl, cap, NULL);
}

int orc_siddon_digest(const muto_grid* g, size_t n, const double* p1,
                      const double* p2, uint32_t* count, uint64_t* digest) {
  if (!g || !count || !digest || (n && (!p1 || !p2))) return MUTO_ERR_ARG;
#pragma omp parallel for schedule(dynamic, 1024)
  for (size_t i = 0; i < n; ++i) {
    uint64_t h = ORC_DIGEST_BASIS;
    count[i] = (uint32_t)siddon_walk(g, p1 + 3 * i, p2 + 3 * i, NULL, NULL,
                                     NULL, NULL, 0, &h);
    digest[i] = h;
  }
  return MUTO_OK;
}

static size_t siddon_walk(const muto_grid* g, const double p1[3],
                          const double p2[3], uint32_t* vx, uint32_t* vy,
                          uint32_t* vz, double* vl, size_t cap,
                          uint64_t* digest) {
  const double mn[3] = {g->x_min, g->y_min, g->z_min};
  const double d[3] = {g->dx, g->dy, g->dz};
  const uint32_t nn[3] = {g->nx, g->ny, g->nz};
  double d21[3];
  size_t cnt = 0;
  const double eps = 1.0e-6; /* :19 */
  double a0[3] = {-0.1, -0.1, -0.1}, an[3] = {1.1, 1.1, 1.1}; /* :18 */
  for (int a = 0; a < 3; ++a) d21[a] = p2[a] - p1[a];         /* :20-22 */
  /* :14 length = sqrt((p2-p1).dot(p2-p1)), Eigen dot order */
  const double length =
      sqrt((d21[0] * d21[0] + d21[1] * d21[1]) + d21[2] * d21[2]);
  /* :25-43 first/last plane alphas; parallel rays outside -> empty */
  for (int a = 0; a < 3; ++a) {
    if (eps < fabs(d21[a])) {
      a0[a] = (mn[a] - p1[a]) / d21[a];
      an[a] = (mn[a] + d[a] * (double)nn[a] - p1[a]) / d21[a];
    } else if (p1[a] < mn[a] || mn[a] + d[a] * (double)nn[a] <= p1[a]) {
      return 0;
    }
  }
  /* :46-51 std::minmax then max/min */
  double lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    if (an[a] < a0[a]) { lo[a] = an[a]; hi[a] = a0[a]; }
    else { lo[a] = a0[a]; hi[a] = an[a]; }
  }
  const double amin = smax(smax(smax(lo[0], lo[1]), lo[2]), 0.0);
  const double amax = smin(smin(smin(hi[0], hi[1]), hi[2]), 1.0);
  if (amax <= amin) return 0; /* :54-56 */

  double alpha = amin;
  uint32_t idx[3];
  for (int a = 0; a < 3; ++a) /* :62-64 */
    idx[a] = orc_x86_u32((alpha * d21[a] + p1[a] - mn[a]) / d[a] +
                         (d21[a] <= 0 ? -eps : eps));
  int di[3] = {0, 0, 0};
  double da[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < 3; ++a) { /* :67-81 */
    if (eps < fabs(d21[a])) {
      da[a] = d[a] / (fabs(d21[a]));
      di[a] = d21[a] < 0 ? -1 : 1;
    }
  }
  double ax[3] = {1.1, 1.1, 1.1}; /* :85-103 */
  for (int a = 0; a < 3; ++a) {
    if (fabs(amin - lo[a]) < eps) {
      ax[a] = amin + da[a];
    } else if (di[a] != 0) {
      ax[a] = ((double)idx[a] * d[a] + mn[a] - p1[a]) / d21[a] +
              (di[a] == 1 ? da[a] : 0.0);
    }
  }
  /* :106-138 the walk.  The reference loop has no iteration bound; a ray
   * crosses at most nx+ny+nz+3 planes inside [amin, amax], so the cap
   * below never binds for finite inputs with non-degenerate steps. */
  double alpha_last = alpha;
  const uint64_t cap_iter = (uint64_t)nn[0] + nn[1] + nn[2] + 8;
  uint64_t it = 0;
  while (alpha < amax && it++ < cap_iter) {
    uint32_t cx = idx[0], cy = idx[1], cz = idx[2];
    if (ax[0] < ax[1] && ax[0] < ax[2]) {
      idx[0] += (uint32_t)di[0]; alpha = ax[0]; ax[0] += da[0];
    } else if (ax[1] < ax[2]) {
      idx[1] += (uint32_t)di[1]; alpha = ax[1]; ax[1] += da[1];
    } else {
      idx[2] += (uint32_t)di[2]; alpha = ax[2]; ax[2] += da[2];
    }
    double dalpha = smin(alpha, amax) - alpha_last;
    if (eps < dalpha) {
      const double L = dalpha * length;
      if (cnt < cap) {
        if (vx) vx[cnt] = cx;
        if (vy) vy[cnt] = cy;
        if (vz) vz[cnt] = cz;
        if (vl) vl[cnt] = L;
      }
      if (digest) {
        uint64_t lb;
        memcpy(&lb, &L, sizeof lb);
        uint64_t h = digest_word(*digest, (uint64_t)cx | ((uint64_t)cy << 32));
        h = digest_word(h, cz);
        *digest = digest_word(h, lb);
      }
      ++cnt;
      alpha_last = alpha;
    }
  }
  return cnt;
}

int orc_siddon_csr(const muto_grid* g, size_t n, const double* p1,
                   const double* p2, uint64_t* row_ptr, uint32_t* vx,
                   uint32_t* vy, uint32_t* vz, double* len) {
  if (!g || !row_ptr || (n && (!p1 || !p2))) return MUTO_ERR_ARG;
  if (!vx) {
    row_ptr[0] = 0;
    for (size_t i = 0; i < n; ++i)
      row_ptr[i + 1] = row_ptr[i] + orc_siddon_path(g, p1 + 3 * i, p2 + 3 * i,
                                                    NULL, NULL, NULL, NULL, 0);
    return MUTO_OK;
  }
  for (size_t i = 0; i < n; ++i) {
    uint64_t o = row_ptr[i];
    size_t cap = (size_t)(row_ptr[i + 1] - o);
    orc_siddon_path(g, p1 + 3 * i, p2 + 3 * i, vx + o, vy ? vy + o : NULL,
                    vz ? vz + o : NULL, len ? len + o : NULL, cap);
  }
  return MUTO_OK;
}

/* ------------------------------------------------------------------ */
/* PoCA: src/MutoPoCA.cpp                                              */
/* ------------------------------------------------------------------ */
static inline double dot3(const double* a, const double* b) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

int orc_poca_point(const muto_ray* r, double eps, double out[3]) {
  /* :142-163 (identical: MutoMLSD.cpp:328-348) */
  double w0[3];
  for (int k = 0; k < 3; ++k) w0[k] = r->pin[k] - r->pout[k];
  double a = dot3(r->din, r->din);
  double b = dot3(r->din, r->dout);
  double c = dot3(r->dout, r->dout);
  double d = dot3(r->din, w0);
  double e = dot3(r->dout, w0);
  double den = a * c - b * b;
  if (fabs(den) < eps) {
    for (int k = 0; k < 3; ++k) out[k] = 0.5 * (r->pin[k] + r->pout[k]);
    return 0;
  }
  double cin = (b * e - c * d) / den;
  double cout = (a * e - b * d) / den;
  for (int k = 0; k < 3; ++k)
    out[k] = 0.5 * (((r->pin[k] + cin * r->din[k]) + r->pout[k]) +
                    cout * r->dout[k]);
  return 1;
}

double orc_scat_angle_sq(const double d1[3], const double d2[3],
                         double angmag) {
  /* :165-174 */
  double axz = (atan2(d1[0], d1[2]) - atan2(d2[0], d2[2])) * angmag;
  double ayz = (atan2(d1[1], d1[2]) - atan2(d2[1], d2[2])) * angmag;
  return (axz * axz + ayz * ayz) / 2.0;
}

/* MutoMLSD.cpp:298-301 selection test */
static int mlsd_accept(const muto_grid* g, const double* P, double eps,
                       double zmin_det, double zmax_det) {
  if (P[0] < g->x_min + eps || g->x_min + g->dx * g->nx - eps < P[0] ||
      P[1] < g->y_min + eps || g->y_min + g->dy * g->ny - eps < P[1] ||
      P[2] < g->z_min + eps || g->z_min + g->dz * g->nz - eps < P[2] ||
      P[2] < zmin_det + eps || zmax_det - eps <= P[2])
    return 0;
  return 1;
}

int orc_poca_points(const muto_grid* g, size_t n, const muto_ray* rays,
                    double eps, int mode, double* p_scat, uint32_t* vox,
                    uint8_t* accept, muto_counters* c) {
  if (!g || (n && !rays)) return MUTO_ERR_ARG;
  if (n == 0) return MUTO_OK;
  const double zmin_det = rays[0].pout[2], zmax_det = rays[0].pin[2];
  int64_t npar = 0, nin = 0;
  for (size_t i = 0; i < n; ++i) {
    double P[3];
    npar += orc_poca_point(&rays[i], eps, P);
    uint32_t xs = orc_x86_u32((P[0] - g->x_min) / g->dx);
    uint32_t ys = orc_x86_u32((P[1] - g->y_min) / g->dy);
    uint32_t zs = orc_x86_u32((P[2] - g->z_min) / g->dz);
    int acc;
    if (mode == 0) /* MutoPoCA.cpp:75-79 */
      acc = !(g->nx <= xs || g->ny <= ys || g->nz <= zs || P[2] < zmin_det ||
              zmax_det < P[2]);
    else
      acc = mlsd_accept(g, P, eps, zmin_det, zmax_det);
    nin += acc;
    if (p_scat) memcpy(p_scat + 3 * i, P, sizeof P);
    if (vox) { vox[3 * i] = xs; vox[3 * i + 1] = ys; vox[3 * i + 2] = zs; }
    if (accept) accept[i] = (uint8_t)acc;
  }
  if (c) {
    c->n_total = (int64_t)n;
    c->n_not_parallel = npar;
    c->n_scat_in_grid = nin;
  }
  return MUTO_OK;
}

static size_t n_vox_of(const muto_grid* g) {
  return (size_t)g->nx * g->ny * g->nz;
}

/* grow-on-demand Siddon scratch */
typedef struct { uint32_t *x, *y, *z; double* l; size_t cap; } path_buf;
static size_t trace(const muto_grid* g, const double* a, const double* b,
                    path_buf* pb) {
  size_t k = orc_siddon_path(g, a, b, pb->x, pb->y, pb->z, pb->l, pb->cap);
  if (k > pb->cap) {
    pb->cap = k * 2;
    pb->x = (uint32_t*)realloc(pb->x, pb->cap * sizeof(uint32_t));
    pb->y = (uint32_t*)realloc(pb->y, pb->cap * sizeof(uint32_t));
    pb->z = (uint32_t*)realloc(pb->z, pb->cap * sizeof(uint32_t));
    pb->l = (double*)realloc(pb->l, pb->cap * sizeof(double));
    k = orc_siddon_path(g, a, b, pb->x, pb->y, pb->z, pb->l, pb->cap);
  }
  return k;
}
static void pb_free(path_buf* pb) {
  free(pb->x); free(pb->y); free(pb->z); free(pb->l);
  memset(pb, 0, sizeof *pb);
}
static inline int in_grid(const muto_grid* g, uint32_t x, uint32_t y,
                          uint32_t z) {
  return x < g->nx && y < g->ny && z < g->nz;
}
static inline size_t tidx(const muto_grid* g, uint32_t x, uint32_t y,
                          uint32_t z) {
  /* Eigen Tensor column-major element (x, y, z) */
  return (size_t)x + (size_t)g->nx * ((size_t)y + (size_t)g->ny * z);
}

int orc_poca_reconstruct(const muto_poca_cfg* cfg, size_t n,
                         const muto_ray* rays, double* img,
                         muto_counters* c) {
  /* MutoPoCA.cpp:51-133 */
  const muto_grid* g = &cfg->grid;
  const size_t nv = n_vox_of(g);
  if (!img || (n && !rays)) return MUTO_ERR_ARG;
  memset(img, 0, nv * sizeof(double));
  if (n == 0) return MUTO_OK;
  double* wimg = (double*)calloc(nv, sizeof(double));
  if (!wimg) return MUTO_ERR_NOMEM;
  /* :56 fNTotal; the reference reads rays[i] unchecked past the end */
  size_t ntot = cfg->total_num_rays > 0 ? (size_t)cfg->total_num_rays : n;
  if (ntot > n) ntot = n;
  const double zmin_det = rays[0].pout[2], zmax_det = rays[0].pin[2];
  int64_t npar = 0, nin = 0, nout = 0;
  path_buf pa = {0}, pbb = {0};
  for (size_t i = 0; i < ntot; ++i) {
    const muto_ray* r = &rays[i];
    double P[3];
    npar += orc_poca_point(r, cfg->eps, P);
    double a2 = orc_scat_angle_sq(r->din, r->dout, cfg->angle_unit_mag);
    uint32_t xs = orc_x86_u32((P[0] - g->x_min) / g->dx);
    uint32_t ys = orc_x86_u32((P[1] - g->y_min) / g->dy);
    uint32_t zs = orc_x86_u32((P[2] - g->z_min) / g->dz);
    if (g->nx <= xs || g->ny <= ys || g->nz <= zs || P[2] < zmin_det ||
        zmax_det < P[2])
      continue;
    ++nin;
    if (cfg->has_energy)
      img[tidx(g, xs, ys, zs)] += a2 * r->ein * r->ein;
    else
      img[tidx(g, xs, ys, zs)] += a2;
    if (cfg->straight_line) {
      size_t k = trace(g, r->pin, r->pout, &pa);
      for (size_t m = 0; m < k; ++m) {
        if (!in_grid(g, pa.x[m], pa.y[m], pa.z[m])) { ++nout; continue; }
        wimg[tidx(g, pa.x[m], pa.y[m], pa.z[m])] +=
            cfg->length_weighted ? pa.l[m] : g->dz;
      }
    } else {
      size_t ki = trace(g, r->pin, P, &pa);
      size_t ko = trace(g, P, r->pout, &pbb);
      for (size_t m = 0; m < ki; ++m) {
        if (!in_grid(g, pa.x[m], pa.y[m], pa.z[m])) { ++nout; continue; }
        wimg[tidx(g, pa.x[m], pa.y[m], pa.z[m])] +=
            cfg->length_weighted ? pa.l[m] : g->dz;
      }
      for (size_t m = 0; m < ko; ++m) {
        if (!in_grid(g, pbb.x[m], pbb.y[m], pbb.z[m])) { ++nout; continue; }
        size_t t = tidx(g, pbb.x[m], pbb.y[m], pbb.z[m]);
        if (cfg->length_weighted)
          wimg[t] += pbb.l[m];
        else if (pbb.x[m] != xs || pbb.y[m] != ys || pbb.z[m] != zs)
          wimg[t] += g->dz;
      }
    }
  }
  for (size_t i = 0; i < nv; ++i) /* :130-132 */
    img[i] = wimg[i] < cfg->eps ? 0 : img[i] / wimg[i];
  free(wimg);
  pb_free(&pa);
  pb_free(&pbb);
  if (c) {
    memset(c, 0, sizeof *c);
    c->n_total = (int64_t)ntot;
    c->n_not_parallel = npar;
    c->n_scat_in_grid = nin;
    c->n_out_of_grid = nout;
  }
  return MUTO_OK;
}

/* ------------------------------------------------------------------ */
/* PoCut: src/MutoPoCut.cpp:19-121                                     */
/* ------------------------------------------------------------------ */
static int cmp_double(const void* a, const void* b) {
  const double x = *(const double*)a, y = *(const double*)b;
  return (x < y) ? -1 : (y < x) ? 1 : 0;
}

int orc_pocut_reconstruct(const muto_poca_cfg* cfg, double cut_percentile,
                          size_t n, const muto_ray* rays, double* img,
                          muto_counters* c) {
  const muto_grid* g = &cfg->grid;
  const size_t nv = n_vox_of(g);
  if (!img || (n && !rays)) return MUTO_ERR_ARG;
  memset(img, 0, nv * sizeof(double));
  if (c) memset(c, 0, sizeof *c);
  if (n == 0) return MUTO_OK;
  size_t ntot = cfg->total_num_rays > 0 ? (size_t)cfg->total_num_rays : n;
  if (ntot > n) ntot = n;
  double* wimg = (double*)calloc(nv, sizeof(double));
  /* per-ray (voxel, value) records; the per-voxel lists of :30-31 are
   * these grouped by voxel (order inside a voxel is irrelevant: sorted) */
  uint64_t* rv = (uint64_t*)malloc(ntot * sizeof(uint64_t));
  double* ra = (double*)malloc(ntot * sizeof(double));
  if (!wimg || !rv || !ra) { free(wimg); free(rv); free(ra); return MUTO_ERR_NOMEM; }
  const double zmin_det = rays[0].pout[2], zmax_det = rays[0].pin[2]; /* :38-39 */
  int64_t npar = 0, nin = 0, nout = 0;
  path_buf pa = {0}, pbb = {0};
  for (size_t i = 0; i < ntot; ++i) { /* :44-101 */
    const muto_ray* r = &rays[i];
    double P[3];
    npar += orc_poca_point(r, cfg->eps, P);
    const double a2 = orc_scat_angle_sq(r->din, r->dout, cfg->angle_unit_mag);
    const uint32_t xs = orc_x86_u32((P[0] - g->x_min) / g->dx);
    const uint32_t ys = orc_x86_u32((P[1] - g->y_min) / g->dy);
    const uint32_t zs = orc_x86_u32((P[2] - g->z_min) / g->dz);
    rv[i] = UINT64_MAX;
    if (g->nx <= xs || g->ny <= ys || g->nz <= zs || P[2] < zmin_det ||
        zmax_det < P[2])
      continue;
    ++nin;
    rv[i] = tidx(g, xs, ys, zs); /* :62-66 */
    ra[i] = cfg->has_energy ? a2 * r->ein * r->ein : a2;
    /* :69-98, two legs (fStraightLine is forced false, :5,14) */
    const size_t ki = trace(g, r->pin, P, &pa);
    const size_t ko = trace(g, P, r->pout, &pbb);
    for (size_t m = 0; m < ki; ++m) {
      if (!in_grid(g, pa.x[m], pa.y[m], pa.z[m])) { ++nout; continue; }
      wimg[tidx(g, pa.x[m], pa.y[m], pa.z[m])] += cfg->length_weighted ? pa.l[m] : g->dz;
    }
    for (size_t m = 0; m < ko; ++m) {
      if (!in_grid(g, pbb.x[m], pbb.y[m], pbb.z[m])) { ++nout; continue; }
      const size_t t = tidx(g, pbb.x[m], pbb.y[m], pbb.z[m]);
      if (cfg->length_weighted)
        wimg[t] += pbb.l[m];
      else if (pbb.x[m] != xs || pbb.y[m] != ys || pbb.z[m] != zs)
        wimg[t] += g->dz;
    }
  }
  /* group by voxel: counting sort */
  size_t* off = (size_t*)calloc(nv + 1, sizeof(size_t));
  double* vals = (double*)malloc((nin ? (size_t)nin : 1) * sizeof(double));
  if (!off || !vals) { free(off); free(vals); free(wimg); free(rv); free(ra); return MUTO_ERR_NOMEM; }
  for (size_t i = 0; i < ntot; ++i)
    if (rv[i] != UINT64_MAX) ++off[rv[i] + 1];
  for (size_t j = 0; j < nv; ++j) off[j + 1] += off[j];
  {
    size_t* fill = (size_t*)calloc(nv, sizeof(size_t));
    for (size_t i = 0; i < ntot; ++i)
      if (rv[i] != UINT64_MAX) vals[off[rv[i]] + fill[rv[i]]++] = ra[i];
    free(fill);
  }
  for (size_t j = 0; j < nv; ++j) { /* :112-117 */
    const size_t m = off[j + 1] - off[j];
    double* a = vals + off[j];
    qsort(a, m, sizeof(double), cmp_double);
    /* threshold = (MTindex)(size * fPercentile), x86 cast; the reference
     * reads past the list when it exceeds the size: clamped here */
    uint64_t thr = orc_x86_u32((double)m * cut_percentile);
    if (thr > m) thr = m;
    double sum = 0.0;
    for (uint64_t k = 0; k < thr; ++k) sum += a[k];
    img[j] = wimg[j] < cfg->eps ? 0 : sum / wimg[j];
  }
  free(off); free(vals); free(wimg); free(rv); free(ra);
  pb_free(&pa);
  pb_free(&pbb);
  if (c) {
    c->n_total = (int64_t)ntot;
    c->n_not_parallel = npar;
    c->n_scat_in_grid = nin;
    c->n_out_of_grid = nout;
  }
  return MUTO_OK;
}

/* ------------------------------------------------------------------ */
/* SLR: src/MutoSLR.cpp:47-103                                         */
/* ------------------------------------------------------------------ */
int orc_slr_reconstruct(const muto_poca_cfg* cfg, size_t n,
                        const muto_ray* rays, double* img, muto_counters* c) {
  const muto_grid* g = &cfg->grid;
  const size_t nv = n_vox_of(g);
  if (!img || (n && !rays)) return MUTO_ERR_ARG;
  memset(img, 0, nv * sizeof(double));
  if (c) memset(c, 0, sizeof *c);
  if (n == 0) return MUTO_OK;
  size_t ntot = cfg->total_num_rays > 0 ? (size_t)cfg->total_num_rays : n;
  if (ntot > n) ntot = n;
  double* wimg = (double*)calloc(nv, sizeof(double));
  if (!wimg) return MUTO_ERR_NOMEM;
  const double zlo = g->z_min; /* :54-55 */
  const double zhi = g->z_min + g->dz * (double)g->nz;
  int64_t nout = 0;
  path_buf pa = {0};
  for (size_t i = 0; i < ntot; ++i) { /* :58-96 */
    const muto_ray* r = &rays[i];
    const double axz = (atan2(r->din[0], r->din[2]) - atan2(r->dout[0], r->dout[2])) * cfg->angle_unit_mag;
    const double ayz = (atan2(r->din[1], r->din[2]) - atan2(r->dout[1], r->dout[2])) * cfg->angle_unit_mag;
    const double tin = (zhi - r->pin[2]) / r->din[2];
    const double tout = (zlo - r->pout[2]) / r->dout[2];
    double pi[3], po[3], d[3];
    for (int k = 0; k < 3; ++k) {
      pi[k] = r->pin[k] + tin * r->din[k];
      po[k] = r->pout[k] + tout * r->dout[k];
      d[k] = pi[k] - po[k];
    }
    const size_t k = trace(g, pi, po, &pa);
    const double ltot = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    const double a2h = (axz * axz + ayz * ayz) * 0.5;
    for (size_t m = 0; m < k; ++m) {
      if (!in_grid(g, pa.x[m], pa.y[m], pa.z[m])) { ++nout; continue; }
      const size_t t = tidx(g, pa.x[m], pa.y[m], pa.z[m]);
      double v;
      if (cfg->length_weighted) {
        wimg[t] += pa.l[m] / g->dz;
        v = a2h * (pa.l[m] / ltot) * (pa.l[m] / ltot);
      } else {
        wimg[t] += 1.0;
        v = a2h / (double)(k * k);
      }
      if (cfg->has_energy) v *= r->ein * r->ein;
      img[t] += v;
    }
  }
  for (size_t j = 0; j < nv; ++j) /* :98-100 */
    img[j] = wimg[j] < cfg->eps ? 0 : img[j] / wimg[j];
  free(wimg);
  pb_free(&pa);
  if (c) {
    c->n_total = (int64_t)ntot;
    c->n_out_of_grid = nout;
  }
  return MUTO_OK;
}

/* ------------------------------------------------------------------ */
/* Muon records: src/MutoFile.cpp:19-82, 248-304                       */
/* ------------------------------------------------------------------ */
/* Line through n points: centroid and the principal axis of the centred
 * points (the first right singular vector of the centred data that
 * src/MutoFile.cpp:270-271 takes from Eigen's JacobiSVD), found by cyclic
 * Jacobi rotations of the 3x3 scatter matrix; the sign is fixed to d.z > 0
 * (Eigen's is arbitrary).  The point is the line at height zref
 * (:272-273). */
void orc_fit_line(const double* x, const double* y, const double* z, int n,
                  double zref, double point[3], double dir[3]) {
  double m[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < n; ++i) { m[0] += x[i]; m[1] += y[i]; m[2] += z[i]; }
  for (int k = 0; k < 3; ++k) m[k] /= n;
  double a[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int i = 0; i < n; ++i) {
    const double d[3] = {x[i] - m[0], y[i] - m[1], z[i] - m[2]};
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) a[r][c] += d[r] * d[c];
  }
  double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int sweep = 0; sweep < 64; ++sweep) {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    const double diag = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
    if (off <= 1e-30 * diag || off == 0.0) break;
    for (int p = 0; p < 2; ++p) {
      for (int q = p + 1; q < 3; ++q) {
        if (a[p][q] == 0.0) continue;
        const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) {
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {
          const double vkp = v[k][p], vkq = v[k][q];
          v[k][p] = c * vkp - s * vkq;
          v[k][q] = s * vkp + c * vkq;
        }
      }
    }
  }
  int best = 0;
  for (int k = 1; k < 3; ++k)
    if (a[k][k] > a[best][best]) best = k;
  double d[3] = {v[0][best], v[1][best], v[2][best]};
  const double nrm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  for (int k = 0; k < 3; ++k) d[k] /= nrm;
  if (d[2] < 0.0)
    for (int k = 0; k < 3; ++k) d[k] = -d[k];
  const double t = (zref - m[2]) / d[2];
  dir[0] = d[0]; dir[1] = d[1]; dir[2] = d[2];
  point[0] = t * d[0] + m[0];
  point[1] = t * d[1] + m[1];
  point[2] = zref;
}

int orc_rays_from_records(size_t n_rec, int n_layers, int n_middle,
                          const double* z, const void* records, int precision,
                          muto_ray* rays, size_t* n_rays) {
  if (!z || !n_rays || (n_rec && (!records || !rays))) return MUTO_ERR_ARG;
  if (precision != 4 && precision != 8) return MUTO_ERR_ARG;
  if (n_middle < 2 || n_layers - n_middle < 2 || n_layers > MUTO_MAX_LAYERS) return MUTO_ERR_ARG;
  const size_t each = 2 * (size_t)n_layers + 2; /* :51 */
  double rec[2 * MUTO_MAX_LAYERS + 2], xs[MUTO_MAX_LAYERS], ys[MUTO_MAX_LAYERS];
  size_t out = 0;
  for (size_t r = 0; r < n_rec; ++r) {
    int ok = 1;
    for (size_t k = 0; k < each; ++k) {
      rec[k] = precision == 8 ? ((const double*)records)[r * each + k]
                              : (double)((const float*)records)[r * each + k];
      ok = ok && !isnan(rec[k]); /* :72 */
    }
    if (!ok) continue;
    muto_ray* ray = &rays[out++];
    ray->ein = rec[0];
    for (int i = 0; i < n_layers; ++i) { xs[i] = rec[1 + 2 * i]; ys[i] = rec[2 + 2 * i]; }
    orc_fit_line(xs, ys, z, n_middle, z[n_middle - 1], ray->pin, ray->din);
    orc_fit_line(xs + n_middle, ys + n_middle, z + n_middle, n_layers - n_middle,
                 z[n_middle], ray->pout, ray->dout);
    ray->eout = rec[2 * n_layers + 1];
  }
  *n_rays = out;
  return MUTO_OK;
}

/* ------------------------------------------------------------------ */
/* MLSD preparation: src/MutoMLSD.cpp:268-436                          */
/* ------------------------------------------------------------------ */
static void scattering_angle(const double* d1, const double* d2, double* sx,
                             double* sy) {
  /* :350-354 */
  *sx = (atan2(d1[0], d1[2]) - atan2(d2[0], d2[2]));
  *sy = (atan2(d1[1], d1[2]) - atan2(d2[1], d2[2]));
}

static void displacement(const muto_grid* g, const muto_ray* r, double* dx,
                         double* dy) {
  /* :356-376 */
  double axin = atan2(r->din[0], r->din[2]);
  double axout = atan2(r->dout[0], r->dout[2]);
  double ayin = atan2(r->din[1], r->din[2]);
  double ayout = atan2(r->dout[1], r->dout[2]);
  double sp = (g->z_min - r->pin[2]) / r->din[2];
  double st = (g->z_min - r->pout[2]) / r->dout[2];
  double xp = r->pin[0] + sp * r->din[0];
  double yp = r->pin[1] + sp * r->din[1];
  double xt = r->pout[0] + st * r->dout[0];
  double yt = r->pout[1] + st * r->dout[1];
  double tx = tan(axin), ty = tan(ayin);
  double lxy = sqrt(1 + tx * tx + ty * ty);
  *dx = (xt - xp) * cos(axin) * lxy * cos(axout) / cos(axout - axin);
  *dy = (yt - yp) * cos(ayin) * lxy * cos(ayout) / cos(ayout - ayin);
}

void orc_em_free(orc_em* em) {
  free(em->sel); free(em->trk4); free(em->p); free(em->pscat);
  free(em->row_ptr); free(em->vox); free(em->len); free(em->tail);
  free(em->col_ptr); free(em->slot); free(em->S);
  memset(em, 0, sizeof *em);
}

/* tails (MutoMLSD.cpp:407-413) and voxel-major slots (:432-436) */
static int em_finish(orc_em* em) {
  const muto_grid* g = &em->grid;
  const size_t nv = n_vox_of(g);
  const uint64_t nnz = em->row_ptr[em->n_sel];
  em->tail = (double*)malloc((nnz ? nnz : 1) * sizeof(double));
  em->col_ptr = (uint64_t*)calloc(nv + 1, sizeof(uint64_t));
  em->slot = (uint64_t*)malloc((nnz ? nnz : 1) * sizeof(uint64_t));
  em->S = (double*)calloc(nnz ? nnz : 1, sizeof(double));
  if (!em->tail || !em->col_ptr || !em->slot || !em->S) return MUTO_ERR_NOMEM;
  for (size_t i = 0; i < em->n_sel; ++i) {
    uint64_t b = em->row_ptr[i], e = em->row_ptr[i + 1];
    if (e == b) continue;
    em->tail[e - 1] = 0.0;
    for (uint64_t k = e - 1; k > b; --k)
      em->tail[k - 1] = em->tail[k] + em->len[k];
  }
  for (uint64_t k = 0; k < nnz; ++k)
    if (em->vox[k] < nv) em->col_ptr[em->vox[k] + 1]++;
  for (size_t j = 0; j < nv; ++j) em->col_ptr[j + 1] += em->col_ptr[j];
  uint64_t* fill = (uint64_t*)malloc((nv ? nv : 1) * sizeof(uint64_t));
  if (!fill) return MUTO_ERR_NOMEM;
  memcpy(fill, em->col_ptr, nv * sizeof(uint64_t));
  for (uint64_t k = 0; k < nnz; ++k) /* track order within each voxel */
    em->slot[k] = em->vox[k] < nv ? fill[em->vox[k]]++ : UINT64_MAX;
  free(fill);
  return MUTO_OK;
}

int orc_em_prepare(const muto_mlsd_cfg* cfg, size_t n, const muto_ray* rays,
                   orc_em* em, muto_counters* c) {
  const muto_grid* g = &cfg->grid;
  memset(em, 0, sizeof *em);
  em->grid = *g;
  if (n == 0) {
    em->row_ptr = (uint64_t*)calloc(1, sizeof(uint64_t));
    return em_finish(em);
  }
  size_t ntot = cfg->total_num_rays > 0 ? (size_t)cfg->total_num_rays : n;
  if (ntot > n) ntot = n;
  /* gatherInformation :268-326 */
  const double zmin_det = rays[0].pout[2], zmax_det = rays[0].pin[2];
  const double pnom = 3.0;
  em->sel = (uint32_t*)malloc(ntot * sizeof(uint32_t));
  em->trk4 = (double*)malloc(ntot * 4 * sizeof(double));
  em->p = (double*)malloc(ntot * sizeof(double));
  em->pscat = (double*)malloc(ntot * 3 * sizeof(double));
  if (!em->sel || !em->trk4 || !em->p || !em->pscat) return MUTO_ERR_NOMEM;
  size_t ns = 0;
  for (size_t i = 0; i < ntot; ++i) {
    const muto_ray* r = &rays[i];
    double P[3];
    orc_poca_point(r, cfg->eps, P);
    if (!mlsd_accept(g, P, cfg->eps, zmin_det, zmax_det)) continue;
    em->sel[ns] = (uint32_t)i;
    memcpy(em->pscat + 3 * ns, P, sizeof P);
    scattering_angle(r->din, r->dout, &em->trk4[4 * ns], &em->trk4[4 * ns + 1]);
    displacement(g, r, &em->trk4[4 * ns + 2], &em->trk4[4 * ns + 3]);
    em->p[ns] = cfg->has_energy ? pnom / r->ein : pnom / pnom;
    ++ns;
  }
  em->n_sel = ns;
  /* getMuonPathAndTail :382-415 for every selected track */
  em->row_ptr = (uint64_t*)malloc((ns + 1) * sizeof(uint64_t));
  if (!em->row_ptr) return MUTO_ERR_NOMEM;
  path_buf pa = {0}, pbb = {0};
  size_t cap = 1024;
  em->vox = (uint32_t*)malloc(cap * sizeof(uint32_t));
  em->len = (double*)malloc(cap * sizeof(double));
  uint64_t nnz = 0;
  int64_t nempty = 0, nout = 0;
  em->row_ptr[0] = 0;
  for (size_t s = 0; s < ns; ++s) {
    const muto_ray* r = &rays[em->sel[s]];
    const double* P = em->pscat + 3 * s;
    size_t need;
    size_t ki = 0, ko = 0;
    if (cfg->straight_line) {
      ki = trace(g, r->pin, r->pout, &pa);
      need = ki;
    } else {
      ki = trace(g, r->pin, P, &pa);
      ko = trace(g, P, r->pout, &pbb);
      need = ki + ko;
    }
    if (nnz + need > cap) {
      while (nnz + need > cap) cap *= 2;
      em->vox = (uint32_t*)realloc(em->vox, cap * sizeof(uint32_t));
      em->len = (double*)realloc(em->len, cap * sizeof(double));
      if (!em->vox || !em->len) return MUTO_ERR_NOMEM;
    }
#define PUSHV(X, Y, Z, L)                                                   \
  do {                                                                      \
    if (!in_grid(g, X, Y, Z)) ++nout;                                       \
    em->vox[nnz] = (X) + (Y) * g->nx + (Z) * g->nx * g->ny;                 \
    em->len[nnz] = (L);                                                     \
    ++nnz;                                                                  \
  } while (0)
    if (cfg->straight_line) {
      for (size_t m = 0; m < ki; ++m) PUSHV(pa.x[m], pa.y[m], pa.z[m], pa.l[m]);
    } else {
      if (ki != 0) {
        pa.l[ki - 1] += ko != 0 ? pbb.l[0] : 0.0;
        for (size_t m = 0; m < ki; ++m)
          PUSHV(pa.x[m], pa.y[m], pa.z[m], pa.l[m]);
      }
      for (size_t m = 1; m < ko; ++m)
        PUSHV(pbb.x[m], pbb.y[m], pbb.z[m], pbb.l[m]);
    }
#undef PUSHV
    /* An empty merged path makes the reference call tail.back() on an
     * empty vector (undefined); such a track contributes no entries here
     * and is counted. */
    if (nnz == em->row_ptr[s]) ++nempty;
    em->row_ptr[s + 1] = nnz;
  }
  pb_free(&pa);
  pb_free(&pbb);
  em->n_empty = nempty;
  em->n_out_of_grid = nout;
  if (c) {
    memset(c, 0, sizeof *c);
    c->n_total = (int64_t)ntot;
    c->n_scat_in_grid = (int64_t)ns;
    c->n_empty_path = nempty;
    c->n_out_of_grid = nout;
    c->nnz = (int64_t)nnz;
  }
  return em_finish(em);
}

int orc_em_from_prepared(const muto_grid* g, size_t n_trk, const double* trk4,
                         const double* p, const uint64_t* row_ptr,
                         const uint32_t* vox, const double* len, orc_em* em) {
  memset(em, 0, sizeof *em);
  em->grid = *g;
  em->n_sel = n_trk;
  uint64_t nnz = row_ptr[n_trk];
  em->trk4 = (double*)malloc((n_trk ? n_trk : 1) * 4 * sizeof(double));
  em->p = (double*)malloc((n_trk ? n_trk : 1) * sizeof(double));
  em->row_ptr = (uint64_t*)malloc((n_trk + 1) * sizeof(uint64_t));
  em->vox = (uint32_t*)malloc((nnz ? nnz : 1) * sizeof(uint32_t));
  em->len = (double*)malloc((nnz ? nnz : 1) * sizeof(double));
  if (!em->trk4 || !em->p || !em->row_ptr || !em->vox || !em->len)
    return MUTO_ERR_NOMEM;
  memcpy(em->trk4, trk4, n_trk * 4 * sizeof(double));
  for (size_t i = 0; i < n_trk; ++i) em->p[i] = p ? p[i] : 1.0;
  memcpy(em->row_ptr, row_ptr, (n_trk + 1) * sizeof(uint64_t));
  memcpy(em->vox, vox, nnz * sizeof(uint32_t));
  memcpy(em->len, len, nnz * sizeof(double));
  return em_finish(em);
}

/* ------------------------------------------------------------------ */
/* EM iteration: src/MutoMLSD.cpp:179-213, 442-517                     */
/* ------------------------------------------------------------------ */
typedef struct { double m00, m01, m10, m11; } mat2;

static mat2 error_matrix(const muto_mlsd_cfg* cfg) {
  /* :168-170 */
  const double e = cfg->error_xy, o = cfg->error_dz_outer,
               i = cfg->error_dz_inner;
  mat2 E;
  E.m00 = 4.0 * e * e / (o * o);
  E.m11 = 2.0 * e * e * (1.0 + (i / o + 1.0) * i / o);
  E.m01 = E.m10 = 2.0 * e * e * i / (o * o);
  return E;
}

static inline void weight(double L, double T, mat2* W) {
  /* generateWeightMatrix :417-430 */
  double LT = L * T;
  W->m00 = L;
  W->m01 = W->m10 = LT + L * L * 0.5;
  W->m11 = L * L * L / 3.0 + (L + T) * LT;
}

/* getCovarianceMatrix :442-449 plus ErrorM */
static mat2 covariance(const orc_em* em, size_t i, const double* img,
                       double p, const mat2* E) {
  const size_t nv = n_vox_of(&em->grid);
  mat2 cov = {0.0, 0.0, 0.0, 0.0};
  for (uint64_t k = em->row_ptr[i]; k < em->row_ptr[i + 1]; ++k) {
    mat2 W;
    weight(em->len[k], em->tail[k], &W);
    double lj = em->vox[k] < nv ? img[em->vox[k]] : 0.0;
    cov.m00 = cov.m00 + lj * W.m00;
    cov.m01 = cov.m01 + lj * W.m01;
    cov.m10 = cov.m10 + lj * W.m10;
    cov.m11 = cov.m11 + lj * W.m11;
  }
  double pp = p * p;
  cov.m00 = pp * cov.m00 + E->m00;
  cov.m01 = pp * cov.m01 + E->m01;
  cov.m10 = pp * cov.m10 + E->m10;
  cov.m11 = pp * cov.m11 + E->m11;
  return cov;
}

/* computeInverseAndDetWithCheck for 2x2 (Eigen: determinant
 * m00*m11 - m10*m01, invertible iff |det| > 1e-12, inverse = adj * 1/det;
 * the inverse is left untouched (zero) when not invertible). */
static int inverse_det(const mat2* m, mat2* inv, double* det) {
  *det = m->m00 * m->m11 - m->m10 * m->m01;
  inv->m00 = inv->m01 = inv->m10 = inv->m11 = 0.0;
  if (!(fabs(*det) > 1e-12)) return 0;
  double invdet = 1.0 / *det;
  inv->m00 = m->m11 * invdet;
  inv->m10 = -m->m10 * invdet;
  inv->m01 = -m->m01 * invdet;
  inv->m11 = m->m00 * invdet;
  return 1;
}

/* updateImageVoxel :493-517 (values in S[0..mj), may be reordered) */
static double voxel_update(const muto_mlsd_cfg* cfg, double* S, uint64_t mj) {
  if (mj == 0) return 0.0;
  if (cfg->update_with_mean) {
    if (cfg->mean_percentile == 100) {
      double s = 0.0;
      for (uint64_t k = 0; k < mj; ++k) s = s + S[k];
      return s / (double)mj * 0.5;
    }
    uint32_t jm = orc_x86_u32(cfg->mean_percentile / 100.0 * (double)mj);
    if (jm < 1) jm = 1;
    qsort(S, mj, sizeof(double), cmp_double);
    double s = 0.0;
    for (uint32_t k = 0; k < jm; ++k) s = s + S[k];
    return s / (double)jm * 0.5;
  }
  /* median: nth_element picks order statistics; a full sort gives the
   * same values */
  qsort(S, mj, sizeof(double), cmp_double);
  double med = S[mj / 2];
  if (mj % 2 == 0) return (med + S[mj / 2 - 1]) * 0.25;
  return med * 0.5;
}

double orc_em_iterate(const muto_mlsd_cfg* cfg, orc_em* em,
                      const double* img_in, double* img_out,
                      int64_t* n_singular) {
  const muto_grid* g = &em->grid;
  const size_t nv = n_vox_of(g);
  const mat2 E = error_matrix(cfg);
  double cost_sum = 0.0;
  int64_t nsing = 0;
  for (size_t i = 0; i < em->n_sel; ++i) { /* :188-195 serial, track order */
    const double p = em->p[i];
    mat2 cov = covariance(em, i, img_in, p, &E);
    /* updateSValueMatrix :451-491 */
    const double Sx = em->trk4[4 * i], Sy = em->trk4[4 * i + 1];
    const double Dx = em->trk4[4 * i + 2], Dy = em->trk4[4 * i + 3];
    double cost = 0.0, det;
    mat2 inv;
    int ok = inverse_det(&cov, &inv, &det);
    if (!ok || det <= 0) {
      ++nsing;
      cost = 0.0;
    } else {
      cost = log(det) + ((Sx * Sx + Sy * Sy) * inv.m00 +
                         (Dx * Dx + Dy * Dy) * inv.m11 +
                         (Sx * Dx + Sy * Dy) * 2.0 * inv.m01) * 0.5;
    }
    for (uint64_t k = em->row_ptr[i]; k < em->row_ptr[i + 1]; ++k) {
      if (em->vox[k] >= nv) continue;
      mat2 W;
      weight(em->len[k], em->tail[k], &W);
      double lj = img_in[em->vox[k]];
      double pr2lj2 = p * p * lj * lj;
      /* (SigmaInv * W).trace() */
      double P00 = inv.m00 * W.m00 + inv.m01 * W.m10;
      double P01 = inv.m00 * W.m01 + inv.m01 * W.m11;
      double P10 = inv.m10 * W.m00 + inv.m11 * W.m10;
      double P11 = inv.m10 * W.m01 + inv.m11 * W.m11;
      double tr = P00 + P11;
      /* SigmaInv * W * SigmaInv */
      double C00 = P00 * inv.m00 + P01 * inv.m10;
      double C01 = P00 * inv.m01 + P01 * inv.m11;
      double C10 = P10 * inv.m00 + P11 * inv.m10;
      double C11 = P10 * inv.m01 + P11 * inv.m11;
      double xy = (Sx * Sx + Sy * Sy) * C00 + (Dx * Dx + Dy * Dy) * C11 +
                  (Sx * Dx + Sy * Dy) * (C01 + C10);
      double sij = 2.0 * lj + (0.5 * xy - tr) * pr2lj2;
      em->S[em->slot[k]] = sij;
    }
    cost_sum += cost;
  }
#pragma omp parallel for schedule(dynamic, 4096)
  for (size_t j = 0; j < nv; ++j) /* :197-200 */
    img_out[j] = voxel_update(cfg, em->S + em->col_ptr[j],
                              em->col_ptr[j + 1] - em->col_ptr[j]);
  if (n_singular) *n_singular = nsing;
  return cost_sum / (double)em->n_sel; /* :202 */
}

void orc_em_mj(const orc_em* em, uint32_t* mj) {
  size_t nv = n_vox_of(&em->grid);
  for (size_t j = 0; j < nv; ++j)
    mj[j] = (uint32_t)(em->col_ptr[j + 1] - em->col_ptr[j]);
}

/* Compact form: c_ij = p^2 <W_ij, G_i>, G = 1/2 Si^-1 D Si^-1 - Si^-1;
 * cost_i as the reference (0 when singular).  Derived from :451-491. */
void orc_em_cij(const muto_mlsd_cfg* cfg, const orc_em* em,
                const double* img, double* cij, double* cost_i) {
  const size_t nv = n_vox_of(&em->grid);
  const mat2 E = error_matrix(cfg);
  for (size_t i = 0; i < em->n_sel; ++i) {
    const double p = em->p[i];
    mat2 cov = covariance(em, i, img, p, &E);
    const double Sx = em->trk4[4 * i], Sy = em->trk4[4 * i + 1];
    const double Dx = em->trk4[4 * i + 2], Dy = em->trk4[4 * i + 3];
    double det;
    mat2 inv;
    int ok = inverse_det(&cov, &inv, &det);
    const double D00 = Sx * Sx + Sy * Sy, D11 = Dx * Dx + Dy * Dy,
                 D01 = Sx * Dx + Sy * Dy;
    if (cost_i)
      cost_i[i] = (!ok || det <= 0)
                      ? 0.0
                      : log(det) + (D00 * inv.m00 + D11 * inv.m11 +
                                    D01 * 2.0 * inv.m01) * 0.5;
    /* M = Si^-1 D Si^-1 */
    double A00 = inv.m00 * D00 + inv.m01 * D01;
    double A01 = inv.m00 * D01 + inv.m01 * D11;
    double A10 = inv.m10 * D00 + inv.m11 * D01;
    double A11 = inv.m10 * D01 + inv.m11 * D11;
    double M00 = A00 * inv.m00 + A01 * inv.m10;
    double M01 = A00 * inv.m01 + A01 * inv.m11;
    double M11 = A10 * inv.m01 + A11 * inv.m11;
    double G00 = 0.5 * M00 - inv.m00, G01 = 0.5 * M01 - inv.m01,
           G11 = 0.5 * M11 - inv.m11;
    for (uint64_t k = em->row_ptr[i]; k < em->row_ptr[i + 1]; ++k) {
      if (em->vox[k] >= nv) { cij[k] = 0.0; continue; }
      mat2 W;
      weight(em->len[k], em->tail[k], &W);
      cij[k] = p * p * (W.m00 * G00 + 2.0 * W.m01 * G01 + W.m11 * G11);
    }
  }
}

/* initailizeImage :226-266 for methods 1 and 2.  Method 1 draws from
 * std::random_device in the reference (not reproducible); here a fixed
 * seed gives N(2, 0.5)*1e-7 floored at 1e-10 by Box-Muller. */
static void init_image(const muto_mlsd_cfg* cfg, const double* init_img,
                       double* img) {
  const size_t nv = n_vox_of(&cfg->grid);
  if (cfg->initialize_method == 0 && init_img) {
    memcpy(img, init_img, nv * sizeof(double));
    return;
  }
  if (cfg->initialize_method == 1 ||
      (cfg->initialize_method == 0 && !init_img)) {
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

int orc_mlsd_reconstruct(const muto_mlsd_cfg* cfg, size_t n,
                         const muto_ray* rays, const double* init_img,
                         double* img_out, double* cost_per_iter,
                         muto_counters* c) {
  /* :91-220 */
  orc_em em;
  int rc = orc_em_prepare(cfg, n, rays, &em, c);
  if (rc) { orc_em_free(&em); return rc; }
  const size_t nv = n_vox_of(&cfg->grid);
  double* img = img_out;
  double* tmp = (double*)malloc(nv * sizeof(double));
  if (!tmp) { orc_em_free(&em); return MUTO_ERR_NOMEM; }
  init_image(cfg, init_img, img);
  double cost_last, cost = INFINITY; /* uninitialised in the reference */
  int64_t nsing = 0;
  int it;
  for (it = 0; it < cfg->num_iteration; ++it) {
    cost_last = cost;
    cost = orc_em_iterate(cfg, &em, img, tmp, &nsing);
    memcpy(img, tmp, nv * sizeof(double));
    if (cost_per_iter) cost_per_iter[it] = cost;
    if (cfg->early_stop && cost_last < cost) { ++it; break; }
  }
  if (c) { c->n_singular = nsing; c->iterations_done = it; }
  free(tmp);
  orc_em_free(&em);
  return MUTO_OK;
}

/* ------------------------------------------------------------------ */
/* MAP: src/MutoMAP.cpp                                                */
/* ------------------------------------------------------------------ */
#define MAP_SCALE 1.0e7 /* fUnitConv :33 */

uint32_t orc_cubic_solve(double* roots, double p3, double p2, double p1,
                         double p0, double tol, int* warn) {
  /* :353-407.  The Newton loop is unbounded in the reference; 10000 steps
   * is far above what converging inputs use, and the cap is counted. */
  if (warn) *warn = 0;
  if (fabs(p3) < tol && warn) *warn = 1;
  double a = p2 / p3, b = p1 / p3, c = p0 / p3;
  double r = 1 + smax(smax(fabs(a), fabs(b)), fabs(c));
  double xinfl = -a / 3;
  double gx = xinfl * xinfl * xinfl + a * xinfl * xinfl + b * xinfl + c;
  double x0 = 0 < gx ? -r : r;
  double xk = x0, rel = tol * 2;
  int guard = 0;
  while (rel > tol) {
    double xk2 = xk * xk;
    double gxk = xk2 * xk + a * xk2 + b * xk + c;
    double gpxk = 3 * xk2 + 2 * a * xk + b;
    double xn = xk - gxk / gpxk;
    rel = fabs((xn - xk) / xk);
    xk = xn;
    if (++guard >= 10000) { if (warn) *warn |= 2; break; }
  }
  double x1 = xk;
  double D = x1 + a;
  double E = D * x1 + b;
  double signb = 0 <= D ? 1 : -1;
  double diak = D * D - 4 * E;
  if (diak < 0) { roots[0] = x1; return 1; }
  double q = -0.5 * (D + signb * sqrt(diak));
  double x2 = q;
  double x3 = tol < fabs(q) ? E / q : x2;
  roots[0] = x1; roots[1] = x2; roots[2] = x3;
  return 3;
}

int orc_map_neighbours(const muto_mlsd_cfg* cfg, const double* img,
                       uint32_t j, double* w, double* lm) {
  /* index2xyz :283-290; getSixNeighbours / getAllNeighbours :292-350 */
  const muto_grid* g = &cfg->grid;
  uint32_t x = j % (g->nx * g->ny) % g->nx;
  uint32_t y = j % (g->nx * g->ny) / g->nx;
  uint32_t z = j / (g->nx * g->ny);
  uint32_t xp = x < g->nx - 1 ? x + 1 : g->nx - 1;
  uint32_t yp = y < g->ny - 1 ? y + 1 : g->ny - 1;
  uint32_t zp = z < g->nz - 1 ? z + 1 : g->nz - 1;
  uint32_t xm = 0 < x ? x - 1 : 0;
  uint32_t ym = 0 < y ? y - 1 : 0;
  uint32_t zm = 0 < z ? z - 1 : 0;
  const uint32_t nb[26][3] = {
      {xp, y, z},   {xm, y, z},   {x, yp, z},   {x, ym, z},   {x, y, zp},
      {x, y, zm},   {xp, yp, z},  {xm, yp, z},  {xp, ym, z},  {xm, ym, z},
      {x, yp, zp},  {x, ym, zp},  {x, yp, zm},  {x, ym, zm},  {xp, y, zp},
      {xm, y, zp},  {xp, y, zm},  {xm, y, zm},  {xp, yp, zp}, {xm, yp, zp},
      {xp, ym, zp}, {xm, ym, zp}, {xp, yp, zm}, {xm, yp, zm}, {xp, ym, zm},
      {xm, ym, zm}};
  int nn = cfg->all_neighbours ? 26 : 6;
  const double w2 = 1.0 / sqrt(2.0), w3 = 1.0 / sqrt(3.0); /* :210-212 */
  for (int k = 0; k < nn; ++k) {
    w[k] = k < 6 ? 1.0 : (k < 18 ? w2 : w3);
    lm[k] = img[tidx(g, nb[k][0], nb[k][1], nb[k][2])] * MAP_SCALE;
  }
  return nn;
}

double orc_map_surrogate(const muto_mlsd_cfg* cfg, uint32_t mj, double ljn,
                         double ljml, double lj, const double* w,
                         const double* lm, int nn) {
  /* surrogateFunctionValue :261-281 */
  double value = 0.0;
  if (lj < 0) return DBL_MAX;
  value += 0 < lj ? (double)mj * (log(lj) + ljml / lj) : 0.0;
  const double ex = (cfg->p - 2.0) / 2.0, dd = cfg->delta * cfg->delta;
  for (int k = 0; k < nn; ++k) {
    double t = ljn - lm[k];
    double u = 2.0 * lj - ljn - lm[k];
    value += 0.5 * cfg->beta * cfg->p * w[k] * pow(t * t + dd, ex) * (u * u);
  }
  if (cfg->add_object_prior)
    value += 0.5 * cfg->beta * cfg->p * lj * lj * pow(ljn * ljn + dd, ex);
  return value;
}

void orc_map_voxel_step(const muto_mlsd_cfg* cfg, const uint32_t* mj,
                        const double* img_last, const double* img_ml,
                        double* img_out, int64_t* n_cubic_warn) {
  const size_t nv = n_vox_of(&cfg->grid);
  int64_t nwarn = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : nwarn)
  for (size_t j = 0; j < nv; ++j) { /* :164-184 */
    if (mj[j] == 0) { img_out[j] = 0.0; continue; }
    double p0 = -img_ml[j] * MAP_SCALE, p1 = 1.0, p2 = 0.0, p3 = 0.0;
    double w[26], lm[26];
    int nn = orc_map_neighbours(cfg, img_last, (uint32_t)j, w, lm);
    /* calculateCoefFromNeighbours :207-237 */
    double lj = img_last[j] * MAP_SCALE;
    const double ex = (cfg->p - 2.0) / 2.0, dd = cfg->delta * cfg->delta;
    for (int k = 0; k < nn; ++k) {
      double t = lj - lm[k];
      double t1 = w[k] * cfg->p * pow(t * t + dd, ex);
      p2 += t1 * 2.0 * (lj + lm[k]);
      p3 += t1 * 4.0;
    }
    if (cfg->add_object_prior) p3 += cfg->p * pow(lj * lj + dd, ex);
    p2 *= -cfg->beta / (double)mj[j];
    p3 *= cfg->beta / (double)mj[j];
    double roots[3];
    int warn;
    uint32_t nr = orc_cubic_solve(roots, p3, p2, p1, p0, cfg->eps, &warn);
    nwarn += (warn & 1);
    /* chooseBestEstimation :239-259 (first minimum wins) */
    double ljn = img_last[j] * MAP_SCALE, ljml = img_ml[j] * MAP_SCALE;
    double best = 0.0;
    uint32_t bi = 0;
    for (uint32_t r = 0; r <= nr; ++r) {
      double cand = r < nr ? roots[r] : ljml;
      double f = orc_map_surrogate(cfg, mj[j], ljn, ljml, cand, w, lm, nn);
      if (r == 0 || f < best) { best = f; bi = r; }
    }
    img_out[j] = bi < nr ? roots[bi] / MAP_SCALE : img_ml[j];
  }
  if (n_cubic_warn) *n_cubic_warn = nwarn;
}

int orc_map_reconstruct(const muto_mlsd_cfg* cfg, size_t n,
                        const muto_ray* rays, const double* init_img,
                        double* img_out, double* cost_per_iter,
                        muto_counters* c) {
  /* :39-205 */
  orc_em em;
  int rc = orc_em_prepare(cfg, n, rays, &em, c);
  if (rc) { orc_em_free(&em); return rc; }
  const size_t nv = n_vox_of(&cfg->grid);
  double* img = img_out;
  double* last = (double*)malloc(nv * sizeof(double));
  double* ml = (double*)malloc(nv * sizeof(double));
  uint32_t* mj = (uint32_t*)malloc(nv * sizeof(uint32_t));
  if (!last || !ml || !mj) { orc_em_free(&em); return MUTO_ERR_NOMEM; }
  init_image(cfg, init_img, img);
  orc_em_mj(&em, mj);
  double cost_last, cost = INFINITY;
  int64_t nsing = 0, nwarn = 0;
  int it;
  for (it = 0; it < cfg->num_iteration; ++it) {
    memcpy(last, img, nv * sizeof(double));
    cost_last = cost;
    cost = orc_em_iterate(cfg, &em, last, ml, &nsing);
    orc_map_voxel_step(cfg, mj, last, ml, img, &nwarn);
    if (cost_per_iter) cost_per_iter[it] = cost;
    if (cfg->early_stop && cost_last < cost) { ++it; break; }
  }
  if (c) {
    c->n_singular = nsing;
    c->n_cubic_warn = nwarn;
    c->iterations_done = it;
  }
  free(last); free(ml); free(mj);
  orc_em_free(&em);
  return MUTO_OK;
}

/* ------------------------------------------------------------------ */
/* MAP objective and gradient (derived from :451-491, :207-281)        */
/*   Phi(lam) = 1/2 sum_i cost_i(lam)                                  */
/*            + beta sum_j sum_{k in N(j)} w_jk psi(l_j - l_k)         */
/*            + [object prior] beta sum_j psi(l_j),                    */
/*   l = 1e7 lam, psi(t) = (t^2 + delta^2)^(p/2), neighbours clamped   */
/*   at the edges, sums over all voxels.  In mean mode the reference's */
/*   surrogate (:261-281) is tangent to Phi at l_n for every voxel     */
/*   with M_j > 0, so grad Phi there is the reference's MAP gradient.  */
/* ------------------------------------------------------------------ */
static inline double psi(double t, double p, double dd) {
  return pow(t * t + dd, p / 2.0);
}
static inline double dpsi(double t, double p, double dd) {
  return p * t * pow(t * t + dd, (p - 2.0) / 2.0);
}

double orc_map_phi(const muto_mlsd_cfg* cfg, const orc_em* em,
                   const double* lambda) {
  const size_t nv = n_vox_of(&em->grid);
  double* ci = (double*)malloc((em->n_sel ? em->n_sel : 1) * sizeof(double));
  uint64_t nnz = em->row_ptr[em->n_sel];
  double* cij = (double*)malloc((nnz ? nnz : 1) * sizeof(double));
  orc_em_cij(cfg, em, lambda, cij, ci);
  double phi = 0.0;
  for (size_t i = 0; i < em->n_sel; ++i) phi += 0.5 * ci[i];
  free(ci); free(cij);
  const double dd = cfg->delta * cfg->delta;
  double prior = 0.0;
  for (size_t j = 0; j < nv; ++j) {
    double w[26], lm[26];
    int nn = orc_map_neighbours(cfg, lambda, (uint32_t)j, w, lm);
    double lj = lambda[j] * MAP_SCALE;
    for (int k = 0; k < nn; ++k) prior += w[k] * psi(lj - lm[k], cfg->p, dd);
    if (cfg->add_object_prior) prior += psi(lj, cfg->p, dd);
  }
  return phi + cfg->beta * prior;
}

void orc_map_grad(const muto_mlsd_cfg* cfg, const orc_em* em,
                  const double* lambda, double* grad) {
  const size_t nv = n_vox_of(&em->grid);
  const muto_grid* g = &em->grid;
  uint64_t nnz = em->row_ptr[em->n_sel];
  double* cij = (double*)malloc((nnz ? nnz : 1) * sizeof(double));
  orc_em_cij(cfg, em, lambda, cij, NULL);
  memset(grad, 0, nv * sizeof(double));
  /* d cost_i / d lam_j = -c_ij for invertible Sigma with det > 0 */
  const mat2 E = error_matrix(cfg);
  for (size_t i = 0; i < em->n_sel; ++i) {
    mat2 cov = covariance(em, i, lambda, em->p[i], &E);
    double det;
    mat2 inv;
    int ok = inverse_det(&cov, &inv, &det);
    if (!ok || det <= 0) continue;
    for (uint64_t k = em->row_ptr[i]; k < em->row_ptr[i + 1]; ++k)
      if (em->vox[k] < nv) grad[em->vox[k]] += -0.5 * cij[k];
  }
  free(cij);
  /* prior: each voxel j with M_j > 0 contributes beta*w*psi(l_j - l_k) for
   * each of its neighbour slots k (k may be j itself at an edge). */
  const double dd = cfg->delta * cfg->delta;
  for (size_t j = 0; j < nv; ++j) {
    uint32_t x = (uint32_t)(j % (g->nx * g->ny) % g->nx);
    uint32_t y = (uint32_t)(j % (g->nx * g->ny) / g->nx);
    uint32_t z = (uint32_t)(j / (g->nx * g->ny));
    uint32_t xp = x < g->nx - 1 ? x + 1 : g->nx - 1;
    uint32_t yp = y < g->ny - 1 ? y + 1 : g->ny - 1;
    uint32_t zp = z < g->nz - 1 ? z + 1 : g->nz - 1;
    uint32_t xm = 0 < x ? x - 1 : 0, ym = 0 < y ? y - 1 : 0,
             zm = 0 < z ? z - 1 : 0;
    const uint32_t nb[26][3] = {
        {xp, y, z},   {xm, y, z},   {x, yp, z},   {x, ym, z},   {x, y, zp},
        {x, y, zm},   {xp, yp, z},  {xm, yp, z},  {xp, ym, z},  {xm, ym, z},
        {x, yp, zp},  {x, ym, zp},  {x, yp, zm},  {x, ym, zm},  {xp, y, zp},
        {xm, y, zp},  {xp, y, zm},  {xm, y, zm},  {xp, yp, zp}, {xm, yp, zp},
        {xp, ym, zp}, {xm, ym, zp}, {xp, yp, zm}, {xm, yp, zm}, {xp, ym, zm},
        {xm, ym, zm}};
    int nn = cfg->all_neighbours ? 26 : 6;
    const double w2 = 1.0 / sqrt(2.0), w3 = 1.0 / sqrt(3.0);
    double lj = lambda[j] * MAP_SCALE;
    for (int k = 0; k < nn; ++k) {
      size_t jk = tidx(g, nb[k][0], nb[k][1], nb[k][2]);
      double w = k < 6 ? 1.0 : (k < 18 ? w2 : w3);
      double dv = w * dpsi(lj - lambda[jk] * MAP_SCALE, cfg->p, dd);
      grad[j] += cfg->beta * MAP_SCALE * dv;
      grad[jk] -= cfg->beta * MAP_SCALE * dv;
    }
    if (cfg->add_object_prior)
      grad[j] += cfg->beta * MAP_SCALE * dpsi(lj, cfg->p, dd);
  }
}

void orc_map_surrogate_grad(const muto_mlsd_cfg* cfg, orc_em* em,
                            const double* lambda, double* grad) {
  /* Reference surrogate (:261-281) differentiated at l = l_n by central
   * differences, with l_ML from one mean-mode ML step (:493-500). */
  const size_t nv = n_vox_of(&em->grid);
  muto_mlsd_cfg mc = *cfg;
  mc.update_with_mean = 1;
  mc.mean_percentile = 100;
  double* ml = (double*)malloc(nv * sizeof(double));
  uint32_t* mj = (uint32_t*)malloc(nv * sizeof(uint32_t));
  orc_em_iterate(&mc, em, lambda, ml, NULL);
  orc_em_mj(em, mj);
  for (size_t j = 0; j < nv; ++j) {
    if (mj[j] == 0) { grad[j] = 0.0; continue; }
    double w[26], lm[26];
    int nn = orc_map_neighbours(cfg, lambda, (uint32_t)j, w, lm);
    double ln = lambda[j] * MAP_SCALE, lml = ml[j] * MAP_SCALE;
    double h = 1e-5 * fabs(ln) + 1e-12;
    double fp = orc_map_surrogate(cfg, mj[j], ln, lml, ln + h, w, lm, nn);
    double fm = orc_map_surrogate(cfg, mj[j], ln, lml, ln - h, w, lm, nn);
    /* Q is in l units and its ML part is M_j(ln l + l_ML/l); 1/2 of it is
     * the 1/2 sum_i cost_i share, the pair part carries beta*p terms.  The
     * derivative is reported in lambda units. */
    grad[j] = (fp - fm) / (2.0 * h) * MAP_SCALE;
  }
  free(ml); free(mj);
}
