// Seeded synthetic cosmic-muon generator (host side).  Declared in
// include/muto_gen.h; the physics choices are listed there.
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/muto_gen.h"

namespace {

inline uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// One independent stream per (seed, candidate index).
struct Rng {
  uint64_t s;
  bool has_spare = false;
  double spare = 0.0;
  Rng(uint64_t seed, uint64_t idx) {
    s = seed ^ (idx * 0xD1B54A32D192ED03ull + 0x8CB92BA72F3D8DD7ull);
    splitmix64(s);
  }
  double uniform() {  // (0, 1)
    return ((splitmix64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  }
  double normal() {
    if (has_spare) { has_spare = false; return spare; }
    double u1 = uniform(), u2 = uniform();
    double r = sqrt(-2.0 * log(u1)), a = 6.283185307179586 * u2;
    spare = r * sin(a);
    has_spare = true;
    return r * cos(a);
  }
};

struct Layout {
  std::vector<uint8_t> mat;  // index into x0 table, 0 = air
  std::vector<double> x0;
};

Layout build_layout(const muto_gen_cfg* c) {
  Layout L;
  const muto_grid& g = c->grid;
  L.x0.push_back(c->air_x0);
  size_t nv = (size_t)g.nx * g.ny * g.nz;
  L.mat.assign(nv, 0);
  for (int b = 0; b < c->n_blocks; ++b) {
    const muto_gen_block& B = c->blocks[b];
    L.x0.push_back(B.x0_mm);
    uint8_t id = (uint8_t)std::min<size_t>(L.x0.size() - 1, 255);
    for (uint32_t z = B.z0; z < std::min(B.z1, g.nz); ++z)
      for (uint32_t y = B.y0; y < std::min(B.y1, g.ny); ++y)
        for (uint32_t x = B.x0; x < std::min(B.x1, g.nx); ++x)
          L.mat[(size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z)] = id;
  }
  return L;
}

// Highland theta0 (rad) for thickness t of radiation length X0, momentum p
// (GeV/c, beta = 1).
inline double highland(double t_mm, double x0, double p_gev) {
  double r = t_mm / x0;
  if (r <= 0) return 0.0;
  double f = 1.0 + 0.038 * log(r);
  if (f < 0) f = 0;
  return 13.6 / (p_gev * 1000.0) * sqrt(r) * f;
}

// Least-squares line x = a + b z, y = c + d z through k points.
void fit_line(const double* z, const double* x, const double* y, int k,
              double z_at, double* p, double* dir) {
  double zm = 0, xm = 0, ym = 0;
  for (int i = 0; i < k; ++i) { zm += z[i]; xm += x[i]; ym += y[i]; }
  zm /= k; xm /= k; ym /= k;
  double szz = 0, szx = 0, szy = 0;
  for (int i = 0; i < k; ++i) {
    szz += (z[i] - zm) * (z[i] - zm);
    szx += (z[i] - zm) * (x[i] - xm);
    szy += (z[i] - zm) * (y[i] - ym);
  }
  double bx = szx / szz, by = szy / szz;
  double nrm = sqrt(bx * bx + by * by + 1.0);
  dir[0] = bx / nrm; dir[1] = by / nrm; dir[2] = 1.0 / nrm;
  p[0] = xm + bx * (z_at - zm);
  p[1] = ym + by * (z_at - zm);
  p[2] = z_at;
}

// Voxels crossed by the straight chord a->b inside the grid (1 + plane
// crossings); 0 if it misses.
int chord_voxels(const muto_grid& g, const double* a, const double* b) {
  const double mn[3] = {g.x_min, g.y_min, g.z_min};
  const double d[3] = {g.dx, g.dy, g.dz};
  const uint32_t n[3] = {g.nx, g.ny, g.nz};
  double t0 = 0.0, t1 = 1.0;
  for (int k = 0; k < 3; ++k) {
    double dv = b[k] - a[k];
    double lo = mn[k], hi = mn[k] + d[k] * n[k];
    if (fabs(dv) < 1e-12) {
      if (a[k] < lo || a[k] >= hi) return 0;
      continue;
    }
    double ta = (lo - a[k]) / dv, tb = (hi - a[k]) / dv;
    if (ta > tb) std::swap(ta, tb);
    t0 = std::max(t0, ta);
    t1 = std::min(t1, tb);
  }
  if (t1 <= t0) return 0;
  long cnt = 1;
  for (int k = 0; k < 3; ++k) {
    double u0 = (a[k] + t0 * (b[k] - a[k]) - mn[k]) / d[k];
    double u1 = (a[k] + t1 * (b[k] - a[k]) - mn[k]) / d[k];
    long i0 = (long)floor(std::min(std::max(u0, 0.0), n[k] - 1e-9));
    long i1 = (long)floor(std::min(std::max(u1, 0.0), n[k] - 1e-9));
    cnt += labs(i1 - i0);
  }
  return (int)cnt;
}

// One candidate; returns true if accepted.
bool make_muon(const muto_gen_cfg* c, const Layout& L, uint64_t idx,
               muto_ray* out) {
  const muto_grid& g = c->grid;
  Rng rng(c->seed, idx);
  const double xlo = g.x_min, xhi = g.x_min + g.dx * g.nx;
  const double ylo = g.y_min, yhi = g.y_min + g.dy * g.ny;
  const double zlo = g.z_min, zhi = g.z_min + g.dz * g.nz;
  const double z_top = zhi + c->tracker_gap, z_bot = zlo - c->tracker_gap;
  // direction
  double c0 = cos(c->theta_max_deg * M_PI / 180.0);
  double c03 = c0 * c0 * c0;
  double ct = cbrt(c03 + rng.uniform() * (1.0 - c03));
  double st = sqrt(std::max(0.0, 1.0 - ct * ct));
  double phi = 2.0 * M_PI * rng.uniform();
  double lnp = log(c->p_median) + c->p_sigma_ln * rng.normal();
  double p = std::min(std::max(exp(lnp), c->p_min), c->p_max);
  // slopes per unit descent: x(z) = x(z_ref) + sx * (z_ref - z)
  double sx = st * cos(phi) / ct, sy = st * sin(phi) / ct;
  // mid-plane crossing uniform over the volume square
  double zmid = 0.5 * (zlo + zhi);
  double xm = xlo + (xhi - xlo) * rng.uniform();
  double ym = ylo + (yhi - ylo) * rng.uniform();
  // position at the top face of the volume (straight above it)
  double x = xm + sx * (zmid - zhi), y = ym + sy * (zmid - zhi);
  // top tracker hits: lines above the volume are straight
  double zt[3], xt[3], yt[3];
  for (int k = 0; k < 3; ++k) {
    zt[k] = z_top + c->layer_spacing * (2 - k);
    xt[k] = x + sx * (zhi - zt[k]) + c->smear_sigma * rng.normal();
    yt[k] = y + sy * (zhi - zt[k]) + c->smear_sigma * rng.normal();
  }
  // propagate down through the volume slab layer by layer; material runs
  // are scattered as one Highland step each
  double tx = atan(sx), ty = atan(sy);
  int cur = -1;
  double run = 0.0;  // path length of the current run
  auto kick = [&](int m, double len) {
    if (m < 0 || len <= 0) return;
    double th0 = highland(len, L.x0[m], p);
    for (int pl = 0; pl < 2; ++pl) {
      double z1 = rng.normal(), z2 = rng.normal();
      double off = len * th0 * (z1 / sqrt(12.0) + z2 / 2.0);
      if (pl == 0) { x += off; tx += z2 * th0; }
      else { y += off; ty += z2 * th0; }
    }
  };
  double zc = zhi;
  for (uint32_t layer = 0; layer < g.nz; ++layer) {
    double dzs = g.dz;
    double sxc = tan(tx), syc = tan(ty);
    double xmid = x + sxc * (0.5 * dzs), ymid = y + syc * (0.5 * dzs);
    double zmidl = zc - 0.5 * dzs;
    int m = 0;
    if (xmid >= xlo && xmid < xhi && ymid >= ylo && ymid < yhi) {
      uint32_t ix = (uint32_t)((xmid - xlo) / g.dx);
      uint32_t iy = (uint32_t)((ymid - ylo) / g.dy);
      uint32_t iz = (uint32_t)((zmidl - zlo) / g.dz);
      if (ix < g.nx && iy < g.ny && iz < g.nz)
        m = L.mat[(size_t)ix + (size_t)g.nx * ((size_t)iy + (size_t)g.ny * iz)];
    }
    if (m != cur) {
      kick(cur, run);
      cur = m;
      run = 0.0;
      sxc = tan(tx); syc = tan(ty);
    }
    double seg = dzs * sqrt(1.0 + sxc * sxc + syc * syc);
    run += seg;
    x += sxc * dzs;
    y += syc * dzs;
    zc -= dzs;
  }
  kick(cur, run);
  double sxo = tan(tx), syo = tan(ty);
  double zb[3], xb[3], yb[3];
  for (int k = 0; k < 3; ++k) {
    zb[k] = z_bot - c->layer_spacing * k;
    xb[k] = x + sxo * (zlo - zb[k]) + c->smear_sigma * rng.normal();
    yb[k] = y + syo * (zlo - zb[k]) + c->smear_sigma * rng.normal();
  }
  muto_ray r;
  r.ein = r.eout = p;
  fit_line(zt, xt, yt, 3, z_top, r.pin, r.din);
  fit_line(zb, xb, yb, 3, z_bot, r.pout, r.dout);
  if (c->min_voxels > 0 || c->max_voxels > 0) {
    int nvx = chord_voxels(g, r.pin, r.pout);
    if (nvx == 0) return false;
    if (c->min_voxels > 0 && nvx < c->min_voxels) return false;
    if (c->max_voxels > 0 && nvx > c->max_voxels) return false;
  } else if (chord_voxels(g, r.pin, r.pout) == 0) {
    return false;
  }
  *out = r;
  return true;
}

}  // namespace

extern "C" {

void muto_gen_default(muto_gen_cfg* cfg, const muto_grid* grid,
                      uint64_t seed) {
  memset(cfg, 0, sizeof *cfg);
  cfg->grid = *grid;
  cfg->seed = seed;
  cfg->tracker_gap = 500.0;
  cfg->layer_spacing = 500.0;
  cfg->smear_sigma = 0.5;
  cfg->p_median = 3.0;
  cfg->p_sigma_ln = 0.8;
  cfg->p_min = 0.5;
  cfg->p_max = 100.0;
  cfg->theta_max_deg = 60.0;
  cfg->air_x0 = 303900.0;
}

int muto_gen_layout(const muto_grid* grid, uint64_t seed, int n,
                    double min_mm, double max_mm, const double* x0_choices,
                    int n_choices, muto_gen_block* out) {
  if (!grid || n < 0 || (n > 0 && (!out || !x0_choices || n_choices <= 0)))
    return MUTO_ERR_ARG;
  Rng rng(seed, 0xB10C);
  const double d[3] = {grid->dx, grid->dy, grid->dz};
  const uint32_t nn[3] = {grid->nx, grid->ny, grid->nz};
  for (int b = 0; b < n; ++b) {
    double side = min_mm + (max_mm - min_mm) * rng.uniform();
    uint32_t lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
      uint32_t s = (uint32_t)llround(side / d[k]);
      if (s < 1) s = 1;
      if (s > nn[k]) s = nn[k];
      uint32_t span = nn[k] - s;
      uint32_t o = (uint32_t)floor(rng.uniform() * (span + 1));
      if (o > span) o = span;
      lo[k] = o;
      hi[k] = o + s;
    }
    out[b].x0 = lo[0]; out[b].y0 = lo[1]; out[b].z0 = lo[2];
    out[b].x1 = hi[0]; out[b].y1 = hi[1]; out[b].z1 = hi[2];
    out[b].x0_mm = x0_choices[b % n_choices];
  }
  return MUTO_OK;
}

int muto_generate(const muto_gen_cfg* cfg, int64_t n, muto_ray* out,
                  int64_t* n_tried) {
  if (!cfg || n < 0 || (n > 0 && !out)) return MUTO_ERR_ARG;
  if (cfg->n_blocks > 0 && !cfg->blocks) return MUTO_ERR_ARG;
  if (cfg->n_blocks > 254) return MUTO_ERR_LIMIT;
  Layout L = build_layout(cfg);
  int64_t have = 0;
  uint64_t next = 0;
  const int64_t batch = 1 << 16;
  std::vector<muto_ray> buf(batch);
  std::vector<uint8_t> ok(batch);
  while (have < n) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t k = 0; k < batch; ++k)
      ok[k] = make_muon(cfg, L, next + (uint64_t)k, &buf[k]) ? 1 : 0;
    for (int64_t k = 0; k < batch && have < n; ++k) {
      if (ok[k]) out[have++] = buf[k];
      if (have == n && n_tried) *n_tried = (int64_t)(next + k + 1);
    }
    next += batch;
    if (next > (uint64_t)n * 1000 + (1u << 24)) return MUTO_ERR_LIMIT;
  }
  if (n == 0 && n_tried) *n_tried = 0;
  return MUTO_OK;
}

int muto_gen_truth(const muto_gen_cfg* cfg, double* img) {
  if (!cfg || !img) return MUTO_ERR_ARG;
  Layout L = build_layout(cfg);
  const double k = (13.6 / 3000.0) * (13.6 / 3000.0);
  for (size_t j = 0; j < L.mat.size(); ++j) img[j] = k / L.x0[L.mat[j]];
  return MUTO_OK;
}

}  // extern "C"
