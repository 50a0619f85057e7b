// MAP objective as a differentiable operator: value and gradient with
// respect to the voxel scattering densities, behind a forward/backward pair
// with a saved context.
//
//   Phi(lambda) = 1/2 sum_i cost_i(lambda) + beta R(lambda)
//
// cost_i is the EM iteration's per-track cost (MutoMLSD.cpp:456-468: log det
// Sigma_i + 1/2 <D_i, Sigma_i^-1>, 0 for a singular or non-positive Sigma_i)
// and R the MAP prior (MutoMAP.cpp:207-237): over every voxel j and its 6 or
// 26 clamp-to-edge neighbours k, w_jk psi(l_j - l_k), plus psi(l_j) with the
// object prior, psi(t) = (t^2 + delta^2)^(p/2), l = 1e7 lambda.  The
// reference never evaluates Phi or its gradient; its MAP step minimises a
// surrogate that is tangent to Phi in mean mode.
//
// Gradient: d cost_i / d lambda_j = -<W_ij, G_i> with
// G_i = p_i^2 (1/2 Sigma_i^-1 D_i Sigma_i^-1 - Sigma_i^-1), the transpose of
// the Siddon-weighted forward map through each track's covariance.  The
// forward pass stores G_i per track; the backward pass scatters
// -1/2 <W_ij, G_i> over each track's path (FP64 atomics) and adds the prior
// gradient.  FP64 reads the stored track-major lists or, when the plan keeps
// none (or the caller asks), walks every path again by Siddon; FP32C runs
// the brick-stream passes of the EM iteration.
#include <math.h>

#include <algorithm>
#include <vector>

#include "em_dev.hpp"
#include "em_plan.hpp"

namespace muto {
int fp32c_build(muto_em_plan* pl);
int fp32c_map_forward(muto_em_plan* pl, const double* lam, float* g_out,
                      unsigned long long* cmax_out, double* cost_sum);
int fp32c_map_backward(muto_em_plan* pl, const float* g, const unsigned long long* cmax,
                       double* acc);
}  // namespace muto

using namespace muto;

struct muto_map_saved {
  muto_em_plan* pl = nullptr;
  int precision = MUTO_FP64;
  int recompute = 0;
  double phi = 0.0;
  DBuf<double> lam;                // [nv] the lambda of the forward call
  DBuf<double> g64;                // FP64: [4 n_trk] G00, G01, G11, path length
  DBuf<float> g32;                 // FP32C: [4 n_trk] the track kernel's G'
  DBuf<unsigned long long> cmax;   // FP32C: the |c_ij| bound of that pass
  DBuf<double> grad;               // [nv] device gradient
  DBuf<double> sum;                // [2] cost sum, prior sum
  DBuf<double> part;               // per-block partial sums
};

namespace {

// Per track: Sigma_i from lambda (stored lists in the reference's order, or
// the path walked again with moments about its start moved to the tail
// basis), Eigen's 2x2 det/inverse, cost_i, and G_i (zero where cost_i is 0).
template <int RC>
__global__ void __launch_bounds__(256)
k_map_fwd64(size_t n_trk, const uint64_t* __restrict__ row_ptr,
            const uint32_t* __restrict__ vox, const double* __restrict__ len,
            const double* __restrict__ tail, GridD g, const double* __restrict__ geo,
            int straight, const double* __restrict__ trk4, const double* __restrict__ pr,
            const double* __restrict__ lam, size_t nv, ErrM E, double* __restrict__ gsave,
            double* __restrict__ cost_part, unsigned long long* __restrict__ cnt) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double cost = 0.0;
  bool sing = false;
  if (i < n_trk) {
    Mat2 cov = {0.0, 0.0, 0.0, 0.0};
    double P = 0.0;
    if (RC == 0) {
      for (uint64_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
        Mat2 W;
        weight(len[k], tail[k], W);
        const uint32_t j = vox[k];
        const double lj = j < nv ? lam[j] : 0.0;
        cov.m00 = cov.m00 + lj * W.m00;
        cov.m01 = cov.m01 + lj * W.m01;
        cov.m10 = cov.m10 + lj * W.m10;
        cov.m11 = cov.m11 + lj * W.m11;
      }
    } else {
      // moments in tau = path length from the start: int (1, tau)(1, tau)^T
      double a0 = 0.0, a1 = 0.0, a2 = 0.0;
      P = walk_path(g, geo + 9 * i, straight, [&](uint32_t j, double L, double s, bool) {
        const double lj = j < nv ? lam[j] : 0.0;
        a0 += lj * L;
        a1 += lj * (L * s + 0.5 * L * L);
        a2 += lj * (L * s * s + L * L * s + L * L * L / 3.0);
      });
      // tail t = P - tau: u(t) = A u(tau), A = [[1, 0], [P, -1]]
      cov.m00 = a0;
      cov.m01 = cov.m10 = P * a0 - a1;
      cov.m11 = P * (P * a0 - 2.0 * a1) + a2;
    }
    const double p = pr[i], pp = p * p;
    cov.m00 = pp * cov.m00 + E.e00;
    cov.m01 = pp * cov.m01 + E.e01;
    cov.m10 = pp * cov.m10 + E.e01;
    cov.m11 = pp * cov.m11 + E.e11;
    const double Sx = trk4[4 * i], Sy = trk4[4 * i + 1];
    const double Dx = trk4[4 * i + 2], Dy = trk4[4 * i + 3];
    const double d00 = Sx * Sx + Sy * Sy, d11 = Dx * Dx + Dy * Dy, d01 = Sx * Dx + Sy * Dy;
    Mat2 inv;
    double det;
    const bool ok = inv_det(cov, inv, det);
    double G00 = 0.0, G01 = 0.0, G11 = 0.0;
    if (!ok || det <= 0) {
      sing = true;
    } else {
      cost = log(det) + (d00 * inv.m00 + d11 * inv.m11 + d01 * 2.0 * inv.m01) * 0.5;
      // G = p^2 (1/2 inv D inv - inv)
      const double A00 = inv.m00 * d00 + inv.m01 * d01, A01 = inv.m00 * d01 + inv.m01 * d11;
      const double A10 = inv.m10 * d00 + inv.m11 * d01, A11 = inv.m10 * d01 + inv.m11 * d11;
      G00 = pp * (0.5 * (A00 * inv.m00 + A01 * inv.m10) - inv.m00);
      G01 = pp * (0.5 * (A00 * inv.m01 + A01 * inv.m11) - inv.m01);
      G11 = pp * (0.5 * (A10 * inv.m01 + A11 * inv.m11) - inv.m11);
    }
    gsave[4 * i] = G00;
    gsave[4 * i + 1] = G01;
    gsave[4 * i + 2] = G11;
    gsave[4 * i + 3] = P;
  }
  block_sum_to(cost_part, cost);
  wave_count(&cnt[C_SINGULAR], sing);
}

// Per track: grad_j += scale <W_ij, G_i> over its path (scale = -1/2 grad_phi).
template <int RC>
__global__ void __launch_bounds__(256)
k_map_bwd64(size_t n_trk, const uint64_t* __restrict__ row_ptr,
            const uint32_t* __restrict__ vox, const double* __restrict__ len,
            const double* __restrict__ tail, GridD g, const double* __restrict__ geo,
            int straight, const double* __restrict__ gsave, size_t nv, double scale,
            double* __restrict__ grad) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_trk) return;
  const double G00 = gsave[4 * i], G01 = gsave[4 * i + 1], G11 = gsave[4 * i + 2];
  if (G00 == 0.0 && G01 == 0.0 && G11 == 0.0) return;
  auto add = [&](uint32_t j, double L, double T) {
    Mat2 W;
    weight(L, T, W);
    atomicAdd(&grad[j], scale * (W.m00 * G00 + 2.0 * W.m01 * G01 + W.m11 * G11));
  };
  if (RC == 0) {
    for (uint64_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
      const uint32_t j = vox[k];
      if (j < nv) add(j, len[k], tail[k]);
    }
  } else {
    const double P = gsave[4 * i + 3];
    walk_path(g, geo + 9 * i, straight, [&](uint32_t j, double L, double s, bool) {
      if (j < nv) add(j, L, P - s - L);
    });
  }
}

__device__ __forceinline__ double psi(double t, double p, double dd) {
  return pow(t * t + dd, p / 2.0);
}
__device__ __forceinline__ double dpsi(double t, double p, double dd) {
  return p * t * pow(t * t + dd, (p - 2.0) / 2.0);
}

// prior R(lambda), summed over all voxels
__global__ void __launch_bounds__(256)
k_prior_phi(MapCfgD c, size_t nv, const double* __restrict__ lam, double* __restrict__ part) {
  size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (j < nv) {
    double w[26], lm[26];
    const int nn = map_neighbours(c, lam, (uint32_t)j, w, lm);
    const double lj = lam[j] * MAP_SCALE, dd = c.delta * c.delta;
    for (int k = 0; k < nn; ++k) v += w[k] * psi(lj - lm[k], c.p, dd);
    if (c.obj_prior) v += psi(lj, c.p, dd);
  }
  block_sum_to(part, v);
}

// out[0] = sum of part[0..n) in a fixed order (one block of 256)
__global__ void __launch_bounds__(256)
k_sum_parts(const double* __restrict__ part, size_t n, double* __restrict__ out) {
  double v = 0.0;
  for (size_t k = threadIdx.x; k < n; k += blockDim.x) v += part[k];
  block_sum_to(out, v);
}

// grad += scale dR/dlambda (scale = grad_phi beta): each pair term
// w psi(l_j - l_k) adds to j and subtracts from k (FP64 atomics)
__global__ void __launch_bounds__(256)
k_prior_grad(MapCfgD c, size_t nv, const double* __restrict__ lam, double scale,
             double* __restrict__ grad) {
  size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nv) return;
  size_t idx[26];
  double w[26];
  const int nn = map_neighbour_index(c, (uint32_t)j, idx, w);
  const double lj = lam[j] * MAP_SCALE, dd = c.delta * c.delta;
  const double s = scale * MAP_SCALE;
  double own = 0.0;
  for (int k = 0; k < nn; ++k) {
    const double dv = w[k] * dpsi(lj - lam[idx[k]] * MAP_SCALE, c.p, dd);
    own += s * dv;
    atomicAdd(&grad[idx[k]], -s * dv);
  }
  if (c.obj_prior) own += s * dpsi(lj, c.p, dd);
  atomicAdd(&grad[j], own);
}

__global__ void k_scale(size_t n, double a, double* __restrict__ x) {
  size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) x[j] *= a;
}

}  // namespace

extern "C" int muto_map_forward(muto_em_plan* pl, const double* lambda, int precision, int flags,
                                double* phi, muto_map_saved** saved) {
  if (!pl) return MUTO_ERR_ARG;
  muto_ctx* ctx = pl->ctx;
  if (!lambda || !phi || !saved) { ctx->set_error("muto_map_forward: null argument"); return MUTO_ERR_ARG; }
  *saved = nullptr;
  if (precision != MUTO_FP64 && precision != MUTO_FP32C) {
    ctx->set_error("muto_map_forward: precision must be MUTO_FP64 or MUTO_FP32C");
    return MUTO_ERR_ARG;
  }
  const bool recompute = (flags & MUTO_MAP_RECOMPUTE_SIDDON) || !pl->has_lists;
  if (recompute && precision != MUTO_FP64) {
    ctx->set_error("muto_map_forward: recomputed paths run in MUTO_FP64 (the FP32C brick stream is "
                   "built from stored lists)");
    return MUTO_ERR_ARG;
  }
  if (recompute && !pl->geo.p && pl->n_trk) {
    ctx->set_error("muto_map_forward: the plan has no track geometry to walk (prepared plan)");
    return MUTO_ERR_STATE;
  }
  hipStream_t s = ctx->stream;
  auto* sv = new muto_map_saved();
  sv->pl = pl;
  sv->precision = precision;
  sv->recompute = recompute;
  const size_t nv = pl->nv, nt = pl->n_trk;
  auto fail = [&](int rc) { delete sv; return rc; };
  const size_t n_part = std::max<size_t>(grid_for(nt, 256), grid_for(nv, 256));
  if (sv->lam.alloc(nv) != hipSuccess || sv->sum.alloc(2) != hipSuccess ||
      sv->grad.alloc(nv) != hipSuccess || sv->part.alloc(n_part) != hipSuccess) {
    ctx->set_error("muto_map_forward: out of device memory");
    return fail(MUTO_ERR_NOMEM);
  }
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(sv->lam.p, lambda, nv * sizeof(double), hipMemcpyHostToDevice, s));
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(sv->sum.p, 0, 2 * sizeof(double), s));
  double cost_sum = 0.0;
  if (nt) {
    if (precision == MUTO_FP64) {
      if (sv->g64.alloc(4 * nt) != hipSuccess) { ctx->set_error("muto_map_forward: out of device memory"); return fail(MUTO_ERR_NOMEM); }
      const ErrM E = error_matrix(pl->cfg);
      if (recompute)
        k_map_fwd64<1><<<grid_for(nt, 256), 256, 0, s>>>(nt, nullptr, nullptr, nullptr, nullptr, pl->g,
            pl->geo.p, pl->cfg.straight_line, pl->trk4.p, pl->pr.p, sv->lam.p, nv, E, sv->g64.p,
            sv->part.p, pl->dcount.p);
      else
        k_map_fwd64<0><<<grid_for(nt, 256), 256, 0, s>>>(nt, pl->row_ptr.p, pl->vox.p, pl->len.p,
            pl->tail.p, pl->g, nullptr, 0, pl->trk4.p, pl->pr.p, sv->lam.p, nv, E, sv->g64.p,
            sv->part.p, pl->dcount.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
      k_sum_parts<<<1, 256, 0, s>>>(sv->part.p, grid_for(nt, 256), sv->sum.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
      MUTO_CHECK_HIP(ctx, hipMemcpyAsync(&cost_sum, sv->sum.p, sizeof(double), hipMemcpyDeviceToHost, s));
      MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
    } else {
      if (int rc = fp32c_build(pl)) return fail(rc);
      if (sv->g32.alloc(4 * nt) != hipSuccess || sv->cmax.alloc(1) != hipSuccess) {
        ctx->set_error("muto_map_forward: out of device memory");
        return fail(MUTO_ERR_NOMEM);
      }
      if (int rc = fp32c_map_forward(pl, sv->lam.p, sv->g32.p, sv->cmax.p, &cost_sum)) return fail(rc);
    }
  }
  double prior = 0.0;
  k_prior_phi<<<grid_for(nv, 256), 256, 0, s>>>(map_cfg(pl), nv, sv->lam.p, sv->part.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  k_sum_parts<<<1, 256, 0, s>>>(sv->part.p, grid_for(nv, 256), sv->sum.p + 1);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(&prior, sv->sum.p + 1, sizeof(double), hipMemcpyDeviceToHost, s));
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  sv->phi = 0.5 * cost_sum + pl->cfg.beta * prior;
  *phi = sv->phi;
  *saved = sv;
  return MUTO_OK;
}

extern "C" int muto_map_backward(muto_map_saved* sv, double grad_phi, double* grad_lambda) {
  if (!sv || !sv->pl) return MUTO_ERR_ARG;
  muto_em_plan* pl = sv->pl;
  muto_ctx* ctx = pl->ctx;
  if (!grad_lambda) { ctx->set_error("muto_map_backward: null argument"); return MUTO_ERR_ARG; }
  hipStream_t s = ctx->stream;
  const size_t nv = pl->nv, nt = pl->n_trk;
  MUTO_CHECK_HIP(ctx, hipMemsetAsync(sv->grad.p, 0, nv * sizeof(double), s));
  if (nt) {
    if (sv->precision == MUTO_FP64) {
      if (sv->recompute)
        k_map_bwd64<1><<<grid_for(nt, 256), 256, 0, s>>>(nt, nullptr, nullptr, nullptr, nullptr, pl->g,
            pl->geo.p, pl->cfg.straight_line, sv->g64.p, nv, -0.5 * grad_phi, sv->grad.p);
      else
        k_map_bwd64<0><<<grid_for(nt, 256), 256, 0, s>>>(nt, pl->row_ptr.p, pl->vox.p, pl->len.p,
            pl->tail.p, pl->g, nullptr, 0, sv->g64.p, nv, -0.5 * grad_phi, sv->grad.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    } else {
      if (int rc = fp32c_map_backward(pl, sv->g32.p, sv->cmax.p, sv->grad.p)) return rc;
      k_scale<<<grid_for(nv, 256), 256, 0, s>>>(nv, -0.5 * grad_phi, sv->grad.p);
      MUTO_CHECK_HIP(ctx, hipGetLastError());
    }
  }
  k_prior_grad<<<grid_for(nv, 256), 256, 0, s>>>(map_cfg(pl), nv, sv->lam.p, grad_phi * pl->cfg.beta,
                                                 sv->grad.p);
  MUTO_CHECK_HIP(ctx, hipGetLastError());
  MUTO_CHECK_HIP(ctx, hipMemcpyAsync(grad_lambda, sv->grad.p, nv * sizeof(double), hipMemcpyDeviceToHost, s));
  MUTO_CHECK_HIP(ctx, hipStreamSynchronize(s));
  return MUTO_OK;
}

extern "C" int muto_map_saved_destroy(muto_map_saved* sv) {
  if (!sv) return MUTO_OK;
  if (sv->pl && sv->pl->ctx) (void)hipStreamSynchronize(sv->pl->ctx->stream);
  delete sv;
  return MUTO_OK;
}
