// Likelihood derivatives fused into the Newton step.
//
// Gaussian: log N(y; f, sigma^2) (probit/utilities.py:60-70).
// Probit:   log(Phi(z2) - Phi(z1) + 1e-10), z = (b - f) / sigma, with the
//           +-inf interval ends guarded (probit/utilities.py:47-57, 195-229);
//           Phi in the reference's erf form (probit/utilities.py:18-19).
//           Derivatives are the exact ones JAX autodiff yields for that
//           expression (probit/approximators.py:92-95).
// Closed:   grad_/hessian_log_probit_likelihood with the piecewise BOUNDS
//           tails (probit/utilities.py:15, 37-192).
#include <math.h>

#include "pgp_internal.hpp"

namespace pgp {

constexpr double OVER_SQRT_2PI = 0.39894228040143267794;   // probit/utilities.py:10
constexpr double LOG_OVER_SQRT_2PI = -0.91893853320467274178;  // probit/utilities.py:11
constexpr double SQRT_2 = 1.4142135623730951;  // probit/utilities.py:12
constexpr double PROBIT_FLOOR = 1e-10;         // probit/utilities.py:57

struct LikParams {
  int kind, J, sp;
  double sigma;
  const double* cut;  // J + 1 cutpoints (global or LDS)
};

struct LikDer {
  double ll, d1, d2, d3;
  // partials in the likelihood parameters: (L_p, L_fp, L_ffp)
  double ps[3], plo[3], phi[3];
};

__device__ __forceinline__ double ndtr_ref(double z) { return 0.5 * (1.0 + erf(z / SQRT_2)); }
__device__ __forceinline__ double npdf(double z) { return OVER_SQRT_2PI * exp(-0.5 * z * z); }

__device__ __forceinline__ void gaussian_der(double f, double y, double s, int order, LikDer& o) {
  const double z = (f - y) / s;
  o.ll = LOG_OVER_SQRT_2PI - z * z / 2.0 - log(s);
  const double r = y - f;
  o.d1 = r / (s * s);
  o.d2 = -1.0 / (s * s);
  o.d3 = 0.0;
  if (order >= 3) {
    o.ps[0] = r * r / (s * s * s) - 1.0 / s;
    o.ps[1] = -2.0 * r / (s * s * s);
    o.ps[2] = 2.0 / (s * s * s);
  }
}

__device__ __forceinline__ void probit_der(const LikParams& lp, double f, int y, int order,
                                           LikDer& o) {
  const double s = lp.sigma;
  const double blo = lp.cut[y], bhi = lp.cut[y + 1];
  const bool lo_inf = isinf(blo) && blo < 0;
  const bool hi_inf = isinf(bhi) && bhi > 0;
  const double z1 = lo_inf ? 0.0 : (blo - f) / s;
  const double z2 = hi_inf ? 0.0 : (bhi - f) / s;
  const double c1 = lo_inf ? 0.0 : ndtr_ref(z1);
  const double c2 = hi_inf ? 1.0 : ndtr_ref(z2);
  const double Ze = (c2 - c1) + PROBIT_FLOOR;
  const double p1 = lo_inf ? 0.0 : npdf(z1);
  const double p2 = hi_inf ? 0.0 : npdf(z2);
  const double Zf = (p1 - p2) / s;
  const double Zff = (z1 * p1 - z2 * p2) / (s * s);
  const double L1 = Zf / Ze;
  const double L2 = Zff / Ze - L1 * L1;
  o.ll = log(Ze);
  o.d1 = L1;
  o.d2 = L2;
  o.d3 = 0.0;
  if (order >= 3) {
    const double s3 = s * s * s;
    const double Zfff = ((z1 * z1 - 1.0) * p1 - (z2 * z2 - 1.0) * p2) / s3;
    o.d3 = Zfff / Ze - 3.0 * L1 * L2 - L1 * L1 * L1;
    auto param = [&](double Zp, double Zfp, double Zffp, double* out) {
      const double Lp = Zp / Ze;
      const double Lfp = Zfp / Ze - L1 * Lp;
      out[0] = Lp;
      out[1] = Lfp;
      out[2] = Zffp / Ze - Zff * Zp / (Ze * Ze) - 2.0 * L1 * Lfp;
    };
    param((z1 * p1 - z2 * p2) / s, ((z1 * z1 - 1.0) * p1 - (z2 * z2 - 1.0) * p2) / (s * s),
          ((z1 * z1 * z1 - 3.0 * z1) * p1 - (z2 * z2 * z2 - 3.0 * z2) * p2) / s3, o.ps);
    param(-p1 / s, -z1 * p1 / (s * s), (1.0 - z1 * z1) * p1 / s3, o.plo);
    param(p2 / s, z2 * p2 / (s * s), -(1.0 - z2 * z2) * p2 / s3, o.phi);
  }
}

// --- closed-form variant (probit/utilities.py:37-192) ---
__device__ __forceinline__ double h_series(double x) {
  // probit/utilities.py:37-44
  const double x2 = 1.0 / (x * x);
  return -1.0 * x2 + 5.0 / 2.0 * x2 * x2 - 37.0 / 3.0 * x2 * x2 * x2;
}
__device__ __forceinline__ double Z_far_tails(double z) {
  // probit/utilities.py:83-85
  if (isinf(z)) return 0.0;
  return OVER_SQRT_2PI / z * exp(-0.5 * z * z + h_series(z));
}
__device__ __forceinline__ double norm_cdf_inf(double x) {
  // probit/utilities.py:31-34
  if (isinf(x)) return x > 0 ? 1.0 : 0.0;
  return ndtr_ref(x);
}
__device__ __forceinline__ double npdf_inf(double z) { return isinf(z) ? 0.0 : npdf(z); }

__device__ __forceinline__ void probit_closed_der(const LikParams& lp, double f, int y,
                                                  LikDer& o) {
  const double ub = lp.sp ? 1.3 : 2.3, ub2 = lp.sp ? 1.8 : 3.6, ub3 = lp.sp ? 2.3 : 4.8;
  const double s = lp.sigma;
  const double chi = lp.cut[y + 1], clo = lp.cut[y];
  const bool hi_inf = isinf(chi) && chi > 0;
  const bool lo_inf = isinf(clo) && clo < 0;
  // _safe_Z (probit/utilities.py:88-148)
  const double z2s = hi_inf ? INFINITY : (chi - f) / s;
  const double z1s = lo_inf ? -INFINITY : (clo - f) / s;
  const double SAFE = 1.0;
  double Z = norm_cdf_inf(z2s) - norm_cdf_inf(z1s);
  {
    const double a1 = (ub < z1s && z1s <= ub2) ? z1s : SAFE;
    const double a2 = (ub < z1s) ? z2s : SAFE;
    const double b2 = (-ub2 <= z2s && z2s < -ub) ? z2s : SAFE;
    const double b1 = (-ub > z2s) ? z1s : SAFE;
    if (z1s > ub) Z = Z_far_tails(a1) - Z_far_tails(a2);
    if (z2s < -ub) Z = Z_far_tails(b1) - Z_far_tails(b2);
    const double c1 = (ub2 < fabs(z1s) && fabs(z1s) < ub3) ? z1s : SAFE;
    const double c2 = (ub2 < fabs(z2s) && fabs(z2s) < ub3) ? z2s : SAFE;
    if (z1s > ub2) Z = Z_far_tails(c1);
    if (z2s < -ub2) Z = Z_far_tails(-c2);
    if (z1s >= ub3) Z = SAFE;
    if (z2s <= -ub3) Z = SAFE;
  }
  const double p1 = npdf_inf(z1s), p2 = npdf_inf(z2s);
  // grad (probit/utilities.py:151-169)
  double E = (p1 - p2) / Z;
  if (z1s > ub3) E = z1s;
  if (z2s < -ub3) E = z2s;
  const double w = E / s;
  // hessian (probit/utilities.py:172-192)
  const double _z1 = isinf(z1s) ? 0.0 : z1s;
  const double _z2 = isinf(z2s) ? 0.0 : z2s;
  double V = -(w * w) + (_z1 * p1 - _z2 * p2) / Z / (s * s);
  if (z1s > ub3) V = -1.0 / (s * s);
  if (z2s < -ub3) V = -1.0 / (s * s);
  // objective still uses log_probit_likelihood (probit/approximators.py:96-98)
  LikDer ex;
  probit_der(lp, f, y, 2, ex);
  o.ll = ex.ll;
  o.d1 = w;
  o.d2 = V;
  o.d3 = 0.0;
}

__device__ __forceinline__ void lik_der(const LikParams& lp, double f, double y, int order,
                                        LikDer& o) {
  if (lp.kind == PGP_LIK_GAUSSIAN) {
    gaussian_der(f, y, lp.sigma, order, o);
  } else {
    int yi = (int)y;
    yi = yi < 0 ? 0 : (yi >= lp.J ? lp.J - 1 : yi);
    if (lp.kind == PGP_LIK_PROBIT)
      probit_der(lp, f, yi, order, o);
    else
      probit_closed_der(lp, f, yi, o);
  }
}

__device__ __forceinline__ LikParams likparams_from_prob(const Prob& P) {
  LikParams lp;
  lp.kind = P.lik_kind;
  lp.J = P.J;
  lp.sp = P.sp_bounds;
  lp.sigma = P.noise_std;
  lp.cut = P.cut;
  return lp;
}

// ---- Newton step: F -> G, W, S, LL, R = G - Z ----
__global__ __launch_bounds__(NTHREADS) void k_lik_newton(Prob* const* probs, Sched sc,
                                                         int final_order) {
  int bx;
  const Prob& P = *probs[sched_member(sc, blockIdx.x, bx)];
  const int i = bx * NTHREADS + threadIdx.x;
  if (i >= P.Np) return;
  const LikParams lp = likparams_from_prob(P);
  double* F = vslot(P, V_F);
  double g = 0.0, w = 0.0, ll = 0.0, d3 = 0.0;
  if (i < P.N) {
    LikDer o;
    lik_der(lp, F[i], P.y[i], final_order >= 3 ? 3 : 2, o);
    g = o.d1;
    w = -o.d2;
    ll = o.ll;
    d3 = o.d3;
    if (w < 0.0) atomicAdd(&P.info[1], 1);
  }
  vslot(P, V_G)[i] = g;
  vslot(P, V_W)[i] = w;
  vslot(P, V_S)[i] = sqrt(fmax(w, 0.0));
  vslot(P, V_LL)[i] = ll;
  if (final_order == 0) vslot(P, V_R)[i] = g - vslot(P, V_Z)[i];
  if (final_order >= 3) vslot(P, V_D3)[i] = d3;
}

static void lik_launch(const Batch& b, int order, hipStream_t s) {
  sched_launch(
      b, [&](int m) { return (b.np(m) + NTHREADS - 1) / NTHREADS; },
      [&](Prob* const* dp, const Sched& sc, int n) {
        hipLaunchKernelGGL(k_lik_newton, dim3(n), dim3(NTHREADS), 0, s, dp, sc, order);
      });
}

void launch_lik_newton(const Batch& b, hipStream_t s) { lik_launch(b, 0, s); }

void launch_lik_final(const Batch& b, int order3, hipStream_t s) {
  lik_launch(b, order3 ? 3 : 2, s);
}

// ---- raw elementwise evaluation (kernel-level ABI) ----
struct LikRaw {
  int kind, J, sp;
  double sigma;
  double cut[PGP_MAX_CUT];
};

__global__ __launch_bounds__(NTHREADS) void k_lik_raw(LikRaw lr, const double* f,
                                                      const double* y, int64_t N, double* ll,
                                                      double* d1, double* d2) {
  __shared__ double cut[PGP_MAX_CUT];
  if (threadIdx.x < PGP_MAX_CUT) cut[threadIdx.x] = lr.cut[threadIdx.x];
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * NTHREADS + threadIdx.x;
  if (i >= N) return;
  LikParams lp;
  lp.kind = lr.kind;
  lp.J = lr.J;
  lp.sp = lr.sp;
  lp.sigma = lr.sigma;
  lp.cut = cut;
  LikDer o;
  lik_der(lp, f[i], y[i], 2, o);
  if (ll) ll[i] = o.ll;
  if (d1) d1[i] = o.d1;
  if (d2) d2[i] = o.d2;
}

static LikRaw likraw_from_api(const pgp_likelihood_t* l) {
  LikRaw lp;
  lp.kind = l->kind;
  lp.J = l->J;
  lp.sp = l->single_precision_bounds;
  lp.sigma = l->noise_std;
  for (int c = 0; c < PGP_MAX_CUT; ++c) lp.cut[c] = l->cutpoints[c];
  return lp;
}

int launch_lik_raw(const pgp_likelihood_t* lik, const double* f, const double* y, int64_t N,
                   double* ll, double* d1, double* d2, hipStream_t s) {
  if (N <= 0) return 0;
  LikRaw lp = likraw_from_api(lik);
  dim3 grid((unsigned)((N + NTHREADS - 1) / NTHREADS));
  hipLaunchKernelGGL(k_lik_raw, grid, dim3(NTHREADS), 0, s, lp, f, y, N, ll, d1, d2);
  return 0;
}

// ---- likelihood-parameter gradient terms ----
// Per point: term_p = -L_p - 1/2 Sigma_ii L_ffp + (K shat)_i L_fp, summed over
// points into out[0] (noise_std) and out[1 + c] (cutpoint c: lower end of
// class y when c == y, upper end when c == y + 1).  One workgroup per problem;
// fixed-order reduction.  Sigma_ii is in V_X1, (K shat) in V_X2.
__global__ __launch_bounds__(NTHREADS) void k_lik_grad_terms(Prob* const* probs, double* out) {
  const Prob& P = *probs[blockIdx.x];
  const LikParams lp = likparams_from_prob(P);
  constexpr int NS = 1 + PGP_MAX_CUT;
  __shared__ double acc[NS][NTHREADS];
  for (int c = 0; c < NS; ++c) acc[c][threadIdx.x] = 0.0;
  const double* F = vslot(P, V_F);
  const double* Sig = vslot(P, V_X1);
  const double* Ks = vslot(P, V_X2);
  for (int i = threadIdx.x; i < P.N; i += NTHREADS) {
    LikDer o;
    lik_der(lp, F[i], P.y[i], 3, o);
    const double sg = Sig[i], ks = Ks[i];
    acc[0][threadIdx.x] += -o.ps[0] - 0.5 * sg * o.ps[2] + ks * o.ps[1];
    if (lp.kind != PGP_LIK_GAUSSIAN) {
      int yi = (int)P.y[i];
      yi = yi < 0 ? 0 : (yi >= lp.J ? lp.J - 1 : yi);
      acc[1 + yi][threadIdx.x] += -o.plo[0] - 0.5 * sg * o.plo[2] + ks * o.plo[1];
      acc[2 + yi][threadIdx.x] += -o.phi[0] - 0.5 * sg * o.phi[2] + ks * o.phi[1];
    }
  }
  __syncthreads();
  for (int c = 0; c < NS; ++c) {
    for (int st = NTHREADS / 2; st > 0; st >>= 1) {
      if ((int)threadIdx.x < st) acc[c][threadIdx.x] += acc[c][threadIdx.x + st];
      __syncthreads();
    }
  }
  if (threadIdx.x < NS) {
    double v = acc[threadIdx.x][0];
    // infinite cutpoints carry no gradient
    if (threadIdx.x >= 1 && isinf(P.cut[threadIdx.x - 1])) v = 0.0;
    out[(size_t)blockIdx.x * NS + threadIdx.x] = v;
  }
}

void launch_lik_grad_terms(const Batch& b, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_lik_grad_terms, dim3(b.count), dim3(NTHREADS), 0, s, b.dprobs, out);
}


// ---- predictive class probabilities (probit/utilities.py:232-249) ----
// out[i * J + j] = Phi((b_{j+1} - mu_i) / s_i) - Phi((b_j - mu_i) / s_i),
// s_i = sqrt(var_i + sigma^2), with the reference's +-inf guards.
__global__ __launch_bounds__(NTHREADS) void k_probit_predictive(LikRaw lr, const double* mean,
                                                                const double* var, int64_t M,
                                                                double* out) {
  const int64_t e = (int64_t)blockIdx.x * NTHREADS + threadIdx.x;
  if (e >= M * lr.J) return;
  const int64_t i = e / lr.J;
  const int j = (int)(e - i * lr.J);
  const double s = sqrt(var[i] + lr.sigma * lr.sigma);
  const double lo = lr.cut[j], hi = lr.cut[j + 1];
  const double c1 = (isinf(lo) && lo < 0) ? 0.0 : ndtr_ref((lo - mean[i]) / s);
  const double c2 = (isinf(hi) && hi > 0) ? 1.0 : ndtr_ref((hi - mean[i]) / s);
  out[e] = c2 - c1;
}

int launch_probit_predictive(const pgp_likelihood_t* lik, const double* mean, const double* var,
                             int64_t M, double* out, hipStream_t s) {
  if (M <= 0) return 0;
  LikRaw lr = likraw_from_api(lik);
  const int64_t n = M * lik->J;
  hipLaunchKernelGGL(k_probit_predictive, dim3((unsigned)((n + NTHREADS - 1) / NTHREADS)),
                     dim3(NTHREADS), 0, s, lr, mean, var, M, out);
  return 0;
}

// Ordinal test metrics (examples/classification.py:325-341) from class
// probabilities P (M x J, row-major, as k_probit_predictive writes them) and
// test labels y: y_pred_i = argmax_j P_ij (the first maximum, as argmax),
// and -- the reference indexes P[:, y_test] -- the M x M matrix
// L_ij = P(i, y_j), whose log-sum is its log predictive probability.  One
// workgroup per row i: its part of that sum (fixed-order reduction), the
// prediction, |y_pred_i - y_i| and [y_pred_i != y_i] into the row workspace.
// A label outside 0..J-1 reads no memory and gives NaN.
__global__ __launch_bounds__(NTHREADS) void k_metrics_rows(const double* P, int J,
                                                           const double* y, int64_t M,
                                                           double* plik, int32_t* y_pred,
                                                           double* rows) {
  __shared__ double red[NTHREADS];
  const int64_t i = blockIdx.x;
  const double* Pi = P + i * J;
  double acc = 0.0;
  for (int64_t j = threadIdx.x; j < M; j += NTHREADS) {
    const int c = (int)y[j];
    const double v = (c >= 0 && c < J && (double)c == y[j]) ? Pi[c] : NAN;
    if (plik) plik[i * M + j] = v;
    acc += log(v);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = NTHREADS / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    int best = 0;
    for (int c = 1; c < J; ++c)
      if (Pi[c] > Pi[best]) best = c;
    y_pred[i] = best;
    rows[i] = red[0];
    rows[M + i] = fabs((double)best - y[i]);
    rows[2 * M + i] = ((double)best != y[i]) ? 1.0 : 0.0;
  }
}

// Column sums of the row workspace in a fixed order (one workgroup):
// metrics = [incorrect, correct, mean absolute error, log predictive
// probability, mean zero-one error].
__global__ __launch_bounds__(NTHREADS) void k_metrics_final(const double* rows, int64_t M,
                                                            double* metrics) {
  __shared__ double red[3][NTHREADS];
  double a[3] = {0.0, 0.0, 0.0};
  for (int64_t i = threadIdx.x; i < M; i += NTHREADS)
    for (int q = 0; q < 3; ++q) a[q] += rows[q * M + i];
  for (int q = 0; q < 3; ++q) red[q][threadIdx.x] = a[q];
  __syncthreads();
  for (int h = NTHREADS / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int q = 0; q < 3; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double n = (double)M;
    metrics[0] = red[2][0];
    metrics[1] = n - red[2][0];
    metrics[2] = red[1][0] / n;
    metrics[3] = red[0][0];
    metrics[4] = red[2][0] / n;
  }
}

int launch_ordinal_metrics(const double* P, int J, const double* y, int64_t M, int32_t* y_pred,
                           double* plik, double* metrics, double* rows, hipStream_t s) {
  hipLaunchKernelGGL(k_metrics_rows, dim3((unsigned)M), dim3(NTHREADS), 0, s, P, J, y, M, plik,
                     y_pred, rows);
  hipLaunchKernelGGL(k_metrics_final, dim3(1), dim3(NTHREADS), 0, s, rows, M, metrics);
  return 0;
}

}  // namespace pgp
