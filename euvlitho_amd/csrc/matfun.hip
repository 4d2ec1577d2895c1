// Matrix functions built from the MFMA GEMM: principal square root by the
// coupled Newton-Schulz iteration and the exp(A), exp(-A) pair by Taylor
// series (Paterson-Stockmeyer in A^2) with scaling and squaring.
#include <cmath>

#include "ew.h"

namespace {
const double2 ONE = {1.0, 0.0}, ZERO = {0.0, 0.0};
inline double2 re(double x) { return make_double2(x, 0.0); }
}  // namespace

namespace euvl {

// Scratch: the caller passes 3 extra n x n x batch buffers in S/Sinv's
// allocation block via the workspace (taken here).
int sqrtm_ns(euvl_handle h, int n, int batch, const double2* D, double2* S, double2* Sinv, double tol,
             int max_iter, int* iters, double* d_c) {
  const size_t mat = sizeof(double2) * (size_t)n * n * batch;
  double2* buf[5] = {S, Sinv, (double2*)ws_take(h, mat), (double2*)ws_take(h, mat), (double2*)ws_take(h, mat)};
  if (!buf[2] || !buf[3] || !buf[4]) {
    set_error("sqrtm_ns: workspace too small");
    return EUVL_ERR_NOMEM;
  }
  if (batch > 128) {
    set_error("sqrtm_ns: batch > 128");
    return EUVL_ERR_ARG;
  }
  // c_b = mean Re diag(D_b): the iteration runs on D/c, whose spectrum is near 1
  std::vector<double> c(batch);
  EUVL_TRY(reduce_to_host(h, 2, n, batch, D, c.data()));
  for (int b = 0; b < batch; ++b)
    if (!(c[b] > 0.0)) {
      set_error("sqrtm_ns: mean Re diag(D) = %g is not positive (evanescent-dominated spectrum)", c[b]);
      return EUVL_ERR_NOT_CONVERGED;
    }
  // per-batch scalars live in the handle's reduction area: [0,batch) 1/c, [batch,2batch) sqrt(c),
  // [2batch,3batch) 1/sqrt(c), complex
  double2* hs = (double2*)h->h_pinned;
  for (int b = 0; b < batch; ++b) {
    hs[b] = re(1.0 / c[b]);
    hs[batch + b] = re(std::sqrt(c[b]));
    hs[2 * batch + b] = re(1.0 / std::sqrt(c[b]));
  }
  double2* ds = (double2*)h->d_red;
  EUVL_HIP_CHECK(hipMemcpyAsync(ds, hs, sizeof(double2) * 3 * batch, hipMemcpyHostToDevice, h->stream));
  if (d_c) EUVL_HIP_CHECK(hipMemcpyAsync(d_c, ds + batch, sizeof(double2) * batch, hipMemcpyDeviceToDevice, h->stream));

  const int64_t sm = (int64_t)n * n;
  double2 *Y = buf[2], *Z = nullptr, *T = buf[3], *W = buf[4], *V = buf[0];
  EUVL_TRY(ew_bscale(h, n, batch, Y, D, ds, ZERO));  // Y0 = D / c
  double res = INFINITY;
  bool converged = false;
  int it = 0;
  // first step (Z0 = I): T0 = 1.5 I - 0.5 Y0, Y1 = Y0 T0, Z1 = T0
  EUVL_TRY(ew_lincomb(h, n, batch, T, re(-0.5), Y, ZERO, nullptr, ZERO, nullptr, re(1.5)));
  EUVL_TRY(zgemm(h, n, n, n, ONE, Y, n, sm, T, n, sm, ZERO, W, n, sm, ZERO, batch));
  std::swap(Y, W);          // Y = Y1 (buf4), W = buf2 free
  Z = T;                    // Z = T0 (buf3)
  T = W;                    // buf2 free for T
  W = buf[1];               // Sinv block as scratch
  V = buf[0];               // S block as scratch
  it = 1;
  for (; it < max_iter; ++it) {
    // T = 1.5 I - 0.5 Z Y
    EUVL_TRY(zgemm(h, n, n, n, re(-0.5), Z, n, sm, Y, n, sm, ZERO, T, n, sm, re(1.5), batch));
    std::vector<double> r(batch);
    EUVL_TRY(reduce_to_host(h, 0, n, batch, T, r.data()));
    res = 0.0;
    for (double x : r) res = std::max(res, std::isnan(x) ? INFINITY : x);
    if (!(res < 4.0)) break;  // diverging
    EUVL_TRY(zgemm(h, n, n, n, ONE, Y, n, sm, T, n, sm, ZERO, W, n, sm, ZERO, batch));
    EUVL_TRY(zgemm(h, n, n, n, ONE, T, n, sm, Z, n, sm, ZERO, V, n, sm, ZERO, batch));
    double2* oy = Y;
    double2* oz = Z;
    Y = W;
    Z = V;
    W = oy;
    V = oz;
    // quadratic convergence: the update made from a residual below tol leaves
    // an error of order tol^2, so stop after it
    if (res < tol) {
      converged = true;
      ++it;
      break;
    }
  }
  if (iters) *iters = it;
  if (!converged) {
    set_error("sqrtm_ns: Newton-Schulz residual %.3e after %d iterations (tol %.1e)", res, it, tol);
    return EUVL_ERR_NOT_CONVERGED;
  }
  // S = sqrt(c) Y, Sinv = Z / sqrt(c).  Y and Z may sit in S/Sinv's own blocks;
  // go through a free buffer where needed.
  double2* freeb = T;
  if (Y == S) {
    EUVL_TRY(ew_bscale(h, n, batch, S, S, ds + batch, ZERO));
  } else if (Y == Sinv) {
    // Y occupies Sinv's block: move it to S first (S is free unless Z lives there)
    if (Z == S) {
      EUVL_TRY(ew_copy(h, sm * batch, freeb, Z));
      Z = freeb;
    }
    EUVL_TRY(ew_bscale(h, n, batch, S, Y, ds + batch, ZERO));
  } else {
    if (Z == S) {
      EUVL_TRY(ew_copy(h, sm * batch, freeb, Z));
      Z = freeb;
    }
    EUVL_TRY(ew_bscale(h, n, batch, S, Y, ds + batch, ZERO));
  }
  EUVL_TRY(ew_bscale(h, n, batch, Sinv, Z, ds + 2 * batch, ZERO));
  return EUVL_OK;
}

// exp(A) and exp(-A) for batch x n x n matrices A.  Uses 7 workspace matrices.
int expm_pair(euvl_handle h, int n, int batch, const double2* A, double2* E, double2* Einv, int* squarings) {
  const size_t mat = sizeof(double2) * (size_t)n * n * batch;
  double2* s[7];
  for (int i = 0; i < 7; ++i) {
    s[i] = (double2*)ws_take(h, mat);
    if (!s[i]) {
      set_error("expm_pair: workspace too small");
      return EUVL_ERR_NOMEM;
    }
  }
  const int64_t sm = (int64_t)n * n;
  double2 *B1 = s[0], *B2 = s[1], *B3 = s[2], *H = s[3], *C1 = s[4], *Ev = s[5], *P = s[6];
  EUVL_TRY(zgemm(h, n, n, n, ONE, A, n, sm, A, n, sm, ZERO, B1, n, sm, ZERO, batch));    // A^2
  EUVL_TRY(zgemm(h, n, n, n, ONE, B1, n, sm, B1, n, sm, ZERO, B2, n, sm, ZERO, batch));  // A^4
  EUVL_TRY(zgemm(h, n, n, n, ONE, B2, n, sm, B1, n, sm, ZERO, B3, n, sm, ZERO, batch));  // A^6
  std::vector<double> n4(batch), n6(batch);
  EUVL_TRY(reduce_to_host(h, 1, n, batch, B2, n4.data()));
  EUVL_TRY(reduce_to_host(h, 1, n, batch, B3, n6.data()));
  double alpha = 0.0;
  for (int b = 0; b < batch; ++b) {
    alpha = std::max(alpha, std::pow(n4[b], 0.25));
    alpha = std::max(alpha, std::pow(n6[b], 1.0 / 6.0));
  }
  if (!std::isfinite(alpha)) {
    set_error("expm_pair: non-finite power norm");
    return EUVL_ERR_NOT_CONVERGED;
  }
  const double theta = 0.7;  // degree-19 Taylor remainder < 1e-21 at this radius
  int sq = 0;
  if (alpha > theta) sq = (int)std::ceil(std::log2(alpha / theta));
  if (sq > 40) {
    set_error("expm_pair: norm %.3e needs %d squarings", alpha, sq);
    return EUVL_ERR_NOT_CONVERGED;
  }
  if (squarings) *squarings = sq;
  const double f = std::ldexp(1.0, -sq);
  EUVL_TRY(ew_lincomb(h, n, batch, B1, re(f * f), B1, ZERO, nullptr, ZERO, nullptr, ZERO));
  EUVL_TRY(ew_lincomb(h, n, batch, B2, re(std::pow(f, 4)), B2, ZERO, nullptr, ZERO, nullptr, ZERO));
  EUVL_TRY(ew_lincomb(h, n, batch, B3, re(std::pow(f, 6)), B3, ZERO, nullptr, ZERO, nullptr, ZERO));
  double ce[10], co[10];
  double fact = 1.0;  // j!
  for (int j = 0; j < 20; ++j) {
    if (j > 0) fact *= j;
    if (j % 2 == 0)
      ce[j / 2] = 1.0 / fact;
    else
      co[j / 2] = 1.0 / fact;
  }
  // polynomial q(B) = sum_{i<10} c_i B^i = Q0 + B3 (Q1 + B3 (Q2 + c9 B3)), B = B1, B^2 = B2, B^3 = B3
  auto poly = [&](const double* cf, double2* out) -> int {
    EUVL_TRY(ew_lincomb(h, n, batch, H, re(cf[9]), B3, re(cf[8]), B2, re(cf[7]), B1, re(cf[6])));
    EUVL_TRY(ew_lincomb(h, n, batch, C1, re(cf[5]), B2, re(cf[4]), B1, ZERO, nullptr, re(cf[3])));
    EUVL_TRY(zgemm(h, n, n, n, ONE, B3, n, sm, H, n, sm, ONE, C1, n, sm, ZERO, batch));
    EUVL_TRY(ew_lincomb(h, n, batch, out, re(cf[2]), B2, re(cf[1]), B1, ZERO, nullptr, re(cf[0])));
    EUVL_TRY(zgemm(h, n, n, n, ONE, B3, n, sm, C1, n, sm, ONE, out, n, sm, ZERO, batch));
    return EUVL_OK;
  };
  EUVL_TRY(poly(ce, Ev));
  EUVL_TRY(poly(co, P));
  // odd part: (A / 2^sq) * P  -> H
  EUVL_TRY(zgemm(h, n, n, n, re(f), A, n, sm, P, n, sm, ZERO, H, n, sm, ZERO, batch));
  EUVL_TRY(ew_lincomb(h, n, batch, E, ONE, Ev, ONE, H, ZERO, nullptr, ZERO));
  EUVL_TRY(ew_lincomb(h, n, batch, Einv, ONE, Ev, re(-1.0), H, ZERO, nullptr, ZERO));
  // squaring
  double2 *e = E, *ei = Einv, *t1 = B1, *t2 = B2;
  for (int q = 0; q < sq; ++q) {
    EUVL_TRY(zgemm(h, n, n, n, ONE, e, n, sm, e, n, sm, ZERO, t1, n, sm, ZERO, batch));
    EUVL_TRY(zgemm(h, n, n, n, ONE, ei, n, sm, ei, n, sm, ZERO, t2, n, sm, ZERO, batch));
    std::swap(e, t1);
    std::swap(ei, t2);
  }
  if (e != E) EUVL_TRY(ew_copy(h, sm * batch, E, e));
  if (ei != Einv) EUVL_TRY(ew_copy(h, sm * batch, Einv, ei));
  return EUVL_OK;
}

}  // namespace euvl

extern "C" int euvl_zsqrtm_batched(euvl_handle h, int n, const double* d_D, double* d_S, double* d_Sinv, int batch,
                                   double tol, int max_iter, int* h_iters) {
  if (!h || n <= 0 || batch <= 0 || !d_D || !d_S || !d_Sinv || max_iter < 1) {
    euvl::set_error("euvl_zsqrtm_batched: bad argument");
    return EUVL_ERR_ARG;
  }
  EUVL_TRY(euvl::ws_reserve(h, 3 * sizeof(double2) * (size_t)n * n * batch + 4096));
  return euvl::sqrtm_ns(h, n, batch, (const double2*)d_D, (double2*)d_S, (double2*)d_Sinv, tol, max_iter, h_iters,
                        nullptr);
}

namespace {
__global__ void shift_scale_kernel(int n, int64_t per, int batch, double2* out, const double2* S, const double* th,
                                   double shift) {
  const int64_t total = per * batch;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / per);
    const int64_t r = e - (int64_t)b * per;
    double2 v = S[e];
    if (r / n == r % n) v.x -= shift;
    out[e] = make_double2(-th[b] * v.y, th[b] * v.x);  // i * d * (S - shift I)
  }
}
}  // namespace

extern "C" int euvl_zexpm_pair_batched(euvl_handle h, int n, const double* d_S, double shift_re,
                                       const double* h_thickness, double* d_E, double* d_Einv, int batch,
                                       int* h_squarings) {
  if (!h || n <= 0 || batch <= 0 || batch > 256 || !d_S || !h_thickness || !d_E || !d_Einv) {
    euvl::set_error("euvl_zexpm_pair_batched: bad argument");
    return EUVL_ERR_ARG;
  }
  const size_t mat = sizeof(double2) * (size_t)n * n * batch;
  EUVL_TRY(euvl::ws_reserve(h, 8 * mat + 4096));
  double2* A = (double2*)euvl::ws_take(h, mat);
  double* dth = (double*)h->d_red;
  EUVL_HIP_CHECK(hipMemcpyAsync(dth, h_thickness, sizeof(double) * batch, hipMemcpyHostToDevice, h->stream));
  const int64_t per = (int64_t)n * n;
  int64_t blocks = (per * batch + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(shift_scale_kernel, dim3((unsigned)blocks), dim3(256), 0, h->stream, n, per, batch, A,
                     (const double2*)d_S, dth, shift_re);
  EUVL_LAUNCH_CHECK();
  return euvl::expm_pair(h, n, batch, A, (double2*)d_E, (double2*)d_Einv, h_squarings);
}
