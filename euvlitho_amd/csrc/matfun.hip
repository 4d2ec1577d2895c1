// Matrix functions of the absorber-layer operator D built from the MFMA GEMM:
//   C  = cos(d sqrt(D)),  F = sin(d sqrt(D)) / sqrt(D),  G2 = D F = sqrt(D) sin(d sqrt(D)).
// All three are even in sqrt(D), i.e. power series in D: no square root, no
// branch choice and no eigendecomposition.  They are evaluated by a degree-12
// Taylor series in x = h^2 D (h = d / 2^s, Paterson-Stockmeyer in blocks of
// x^4, or of x^6 where the products by D run as FFT correlations) followed by
// s doublings C <- 2 C^2 - I, F <- 2 F C; s is set by the norms of D^3 and
// D^4 rather than of D.
#include <cmath>

#include "ew.h"

namespace {
const double2 ZERO = {0.0, 0.0};
inline double2 re(double x) { return make_double2(x, 0.0); }
}  // namespace

namespace euvl {

// D, C, F, G2: batch x n x n.  Takes 6 workspace matrices.  With xd, the
// products with D as a factor (D^2 .. D^6 and G2) are FFT correlations over
// the order grid (xd_fft), and the series step and the doublings stay MFMA
// GEMMs.
double cos_sinc_theta(euvl_handle h) {
  // The degree-12 remainder is bounded by (h^2 alpha)^13 / 26!: h^2 alpha <= 5.9
  // keeps it below 2.6e-17 (FP64 products); with FP32 products (the precision
  // form) it only has to stay below FP32 rounding; there the rounding of the
  // series grows with its terms (up to ~cosh(sqrt(h^2 alpha))), so h^2 alpha is
  // kept at 8 (one doubling more than a bound of 24, images 1.2e-4 of peak from
  // FP64 instead of 5.8e-4 at the 1024 nm clip).
  PrecScope scope(h, EUVL_PREC_FP32_MATFUN);
  return (h->prec_mask & h->prec_scope) ? 8.0 : 5.9;
}

int cos_sinc(euvl_handle h, int n, int batch, const double2* D, double d, double2* C, double2* F, double2* G2,
             int* doublings, int fixed_s, double* d_red, int* d_flags, double* alpha_out, const XdOp* xd) {
  const double theta = cos_sinc_theta(h);
  PrecScope scope(h, EUVL_PREC_FP32_MATFUN);
  const size_t mat = sizeof(double2) * (size_t)n * n * batch;
  // six scratch matrices in one block, so that neighbours are contiguous
  // (pairs of them are the operands of the paired launches below)
  double2* s0 = (double2*)ws_take(h, 6 * mat);
  if (!s0) {
    set_error("cos_sinc: workspace too small");
    return EUVL_ERR_NOMEM;
  }
  const int64_t sm = (int64_t)n * n;
  double2* s[6];
  for (int i = 0; i < 6; ++i) s[i] = s0 + (int64_t)i * sm * batch;
  // Unscaled powers P2 = D^2, P3 = D^3, P4 = D^4 (the Paterson-Stockmeyer
  // blocks).  For every j >= 6, ||D^j||^(1/j) <= alpha = max(||D^3||^(1/3),
  // ||D^4||^(1/4)) (Al-Mohy and Higham, SIAM J. Matrix Anal. Appl. 31 (2009),
  // Lemma 4.1), which is often well below ||D||: s is the smallest with
  // h^2 alpha <= theta, so that the remainder of both series (terms j >= 13) is
  // below rounding.
  //
  // With xd (C and F contiguous) the products by D are cheap FFT correlations,
  // so the blocks are taken in powers of D^6 instead: D^2 .. D^6 by xd_fft
  // (D^(j+1) = D^j D) and q = O + D^6 T, one paired GEMM for both series in
  // place of the two below.
  if (xd && F != C + sm * batch) xd = nullptr;
  const double2* X1 = D;
  double2 *X2 = s[1], *X3 = s[2], *H = s[3], *Q = s[4], *X4 = s[5];
  if (xd) {
    X4 = s[3];
    const double2* in = D;
    for (int j = 1; j <= 5; ++j) {  // s[j] = D^(j+1)
      EUVL_TRY(xd_fft(h, *xd, batch, in, s[j]));
      in = s[j];
    }
  } else {
    EUVL_TRY(zgemm(h, n, n, n, re(1.0), D, n, sm, D, n, sm, ZERO, X2, n, sm, ZERO, batch));
    EUVL_TRY(zgemm(h, n, n, n, re(1.0), X2, n, sm, D, n, sm, ZERO, X3, n, sm, ZERO, batch));
    EUVL_TRY(zgemm(h, n, n, n, re(1.0), X2, n, sm, X2, n, sm, ZERO, X4, n, sm, ZERO, batch));
  }
  int sq = 0;
  if (fixed_s >= 0) {
    // the bound is checked on the device; no read-back
    sq = fixed_s;
    EUVL_TRY(reduce(h, 1, n, batch, X3, d_red));
    EUVL_TRY(reduce(h, 1, n, batch, X4, d_red + batch));
    const double hh = std::ldexp(d, -sq);
    EUVL_TRY(flag_check(h, 0, batch, d_red, d_red + batch, nullptr, hh * hh, theta, FLAG_DOUBLINGS, d_flags));
  } else {
    if (batch > 256) {
      set_error("cos_sinc: batch %d above the 256 of the host read-back", batch);
      return EUVL_ERR_ARG;
    }
    EUVL_TRY(reduce(h, 1, n, batch, X3, h->d_red));
    EUVL_TRY(reduce(h, 1, n, batch, X4, h->d_red + 256));
    EUVL_HIP_CHECK(hipMemcpyAsync(h->h_pinned, h->d_red, sizeof(double) * 512, hipMemcpyDeviceToHost, h->stream));
    EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));
    double alpha = 0.0;
    for (int b = 0; b < batch; ++b) {
      const double a3 = std::cbrt(h->h_pinned[b]), a4 = std::sqrt(std::sqrt(h->h_pinned[256 + b]));
      alpha = std::max(alpha, (std::isnan(a3) || std::isnan(a4)) ? INFINITY : std::max(a3, a4));
    }
    if (!std::isfinite(alpha)) {
      set_error("cos_sinc: non-finite ||D^3|| or ||D^4||");
      return EUVL_ERR_NOT_CONVERGED;
    }
    while (std::ldexp(d, -sq) * std::ldexp(d, -sq) * alpha > theta && sq < 60) ++sq;
    if (alpha_out) *alpha_out = alpha;
  }
  if (doublings) *doublings = sq;
  const double hh = std::ldexp(d, -sq);
  // q(x) = sum c_j x^j with x = h^2 D is evaluated as sum (c_j h^(2j)) D^j on the unscaled powers
  double cc[13], cf[13];
  double fact = 1.0;  // m!
  for (int m = 0; m < 26; ++m) {
    if (m > 0) fact *= m;
    const double sgn = ((m / 2) % 2) ? -1.0 : 1.0;
    const double scale = std::pow(hh * hh, m / 2);
    if (m % 2 == 0)
      cc[m / 2] = sgn * scale / fact;
    else
      cf[m / 2] = sgn * hh * scale / fact;
  }
  if (xd) {
    // O (into C, F) and T (over D^2, D^3: contiguous, the right operands of
    // the paired launch), then [C; F] += D^6 [Tc; Tf]
    const double2* P[6] = {D, s[1], s[2], s[3], s[4], s[5]};
    EUVL_TRY(ew_ps6_blocks(h, n, batch, P, cc, cf, C, s[1], F, s[2]));
    EUVL_TRY(zgemm_periodic(h, n, n, n, re(1.0), s[5], n, sm, s[1], n, sm, re(1.0), C, n, sm, ZERO, 2 * batch,
                            batch, 0));
    // doublings as below
    double2 *c = C, *t = s[0];
    for (int q = 0; q < sq; ++q) {
      EUVL_TRY(zgemm_periodic(h, n, n, n, re(2.0), c, n, sm, c, n, sm, ZERO, t, n, sm, ZERO, 2 * batch, 0, batch));
      EUVL_TRY(ew_diag_add(h, n, batch, t, re(-1.0)));
      std::swap(c, t);
    }
    if (c != C) EUVL_TRY(ew_copy(h, 2 * sm * batch, C, c));
  } else if (F == C + sm * batch) {
    // q = Q0 + P4 (Q1 + P4 Q2), Qi = c_4i I + c_4i+1 D + c_4i+2 P2 + c_4i+3 P3 (+ c12 P4 in Q2).
    // The blocks of both polynomials come from one pass over D..P4 (the sinc
    // Q1 over P2, which is not read after it); then two products each.
    // C and F contiguous: the cos and sinc products of a step share their
    // left operand (P4, or C in a doubling) and run as one launch of 2 batch
    // entries (the last batch ones read the same operand again); blocks
    // Hc, Hf in s[3], s[4] and Qc, Qf in s[0], s[1] (= P2), also contiguous
    EUVL_TRY(ew_ps_blocks(h, n, batch, X1, X2, X3, X4, cc, cf, s[3], s[0], C, s[4], s[1], F));
    EUVL_TRY(zgemm_periodic(h, n, n, n, re(1.0), X4, n, sm, s[3], n, sm, re(1.0), s[0], n, sm, ZERO, 2 * batch,
                            batch, 0));
    EUVL_TRY(zgemm_periodic(h, n, n, n, re(1.0), X4, n, sm, s[0], n, sm, re(1.0), C, n, sm, ZERO, 2 * batch,
                            batch, 0));
    // doublings: [C; F] <- 2 [C; F] C, then C -= I  (angle h -> 2h)
    double2 *c = C, *t = s[0];
    for (int q = 0; q < sq; ++q) {
      EUVL_TRY(zgemm_periodic(h, n, n, n, re(2.0), c, n, sm, c, n, sm, ZERO, t, n, sm, ZERO, 2 * batch, 0, batch));
      EUVL_TRY(ew_diag_add(h, n, batch, t, re(-1.0)));
      std::swap(c, t);
    }
    if (c != C) EUVL_TRY(ew_copy(h, 2 * sm * batch, C, c));
  } else {
    double2* Hf = s[0];
    double2* Qf = X2;
    EUVL_TRY(ew_ps_blocks(h, n, batch, X1, X2, X3, X4, cc, cf, H, Q, C, Hf, Qf, F));
    EUVL_TRY(zgemm(h, n, n, n, re(1.0), X4, n, sm, H, n, sm, re(1.0), Q, n, sm, ZERO, batch));
    EUVL_TRY(zgemm(h, n, n, n, re(1.0), X4, n, sm, Q, n, sm, re(1.0), C, n, sm, ZERO, batch));
    EUVL_TRY(zgemm(h, n, n, n, re(1.0), X4, n, sm, Hf, n, sm, re(1.0), Qf, n, sm, ZERO, batch));
    EUVL_TRY(zgemm(h, n, n, n, re(1.0), X4, n, sm, Qf, n, sm, re(1.0), F, n, sm, ZERO, batch));
    // doublings: F <- 2 F C, C <- 2 C^2 - I  (angle h -> 2h)
    double2 *c = C, *f = F, *tc = s[0], *tf = X2;
    for (int q = 0; q < sq; ++q) {
      EUVL_TRY(zgemm(h, n, n, n, re(2.0), f, n, sm, c, n, sm, ZERO, tf, n, sm, ZERO, batch));
      EUVL_TRY(zgemm(h, n, n, n, re(2.0), c, n, sm, c, n, sm, ZERO, tc, n, sm, re(-1.0), batch));
      std::swap(c, tc);
      std::swap(f, tf);
    }
    if (c != C) EUVL_TRY(ew_copy(h, sm * batch, C, c));
    if (f != F) EUVL_TRY(ew_copy(h, sm * batch, F, f));
  }
  if (G2) {
    if (xd)  // F is a power series in D, so D F = F D
      EUVL_TRY(xd_fft(h, *xd, batch, F, G2));
    else
      EUVL_TRY(zgemm(h, n, n, n, re(1.0), D, n, sm, F, n, sm, ZERO, G2, n, sm, ZERO, batch));
  }
  return EUVL_OK;
}

}  // namespace euvl

extern "C" int euvl_zcos_sinc_batched(euvl_handle h, int n, const double* d_D, double thickness, double* d_C,
                                      double* d_F, double* d_G2, int batch, int* h_doublings) {
  if (!h || n <= 0 || batch <= 0 || batch > 256 || !d_D || !d_C || !d_F || !(thickness >= 0.0)) {
    euvl::set_error("euvl_zcos_sinc_batched: bad argument");
    return EUVL_ERR_ARG;
  }
  EUVL_TRY(euvl::ws_reserve(h, 6 * sizeof(double2) * (size_t)n * n * batch + 4096));
  return euvl::cos_sinc(h, n, batch, (const double2*)d_D, thickness, (double2*)d_C, (double2*)d_F,
                        (double2*)d_G2, h_doublings);
}
