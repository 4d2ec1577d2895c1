// Matrix functions of the absorber-layer operator D built from the MFMA GEMM:
//   C  = cos(d sqrt(D)),  F = sin(d sqrt(D)) / sqrt(D),  G2 = D F = sqrt(D) sin(d sqrt(D)).
// All three are even in sqrt(D), i.e. power series in D: no square root, no
// branch choice and no eigendecomposition.  They are evaluated by a degree-9
// Taylor series in x = h^2 D (h = d / 2^s, Paterson-Stockmeyer in blocks of
// x^3) followed by s doublings C <- 2 C^2 - I, F <- 2 F C.
#include <cmath>

#include "ew.h"

namespace {
const double2 ZERO = {0.0, 0.0};
inline double2 re(double x) { return make_double2(x, 0.0); }
}  // namespace

namespace euvl {

// D, C, F, G2: batch x n x n.  Takes 6 workspace matrices.
int cos_sinc(euvl_handle h, int n, int batch, const double2* D, double d, double2* C, double2* F, double2* G2,
             int* doublings) {
  PrecScope scope(h, EUVL_PREC_FP32_MATFUN);
  const size_t mat = sizeof(double2) * (size_t)n * n * batch;
  double2* s[6];
  for (int i = 0; i < 6; ++i) {
    s[i] = (double2*)ws_take(h, mat);
    if (!s[i]) {
      set_error("cos_sinc: workspace too small");
      return EUVL_ERR_NOMEM;
    }
  }
  const int64_t sm = (int64_t)n * n;
  // ||D||_inf bounds every ||D^j||^(1/j); choose s so that ||x||_inf <= 1.5,
  // where the degree-9 remainder ||x||^10/20! is below 3e-17.
  std::vector<double> nr(batch);
  EUVL_TRY(reduce_to_host(h, 1, n, batch, D, nr.data()));
  double nd = 0.0;
  for (double v : nr) nd = std::max(nd, std::isnan(v) ? INFINITY : v);
  if (!std::isfinite(nd)) {
    set_error("cos_sinc: non-finite ||D||");
    return EUVL_ERR_NOT_CONVERGED;
  }
  int sq = 0;
  while (std::ldexp(d, -sq) * std::ldexp(d, -sq) * nd > 1.5 && sq < 60) ++sq;
  if (doublings) *doublings = sq;
  const double hh = std::ldexp(d, -sq);
  double2 *X1 = s[0], *X2 = s[1], *X3 = s[2], *H = s[3], *Q = s[4], *T = s[5];
  EUVL_TRY(ew_lincomb(h, n, batch, X1, re(hh * hh), D, ZERO, nullptr, ZERO, nullptr, ZERO));
  EUVL_TRY(zgemm(h, n, n, n, re(1.0), X1, n, sm, X1, n, sm, ZERO, X2, n, sm, ZERO, batch));
  EUVL_TRY(zgemm(h, n, n, n, re(1.0), X2, n, sm, X1, n, sm, ZERO, X3, n, sm, ZERO, batch));
  double cc[10], cf[10];
  double fact = 1.0;  // m!
  for (int m = 0; m < 20; ++m) {
    if (m > 0) fact *= m;
    const double sgn = ((m / 2) % 2) ? -1.0 : 1.0;
    if (m % 2 == 0)
      cc[m / 2] = sgn / fact;
    else
      cf[m / 2] = sgn * hh / fact;
  }
  // q(x) = sum_{j<10} c_j x^j = Q0 + x^3 (Q1 + x^3 (Q2 + c9 x^3))
  auto poly = [&](const double* c, double2* out) -> int {
    EUVL_TRY(ew_lincomb(h, n, batch, H, re(c[9]), X3, re(c[8]), X2, re(c[7]), X1, re(c[6])));
    EUVL_TRY(ew_lincomb(h, n, batch, Q, re(c[5]), X2, re(c[4]), X1, ZERO, nullptr, re(c[3])));
    EUVL_TRY(zgemm(h, n, n, n, re(1.0), X3, n, sm, H, n, sm, re(1.0), Q, n, sm, ZERO, batch));
    EUVL_TRY(ew_lincomb(h, n, batch, out, re(c[2]), X2, re(c[1]), X1, ZERO, nullptr, re(c[0])));
    EUVL_TRY(zgemm(h, n, n, n, re(1.0), X3, n, sm, Q, n, sm, re(1.0), out, n, sm, ZERO, batch));
    return EUVL_OK;
  };
  EUVL_TRY(poly(cc, C));
  EUVL_TRY(poly(cf, F));
  // doublings: F <- 2 F C, C <- 2 C^2 - I  (angle h -> 2h)
  double2 *c = C, *f = F, *tc = X1, *tf = X2;
  for (int q = 0; q < sq; ++q) {
    EUVL_TRY(zgemm(h, n, n, n, re(2.0), f, n, sm, c, n, sm, ZERO, tf, n, sm, ZERO, batch));
    EUVL_TRY(zgemm(h, n, n, n, re(2.0), c, n, sm, c, n, sm, ZERO, tc, n, sm, re(-1.0), batch));
    std::swap(c, tc);
    std::swap(f, tf);
  }
  if (c != C) EUVL_TRY(ew_copy(h, sm * batch, C, c));
  if (f != F) EUVL_TRY(ew_copy(h, sm * batch, F, f));
  (void)T;
  if (G2) EUVL_TRY(zgemm(h, n, n, n, re(1.0), D, n, sm, F, n, sm, ZERO, G2, n, sm, ZERO, batch));
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
