// The coupling product H G of the top solve by 2D FFTs in LDS.
//
// G = diag(kt) sigma^ diag(kt) / k^2, where sigma^(dl, dm) is the Fourier
// coefficient of the inverse permittivity at the order difference (the
// reference's Toeplitz block, header.h:979-1052).  Row i of H G is therefore
//   (H G)[i][j] = kt_j / k^2 sum_q h_i(q) sigma^(l_q - l_j, m_q - m_j),
//   h_i(q) = H[i][q] kt_q,
// a 2D cross-correlation of h_i with sigma^ over the (l, m) grid of the
// orders.  On a periodic Pl x Pm grid with Pl >= 4 lmax + 1 and
// Pm >= 4 mmax + 1 the window |dl| <= 2 lmax, |dm| <= 2 mmax of sigma^ wraps
// onto distinct points, so the circular correlation equals the linear one at
// every order:
//   c = IDFT(DFT(h) S) / (Pl Pm),  S(k) = sum_z sigma^(z) e^{+2 pi i k.z / P}.
// One workgroup per row of H transforms its grid in LDS (about 15 GFLOP and
// no HBM traffic beyond reading H and updating V, against 8 N^3 flops per
// batch entry for the product), and subtracts the row of H G from V.
//
// The 1D transforms are mixed-radix Stockham passes (radices 9, 4, 2 and 3;
// grid sizes with no other prime factor), in place through registers: every
// thread reads the inputs of its butterflies, a barrier, then it writes the
// outputs.  Along m only the 2 lmax + 1 rows that hold orders are transformed
// (the others are zero on the way in and not read on the way out).
#include <algorithm>
#include <atomic>
#include <cstdlib>

#include "ew.h"

namespace {

constexpr int FFT_MAX_PTS = 8192;  // Pl Pm: the grid in LDS (128 KiB)
constexpr int FFT_MAX_P = 128;     // points along one axis
constexpr int FFT_THREADS = 1024;

// The grid and the root tables live in LDS; the out-of-line passes receive
// them as generic pointers and cast them back, so that their accesses are
// LDS instructions with 32-bit addresses rather than flat ones.
// Elements are accessed as a native 16-byte vector (one ds_read_b128 /
// ds_write_b128 each: two 8-byte accesses per element conflicted on the LDS
// banks 9 times as often); the vector class has no constructors for
// another address space.
typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) dvec2 lds_f64;

__device__ __forceinline__ double2 lds_ld(const lds_f64* p, int i) {
  const dvec2 t = p[i];
  return make_double2(t.x, t.y);
}
__device__ __forceinline__ void lds_st(lds_f64* p, int i, double2 v) { p[i] = (dvec2){v.x, v.y}; }

__device__ __forceinline__ double2 root_of(const lds_f64* root, int t, bool inv) {
  const double2 w = lds_ld(root, t);
  return inv ? make_double2(w.x, -w.y) : w;
}

// a / b for 0 <= a < 2^13 and 0 < b <= 2^7 from rb ~ 1/b in FP32 (v_rcp_f32,
// within 1 ulp): (a + 1/2) / b is at least 1/(2b) from an integer, and the
// product's relative error is below 2^-22, i.e. its absolute error below
// 2^13 2^-22 / b < 1/(2b), so the truncation is exact (and far cheaper than
// an integer division)
__device__ __forceinline__ int fdiv(int a, float rb) { return (int)(((float)a + 0.5f) * rb); }
__device__ __forceinline__ float recip(int b) { return __builtin_amdgcn_rcpf((float)b); }

// y = the R-point DFT of x, e^(-+2 pi i r k / R), R = 2, 3, 4, 9, with its
// constant roots (9 as 3 x 3: DFTs of the three stride-3 subsequences,
// twiddles e^(-+2 pi i r1 k1 / 9), DFTs across them)
template <int R>
__device__ __forceinline__ void dft_small(const double2* x, double2* y, bool inv) {
  if constexpr (R == 9) {
    // cos and sin of 2 pi t / 9 for t = 1, 2, 4
    constexpr double c1 = 0.766044443118978, s1 = 0.6427876096865393;
    constexpr double c2 = 0.17364817766693041, s2 = 0.984807753012208;
    constexpr double c4 = -0.9396926207859083, s4 = 0.3420201433256689;
    const double sg = inv ? 1.0 : -1.0;
    double2 a[3][3];
#pragma unroll
    for (int r1 = 0; r1 < 3; ++r1) {
      const double2 in[3] = {x[r1], x[r1 + 3], x[r1 + 6]};
      dft_small<3>(in, a[r1], inv);
    }
    a[1][1] = cmul(a[1][1], make_double2(c1, sg * s1));
    a[1][2] = cmul(a[1][2], make_double2(c2, sg * s2));
    a[2][1] = cmul(a[2][1], make_double2(c2, sg * s2));
    a[2][2] = cmul(a[2][2], make_double2(c4, sg * s4));
#pragma unroll
    for (int k1 = 0; k1 < 3; ++k1) {
      const double2 in[3] = {a[0][k1], a[1][k1], a[2][k1]};
      double2 out[3];
      dft_small<3>(in, out, inv);
#pragma unroll
      for (int k2 = 0; k2 < 3; ++k2) y[k1 + 3 * k2] = out[k2];
    }
  } else if constexpr (R == 2) {
    y[0] = cadd(x[0], x[1]);
    y[1] = csub(x[0], x[1]);
  } else if constexpr (R == 3) {
    constexpr double h = 0.86602540378443864676;  // sin(2 pi / 3)
    const double2 s = cadd(x[1], x[2]), d = csub(x[1], x[2]);
    const double2 t = make_double2(x[0].x - 0.5 * s.x, x[0].y - 0.5 * s.y);
    // -i h d (forward), +i h d (inverse)
    const double2 e = inv ? make_double2(-h * d.y, h * d.x) : make_double2(h * d.y, -h * d.x);
    y[0] = cadd(x[0], s);
    y[1] = cadd(t, e);
    y[2] = csub(t, e);
  } else {
    const double2 a = cadd(x[0], x[2]), b = csub(x[0], x[2]);
    const double2 c = cadd(x[1], x[3]), d = csub(x[1], x[3]);
    // -i d (forward), +i d (inverse)
    const double2 e = inv ? make_double2(-d.y, d.x) : make_double2(d.y, -d.x);
    y[0] = cadd(a, c);
    y[1] = cadd(b, e);
    y[2] = csub(a, c);
    y[3] = csub(b, e);
  }
}

// One Stockham pass of radix R over `nlines` lines of the grid X (Pl x Pm,
// row-major).  Along m a line is a row (the kept rows when keep > 0: rows
// 0 .. keep/2 and Pl - keep/2 .. Pl - 1), along l a column.  Butterfly j of
// a line reads positions j + r P/R, multiplies input r by w^(r (j mod Ns)),
// w = e^(-+2 pi i / (Ns R)), takes the R-point DFT and writes output r to
// (j / Ns) Ns R + (j mod Ns) + r Ns (Ns: the product of the radices before).
// (Out of line: inlined into the four axis loops of a kernel it spills.)
// INV: the inverse direction; FIRST: the first pass of an axis (Ns = 1, every
// twiddle is 1).
template <int R, bool INV, bool FIRST>
__device__ __attribute__((noinline)) void fft_pass(double2* Xg, const double2* rootg, int P, int Ns,
                                                   bool along_m, int nlines, int Pl, int Pm, int keep) {
  constexpr bool inv = INV;
  constexpr int MAXB = (FFT_MAX_PTS / R + FFT_THREADS - 1) / FFT_THREADS;
  lds_f64* X = (lds_f64*)Xg;
  const lds_f64* root = (const lds_f64*)rootg;
  const int per = P / R, total = nlines * per, stride = along_m ? 1 : Pm;
  const int step = P / (Ns * R);
  const float rper = recip(per), rlines = recip(nlines), rns = recip(Ns);
  double2 v[MAXB][R];
  int dst[MAXB];
#pragma unroll
  for (int u = 0; u < MAXB; ++u) {
    const int beta = threadIdx.x + u * FFT_THREADS;
    dst[u] = -1;
    if (beta < total) {
      int line, j;
      if (along_m) {  // consecutive threads along a row
        line = fdiv(beta, rper);
        j = beta - line * per;
      } else {  // consecutive threads across columns
        j = fdiv(beta, rlines);
        line = beta - j * nlines;
      }
      const int base = along_m ? ((keep > 0 && line > keep / 2) ? Pl - keep + line : line) * Pm : line;
      const int jq = FIRST ? j : fdiv(j, rns), jn = FIRST ? 0 : j - jq * Ns;
      double2 x[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double2 e = lds_ld(X, base + (j + r * per) * stride);
        x[r] = (r && !FIRST) ? cmul(e, root_of(root, r * jn * step, inv)) : e;  // r jn step < P
      }
      dft_small<R>(x, v[u], inv);
      dst[u] = base + (jq * Ns * R + jn) * stride;
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < MAXB; ++u)
    if (dst[u] >= 0)
#pragma unroll
      for (int r = 0; r < R; ++r) lds_st(X, dst[u] + r * Ns * stride, v[u][r]);
  __syncthreads();
}

// all passes along one axis
__device__ void fft_axis(double2* X, const double2* root, const euvl::FftGrid& g, int axis, bool inv, int nlines,
                         int keep) {
  const bool along_m = axis == 1;
  const int P = g.P[axis];
  int Ns = 1;
  for (int s = 0; s < g.nst[axis]; ++s) {
    const int R = g.rad[axis][s];
#define EUVL_FFT_PASS(RR, INV)                                                        \
  if (s == 0)                                                                        \
    fft_pass<RR, INV, true>(X, root, P, Ns, along_m, nlines, g.P[0], g.P[1], keep);  \
  else                                                                               \
    fft_pass<RR, INV, false>(X, root, P, Ns, along_m, nlines, g.P[0], g.P[1], keep);
    if (inv) {
      switch (R) {
        case 2: EUVL_FFT_PASS(2, true) break;
        case 3: EUVL_FFT_PASS(3, true) break;
        case 9: EUVL_FFT_PASS(9, true) break;
        default: EUVL_FFT_PASS(4, true) break;
      }
    } else {
      switch (R) {
        case 2: EUVL_FFT_PASS(2, false) break;
        case 3: EUVL_FFT_PASS(3, false) break;
        case 9: EUVL_FFT_PASS(9, false) break;
        default: EUVL_FFT_PASS(4, false) break;
      }
    }
#undef EUVL_FFT_PASS
    Ns *= R;
  }
}

// root tables e^(-2 pi i t / P) of both axes
__device__ void fill_roots(const euvl::FftGrid& g, double2 (*root)[FFT_MAX_P]) {
  for (int t = threadIdx.x; t < g.P[0] + g.P[1]; t += blockDim.x) {
    const int a = t < g.P[0] ? 0 : 1, u = t - (a ? g.P[0] : 0);
    double s, c;
    sincospi(-2.0 * u / g.P[a], &s, &c);
    root[a][u] = make_double2(c, s);
  }
}

// x mod P for |x| < P
__device__ __forceinline__ int wrap(int x, int P) { return x < 0 ? x + P : x; }

// S_b = sum_z sigma^(z) e^{+2 pi i k.z / P} for batch entry b, with
// sigma^ = ea^-1 delta + (1 - ea^-1) m^ (m^: the clip's indicator spectrum,
// window (4 lmax + 1) x (4 mmax + 1))
__global__ __launch_bounds__(FFT_THREADS) void sigma_spectrum_kernel(euvl::FftGrid g, int lmax, int mmax,
                                                                     const double2* W, int per_clip,
                                                                     int64_t w_stride, double2 ea_inv,
                                                                     double2* S) {
  __shared__ double2 X[FFT_MAX_PTS];
  __shared__ double2 root[2][FFT_MAX_P];
  const int b = blockIdx.x, Pl = g.P[0], Pm = g.P[1], n = Pl * Pm;
  fill_roots(g, root);
  for (int e = threadIdx.x; e < n; e += blockDim.x) X[e] = make_double2(0.0, 0.0);
  __syncthreads();
  const int wl = 4 * lmax + 1, wm = 4 * mmax + 1;
  const double2* Wc = W + (int64_t)(b / per_clip) * w_stride;
  const double2 om = make_double2(1.0 - ea_inv.x, -ea_inv.y);
  for (int e = threadIdx.x; e < wl * wm; e += blockDim.x) {
    const int dl = e / wm - 2 * lmax, dm = e % wm - 2 * mmax;
    double2 s = cmul(om, Wc[e]);
    if (dl == 0 && dm == 0) s = cadd(s, ea_inv);
    X[wrap(dl, Pl) * Pm + wrap(dm, Pm)] = s;
  }
  __syncthreads();
  fft_axis(X, root[1], g, 1, true, Pl, 0);
  fft_axis(X, root[0], g, 0, true, Pm, 0);
  for (int e = threadIdx.x; e < n; e += blockDim.x) S[(int64_t)b * n + e] = X[e];
}

// V[b][i][:] -= (H G)[b][i][:], one workgroup per (i, b)
__global__ __launch_bounds__(FFT_THREADS) void coupling_fft_kernel(euvl::FftGrid g, int N, int lmax,
                                                                   const int* __restrict__ lidx,
                                                                   const int* __restrict__ midx,
                                                                   const double* __restrict__ kt, double scale,
                                                                   const double2* __restrict__ H,
                                                                   const double2* __restrict__ S, double2* V,
                                                                   int64_t ldv, int64_t sv) {
  __shared__ double2 X[FFT_MAX_PTS];
  __shared__ double2 root[2][FFT_MAX_P];
  const int i = blockIdx.x, b = blockIdx.y, Pl = g.P[0], Pm = g.P[1], n = Pl * Pm;
  const int keep = (2 * lmax + 1 < Pl) ? 2 * lmax + 1 : 0;
  const int rows = keep ? keep : Pl;
  fill_roots(g, root);
  for (int e = threadIdx.x; e < n; e += blockDim.x) X[e] = make_double2(0.0, 0.0);
  __syncthreads();
  const double2* Hi = H + ((int64_t)b * N + i) * N;
  const double* ktb = kt + (int64_t)b * N;
  for (int q = threadIdx.x; q < N; q += blockDim.x)
    X[wrap(lidx[q], Pl) * Pm + wrap(midx[q], Pm)] = cscale(Hi[q], ktb[q]);
  __syncthreads();
  fft_axis(X, root[1], g, 1, false, rows, keep);
  fft_axis(X, root[0], g, 0, false, Pm, 0);
  const double2* Sb = S + (int64_t)b * n;
  for (int e = threadIdx.x; e < n; e += blockDim.x) X[e] = cmul(X[e], Sb[e]);
  __syncthreads();
  fft_axis(X, root[0], g, 0, true, Pm, 0);
  fft_axis(X, root[1], g, 1, true, rows, keep);
  double2* Vi = V + (int64_t)b * sv + (int64_t)i * ldv;
  for (int j = threadIdx.x; j < N; j += blockDim.x) {
    const double2 c = X[wrap(lidx[j], Pl) * Pm + wrap(midx[j], Pm)];
    Vi[j] = csub(Vi[j], cscale(c, ktb[j] * scale));
  }
}

// Y[b][i][:] = (X D)[b][i][:] for the absorber operator D of the EM solve,
// one workgroup per (i, b).  With x = X[b][i][:], P = x * m^ and
// Q = (u x) * m^ (correlations with the indicator spectrum over the order
// grid, as in the coupling product), row i of X D is
//   y_j = (a_p + b_u kt_j u_j) P_j - b_u kt_j Q_j + (a_x - kxy2_j) x_j.
// The two correlations share the spectrum S of m^; each is a forward and an
// inverse 2D transform of the grid in LDS, and P, Q are kept in registers
// (outputs j = thread + t FFT_THREADS, t < XD_PER).
constexpr int XD_PER = 2;

__global__ __launch_bounds__(FFT_THREADS) void xd_fft_kernel(euvl::FftGrid g, int N, int lmax,
                                                             const int* __restrict__ lidx,
                                                             const int* __restrict__ midx,
                                                             const int* __restrict__ uidx,
                                                             const double* __restrict__ kt,
                                                             const double* __restrict__ kxy2, double2 a_p,
                                                             double2 b_u, double2 a_x, double scale,
                                                             const double2* __restrict__ X, int64_t sx,
                                                             const double2* __restrict__ S, double2* __restrict__ Y,
                                                             int64_t sy) {
  __shared__ double2 G[FFT_MAX_PTS];
  __shared__ double2 root[2][FFT_MAX_P];
  const int i = blockIdx.x, b = blockIdx.y, Pl = g.P[0], Pm = g.P[1], n = Pl * Pm;
  const int keep = (2 * lmax + 1 < Pl) ? 2 * lmax + 1 : 0;
  const int rows = keep ? keep : Pl;
  fill_roots(g, root);
  const double2* Xi = X + (int64_t)b * sx + (int64_t)i * N;
  const double2* Sb = S + (int64_t)b * n;
  double2 P[XD_PER], Q[XD_PER];
#pragma unroll 1
  for (int pass = 0; pass < 2; ++pass) {
    for (int e = threadIdx.x; e < n; e += blockDim.x) G[e] = make_double2(0.0, 0.0);
    __syncthreads();
    for (int q = threadIdx.x; q < N; q += blockDim.x) {
      const double2 x = Xi[q];
      G[wrap(lidx[q], Pl) * Pm + wrap(midx[q], Pm)] = pass ? cscale(x, (double)uidx[q]) : x;
    }
    __syncthreads();
    fft_axis(G, root[1], g, 1, false, rows, keep);
    fft_axis(G, root[0], g, 0, false, Pm, 0);
    for (int e = threadIdx.x; e < n; e += blockDim.x) G[e] = cmul(G[e], Sb[e]);
    __syncthreads();
    fft_axis(G, root[0], g, 0, true, Pm, 0);
    fft_axis(G, root[1], g, 1, true, rows, keep);
#pragma unroll
    for (int t = 0; t < XD_PER; ++t) {
      const int j = threadIdx.x + t * FFT_THREADS;
      const double2 c = j < N ? G[wrap(lidx[j], Pl) * Pm + wrap(midx[j], Pm)] : make_double2(0.0, 0.0);
      if (pass)
        Q[t] = c;
      else
        P[t] = c;
    }
    __syncthreads();  // every read of the grid precedes the next pass's clearing
  }
  double2* Yi = Y + (int64_t)b * sy + (int64_t)i * N;
#pragma unroll
  for (int t = 0; t < XD_PER; ++t) {
    const int j = threadIdx.x + t * FFT_THREADS;
    if (j < N) {
      const double ktj = kt[(int64_t)b * N + j];
      const double2 cp = cadd(a_p, cscale(b_u, ktj * (double)uidx[j]));
      const double2 cq = cscale(b_u, -ktj);
      const double2 cx = make_double2(a_x.x - kxy2[(int64_t)b * N + j], a_x.y);
      Yi[j] = cadd(cscale(cadd(cmul(cp, P[t]), cmul(cq, Q[t])), scale), cmul(cx, Xi[j]));
    }
  }
}

// smallest size >= lo with no prime factor above 3, and its radices (nines
// first, then fours, a two, threes); 0 when above FFT_MAX_P
int fft_size(int lo, int* rad, int* nst) {
  for (int P = std::max(lo, 1); P <= FFT_MAX_P; ++P) {
    int p = P, k = 0;
    for (int r : {9, 4, 2, 3})
      while (p % r == 0 && k < euvl::FFT_MAX_STAGES) {
        rad[k++] = r;
        p /= r;
      }
    if (p == 1) {
      *nst = k;
      return P;
    }
  }
  return 0;
}

}  // namespace

namespace euvl {

bool fft_grid_for(int lmax, int mmax, FftGrid& g) {
  g.P[0] = fft_size(4 * lmax + 1, g.rad[0], &g.nst[0]);
  g.P[1] = fft_size(4 * mmax + 1, g.rad[1], &g.nst[1]);
  return g.P[0] > 0 && g.P[1] > 0 && g.P[0] * g.P[1] <= FFT_MAX_PTS;
}

size_t coupling_fft_bytes(int E) { return sizeof(double2) * (size_t)FFT_MAX_PTS * E; }

int coupling_fft(euvl_handle h, const FftGrid& g, int N, int E, const int* lidx, const int* midx, int lmax,
                 int mmax, const double2* W, int per_clip, int64_t w_stride, double2 ea_inv, const double* kt,
                 double k, const double2* H, double2* V, int64_t ldv, int64_t sv, double2* S) {
  hipLaunchKernelGGL(sigma_spectrum_kernel, dim3(E), dim3(FFT_THREADS), 0, h->stream, g, lmax, mmax, W, per_clip,
                     w_stride, ea_inv, S);
  EUVL_LAUNCH_CHECK();
  const double scale = 1.0 / (k * k * (double)g.P[0] * (double)g.P[1]);
  hipLaunchKernelGGL(coupling_fft_kernel, dim3(N, E), dim3(FFT_THREADS), 0, h->stream, g, N, lmax, lidx, midx, kt,
                     scale, H, (const double2*)S, V, ldv, sv);
  EUVL_LAUNCH_CHECK();
  return EUVL_OK;
}

bool xd_fft_fits(int N) { return N <= XD_PER * FFT_THREADS; }

// The FFT kernels hold a grid of FFT_MAX_PTS points in LDS and run one
// workgroup per CU, so below some N the GEMMs they replace are faster.  EM
// solve on MI355X (FP64, ms, products by D and H G both by FFT / both by
// GEMM): N = 289 x 16 entries 3.9 / 2.1, 625 x 9: 9.4 / 6.0, 1089 x 4:
// 14.2 / 12.9, 1681 x 4: 31.9 / 38.2; the products by D alone pay from
// N ~ 1200, H G alone from N ~ 1000 (1089 x 4: 12.8 against 12.9; 625 x 9:
// 6.3 against 6.0).  EUVL_FFT_MIN_N (products by D) and
// EUVL_COUPLING_FFT_MIN_N (H G) override the defaults (read once).
std::atomic<int> g_fft_min_n[2] = {-1, -1};  // -1: not read from the environment yet

int fft_min_n(bool coupling) {
  const int k = coupling ? 1 : 0;
  int v = g_fft_min_n[k].load(std::memory_order_relaxed);
  if (v < 0) {
    const char* e = getenv(coupling ? "EUVL_COUPLING_FFT_MIN_N" : "EUVL_FFT_MIN_N");
    v = e ? std::max(0, atoi(e)) : (coupling ? 800 : 1200);
    int expect = -1;
    if (!g_fft_min_n[k].compare_exchange_strong(expect, v)) v = expect;
  }
  return v;
}

int xd_fft_spectra(euvl_handle h, const FftGrid& g, int lmax, int mmax, const double2* W, int per_clip,
                   int64_t w_stride, int E, double2* S) {
  // the spectrum of m^ alone: sigma_spectrum_kernel with ea^-1 = 0
  hipLaunchKernelGGL(sigma_spectrum_kernel, dim3(E), dim3(FFT_THREADS), 0, h->stream, g, lmax, mmax, W, per_clip,
                     w_stride, make_double2(0.0, 0.0), S);
  EUVL_LAUNCH_CHECK();
  return EUVL_OK;
}

int xd_fft(euvl_handle h, const XdOp& op, int E, const double2* X, double2* Y) {
  if (!xd_fft_fits(op.N) || X == Y) {
    set_error("xd_fft: N = %d above %d, or Y aliases X", op.N, XD_PER * FFT_THREADS);
    return EUVL_ERR_ARG;
  }
  const double scale = 1.0 / ((double)op.g.P[0] * (double)op.g.P[1]);
  const int64_t sm = (int64_t)op.N * op.N;
  hipLaunchKernelGGL(xd_fft_kernel, dim3(op.N, E), dim3(FFT_THREADS), 0, h->stream, op.g, op.N, op.lmax, op.lidx,
                     op.midx, op.uidx, op.kt, op.kxy2, op.a_p, op.b_u, op.a_x, scale, X, sm, op.S, Y, sm);
  EUVL_LAUNCH_CHECK();
  return EUVL_OK;
}

}  // namespace euvl

extern "C" int euvl_set_fft_min_n(int products, int coupling) {
  if (products >= 0) euvl::g_fft_min_n[0].store(products);
  if (coupling >= 0) euvl::g_fft_min_n[1].store(coupling);
  return EUVL_OK;
}
