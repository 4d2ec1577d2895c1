// Rigorous EM solve of one absorber layer on the Mo/Si multilayer: mask
// spectrum, per-mode multilayer cascade, Toeplitz assembly and the
// eigen-free ampS of cnn/include/header.h:1208-1466.
//
// Math (all matrices N x N over the order set, per incidence offset):
//   D  = k^2 eps^ - i eta^ diag(kt) - diag(kxy2)          (header.h:806-821)
//   S  = D^(1/2), E = exp(i d S)                           (V, al1, gamma of header.h:834-921, 1088)
//   M  = i (k I - diag(kt/k) Sigma), Sigma = sigma^ diag(kt) (header.h:979-1052)
//   W1 = M^-1 diag(b2 (uu+ub)), W2 = S^-1 diag(alru (uu-ub)) (Cj, al2/al1 Cjp of header.h:1053-1144)
//   P  = 1/2 E^-1 (W1+W2) = V U1U,  Q = 1/2 E (W1-W2) = V U1B
//   R  = S (P-Q),  U0 = diag(klm)(P+Q) + R                  (header.h:1355-1370)
//   FG = diag(1/klm) R U0^-1                                (header.h:1371-1391)
//   Ax(:, inc) = -2 kz FG(:, idx(inc)) + e_idx(inc)         (header.h:1394-1465)
// With C = cos(d S), F = sin(d S) S^-1, G2 = D F (even in S: power series in D):
//   P - Q = C W2 - i S F W1,  P + Q = C W1 - i F diag(alru (uu-ub))
//   R = C diag(alru (uu-ub)) - i G2 W1,  U0 = diag(klm)(C W1 - i F diag(alru (uu-ub))) + R
// so neither the eigenvectors, nor S, nor a square-root branch is needed.
#include <cmath>
#include <algorithm>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "ew.h"
#include "optics.h"

namespace {

const double2 ONE = {1.0, 0.0}, ZERO = {0.0, 0.0}, MONE = {-1.0, 0.0}, MI = {0.0, -1.0};

__device__ inline int mask_bit(const uint8_t* bits, int64_t idx) { return (bits[idx >> 3] >> (7 - (idx & 7))) & 1; }

// Reference index rule of fourier(): j = (i*l) % N for l >= 0, N - ((-i*l) % N) otherwise (header.h:1629-1631)
__device__ inline int tw_index(int64_t i, int l, int N) {
  if (l >= 0) return (int)((i * l) % N);
  return N - (int)((-i * l) % N);
}

// T[c][i][mm] = sum_j m(i, j) tw_y[idx(j, m)], m = mm - mw; y row j is bitmap row NDIVY-1-j.
// A real-by-complex product (bits 0/1 times the twiddle table) on FP64 MFMAs:
// a workgroup of 4 waves covers 64 x rows and 64 orders over one of S splits
// of the y range; lane l feeds bit (i0 + l%16, j0 + l/16) as A and the
// table entry of (j0 + l/16, order l%16) as B (re and im accumulated
// separately).  Split s writes partial sums P[s][c][i][mm], summed in split
// order by mask_dft_reduce_kernel (deterministic).
constexpr int DFT_TILE = 64;
__global__ __launch_bounds__(256) void mask_dft_y_kernel(const uint8_t* __restrict__ bits, int nx, int ny, int mw,
                                                         const double2* __restrict__ twy, double2* __restrict__ P,
                                                         int n_clips, int S, int ks) {
  const int nm = 2 * mw + 1;
  const int s = blockIdx.z % S, c = blockIdx.z / S;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const int i = blockIdx.x * DFT_TILE + w * 16 + li;
  const int m0 = blockIdx.y * DFT_TILE;
  const uint8_t* b = bits + (int64_t)c * nx * ny / 8;
  double4_t re[4], im[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    re[t] = (double4_t){0, 0, 0, 0};
    im[t] = (double4_t){0, 0, 0, 0};
  }
  const int j0 = s * ks, j1 = min(ny, j0 + ks);
  for (int k0 = j0; k0 < j1; k0 += 4) {
    const int j = k0 + lk;
    const bool jin = j < j1;
    const double a = (jin && i < nx) ? (double)mask_bit(b, (int64_t)nx * (ny - 1 - j) + i) : 0.0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int mm = m0 + t * 16 + li;
      const double2 tw = (jin && mm < nm) ? twy[tw_index(j, mm - mw, ny)] : make_double2(0.0, 0.0);
      re[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, tw.x, re[t], 0, 0, 0);
      im[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, tw.y, im[t], 0, 0, 0);
    }
  }
  // accumulator element r of lane l: row 4r + l/16 (x), column l%16 (order)
  double2* out = P + ((int64_t)s * n_clips + c) * nx * nm;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int ii = blockIdx.x * DFT_TILE + w * 16 + 4 * r + lk, mm = m0 + t * 16 + li;
      if (ii < nx && mm < nm) out[(int64_t)ii * nm + mm] = make_double2(re[t][r], im[t][r]);
    }
}

__global__ void mask_dft_reduce_kernel(const double2* __restrict__ P, int S, int64_t per, double2* __restrict__ T) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < per; e += (int64_t)gridDim.x * blockDim.x) {
    double2 acc = P[e];
    for (int s = 1; s < S; ++s) acc = cadd(acc, P[(int64_t)s * per + e]);
    T[e] = acc;
  }
}

// out[c][l][mm] = (1/(nx ny)) sum_i T[c][i][mm] tw_x[idx(i, l)]: one
// workgroup per output, a strided sum per thread and a tree in LDS
__global__ __launch_bounds__(256) void mask_dft_x_kernel(const double2* __restrict__ T, int nx, int ny, int lw,
                                                         int mw, const double2* __restrict__ twx,
                                                         double2* __restrict__ out, int n_clips) {
  const int nm = 2 * mw + 1, nl = 2 * lw + 1;
  const int64_t e = blockIdx.x;
  const int c = (int)(e / ((int64_t)nl * nm));
  const int64_t r = e - (int64_t)c * nl * nm;
  const int mm = (int)(r % nm);
  const int l = (int)(r / nm) - lw;
  const double2* t = T + (int64_t)c * nx * nm + mm;
  double2 s = make_double2(0.0, 0.0);
  for (int i = threadIdx.x; i < nx; i += blockDim.x) s = cadd(s, cmul(t[(int64_t)i * nm], twx[tw_index(i, l, nx)]));
  __shared__ double2 red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] = cadd(red[threadIdx.x], red[threadIdx.x + k]);
    __syncthreads();
  }
  // the reference divides by FDIVY after the y sum and by FDIVX after the x sum
  if (threadIdx.x == 0) out[e] = cscale(red[0], 1.0 / ((double)nx * (double)ny));
}

struct ModeVecs {
  double* kt;      // kx (X) or ky (Y)
  double* kxy2;
  double2* klm;
  double2* uu;
  double2* ub;
  double2* alru;
  double2* b2;
};

// Per-mode vectors and the multilayer cascade (header.h:5-363, 1277-1303).
__global__ void mode_vectors_kernel(int n, int batch, char polar, double k, double dx, double dy, const int* lidx,
                                    const int* midx, const double* sx0, const double* sy0, int n_layers,
                                    const double2* leps, const double* ld, ModeVecs v, double* kx_out,
                                    double* ky_out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * batch) return;
  const int b = (int)(e / n), i = (int)(e % n);
  const double kx = sx0[b] + 2.0 * PI * lidx[i] / dx;
  const double ky = sy0[b] + 2.0 * PI * midx[i] / dy;
  const double kxy2 = kx * kx + ky * ky;
  const double kt = (polar == 'X') ? kx : ky;
  const double k2 = k * k;
  // cascade from the substrate: (U, B) = (1, 0), then T(layer j, layer j-1) for j = 1..n_layers-1
  double2 U = make_double2(1.0, 0.0), B = make_double2(0.0, 0.0);
  double2 e_prev = leps[0];
  double2 al_prev = csqrt_principal(make_double2(k2 * e_prev.x - kxy2, k2 * e_prev.y));
  // f(e) = k - kt^2/(k e)
  double2 f_prev = csub(make_double2(k, 0.0), cdiv(make_double2(kt * kt, 0.0), cscale(e_prev, k)));
  for (int j = 1; j < n_layers; ++j) {
    const double2 ej = leps[j];
    const double2 al = csqrt_principal(make_double2(k2 * ej.x - kxy2, k2 * ej.y));
    const double2 f = csub(make_double2(k, 0.0), cdiv(make_double2(kt * kt, 0.0), cscale(ej, k)));
    const double2 cj = cdiv(f_prev, f);
    const double2 r = cdiv(al_prev, al);
    const double2 g = cexpi(cmul(make_double2(0.0, ld[j]), al));  // exp(i al d)
    const double2 ginv = cdiv(ONE, g);
    const double2 p = cscale(cadd(cj, r), 0.5), m = cscale(csub(cj, r), 0.5);
    const double2 nU = cadd(cmul(cmul(p, ginv), U), cmul(cmul(m, ginv), B));
    const double2 nB = cadd(cmul(cmul(m, g), U), cmul(cmul(p, g), B));
    U = nU;
    B = nB;
    al_prev = al;
    f_prev = f;
  }
  const double2 eru = leps[n_layers - 1];
  v.kt[e] = kt;
  v.kxy2[e] = kxy2;
  // klm = sqrt(k^2 - kxy2); evanescent orders take the Im >= 0 branch (the reference's real sqrt gives NaN there)
  const double d = k2 - kxy2;
  v.klm[e] = (d >= 0.0) ? make_double2(sqrt(d), 0.0) : make_double2(0.0, sqrt(-d));
  v.uu[e] = U;
  v.ub[e] = B;
  v.alru[e] = csqrt_principal(make_double2(k2 * eru.x - kxy2, k2 * eru.y));
  // b2 = i k - i kt^2/(k eru)  (header.h:1294-1297)
  v.b2[e] = csub(make_double2(0.0, k), cmul(make_double2(0.0, 1.0), cdiv(make_double2(kt * kt, 0.0), cscale(eru, k))));
  if (kx_out) kx_out[e] = kx;
  if (ky_out) ky_out[e] = ky;
}

// D and G = diag(kt) sigma^ diag(kt) / k^2 from the indicator spectrum W
// (window |dl| <= 2 lmax, |dm| <= 2 mmax).
// Batch entry b reads the spectrum W + (b / per_clip) * w_stride (the
// offsets of one clip share its spectrum).
__global__ void toeplitz_kernel(int n, int batch, char polar, double k, double dx, double dy, const int* lidx,
                                const int* midx, int lmax, int mmax, const double2* W, int per_clip, int64_t w_stride,
                                double2 ea, double2 ea_inv, double2 log_ea, ModeVecs v, double2* D, double2* G) {
  const int64_t per = (int64_t)n * n;
  const int64_t total = per * batch;
  const int wm = 4 * mmax + 1;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / per);
    const int64_t r = e - (int64_t)b * per;
    const int i = (int)(r / n), j = (int)(r % n);
    const int dl = lidx[i] - lidx[j], dm = midx[i] - midx[j];
    const double2 mh = W[(int64_t)(b / per_clip) * w_stride + (int64_t)(dl + 2 * lmax) * wm + (dm + 2 * mmax)];
    const bool z = (dl == 0 && dm == 0);
    // eps^ = ea delta + (1-ea) m^ ; log-eps^ = log(ea)(delta - m^) ; sigma^ = ea^-1 delta + (1-ea^-1) m^
    double2 eps = cmul(make_double2(1.0 - ea.x, -ea.y), mh);
    double2 sig = cmul(make_double2(1.0 - ea_inv.x, -ea_inv.y), mh);
    double2 lep = cmul(log_ea, make_double2(-mh.x, -mh.y));
    if (z) {
      eps = cadd(eps, ea);
      sig = cadd(sig, ea_inv);
      lep = cadd(lep, log_ea);
    }
    const double ktj = v.kt[(int64_t)b * n + j], kti = v.kt[(int64_t)b * n + i];
    const double fac = (polar == 'X') ? 2.0 * PI * dl / dx : 2.0 * PI * dm / dy;
    // eta^ = i fac lep ; D = k^2 eps - i eta^ kt_j = k^2 eps + fac lep kt_j
    double2 dv = cadd(cscale(eps, k * k), cscale(lep, fac * ktj));
    if (i == j) dv.x -= v.kxy2[(int64_t)b * n + i];
    D[e] = dv;
    G[e] = cscale(sig, kti * ktj / (k * k));
  }
}

// lam_j = alru (uu - ub) i k / (b2 (uu + ub)): L2 S^-1 of the top solve, with
// L2 = diag(alru (uu - ub)) and S = diag(b2 (uu + ub) / (i k)).  The cascade's
// scale |uu| cancels: only the multilayer's reflection ratio ub/uu enters.
__device__ __forceinline__ double2 top_lambda(const ModeVecs& v, int64_t j, double k) {
  const double2 uu = v.uu[j], ub = v.ub[j];
  return cdiv(cmul(cmul(v.alru[j], csub(uu, ub)), make_double2(0.0, k)), cmul(v.b2[j], cadd(uu, ub)));
}

// H = (C - i diag(klm) F) diag(lam),  V = diag(klm) C - i G2 + H  (V gets - H G next)
__global__ void top_hv_kernel(int n, int batch, double k, const double2* C, const double2* F, const double2* G2,
                              double2* H, double2* V, int64_t ldv, int64_t sv, ModeVecs v) {
  const int64_t per = (int64_t)n * n;
  const int64_t total = per * batch;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / per);
    const int64_t r = e - (int64_t)b * per;
    const int i = (int)(r / n), j = (int)(r % n);
    const double2 klm = v.klm[(int64_t)b * n + i];
    const double2 c = C[e], kf = cmul(klm, F[e]), g2 = G2[e], kc = cmul(klm, c);
    const double2 hv = cmul(make_double2(c.x + kf.y, c.y - kf.x), top_lambda(v, (int64_t)b * n + j, k));
    H[e] = hv;
    V[(int64_t)b * sv + (int64_t)i * ldv + j] = make_double2(kc.x + g2.y + hv.x, kc.y - g2.x + hv.y);  // klm c - i g2 + h
  }
}

// Y = diag(lam) (Z - G Z), in place over GZ (n x ninc)
__global__ void top_y_kernel(int n, int ninc, int batch, double k, const double2* Z, double2* GZ, ModeVecs v) {
  const int64_t per = (int64_t)n * ninc;
  const int64_t total = per * batch;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / per);
    const int i = (int)((e - (int64_t)b * per) / ninc);
    GZ[e] = cmul(top_lambda(v, (int64_t)b * n + i, k), csub(Z[e], GZ[e]));
  }
}

__global__ void fill_kernel(int64_t count, double2* out, double2 v) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (int64_t)gridDim.x * blockDim.x)
    out[e] = v;
}

// right-hand sides e_idx(c) for every batch entry: Bm (n x ninc, leading
// dimension ld, batch stride sb)
__global__ void rhs_kernel(int n, int ninc, int batch, const int* idx, double2* Bm, int64_t ld, int64_t sb) {
  const int64_t per = (int64_t)n * ninc;
  const int64_t total = per * batch;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / per);
    const int64_t r = e - (int64_t)b * per;
    const int i = (int)(r / ninc), c = (int)(r % ninc);
    Bm[(int64_t)b * sb + (int64_t)i * ld + c] = (idx[c] == i) ? make_double2(1.0, 0.0) : make_double2(0.0, 0.0);
  }
}

// dst = src for a rows x cols block of every batch entry (leading dimensions
// and batch strides of their own)
__global__ void copy2d_kernel(int rows, int cols, int batch, const double2* src, int64_t lds, int64_t ss, double2* dst,
                              int64_t ldd, int64_t sd) {
  const int64_t per = (int64_t)rows * cols;
  const int64_t total = per * batch;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / per);
    const int64_t r = e - (int64_t)b * per;
    const int i = (int)(r / cols), c = (int)(r % cols);
    dst[(int64_t)b * sd + (int64_t)i * ldd + c] = src[(int64_t)b * ss + (int64_t)i * lds + c];
  }
}

// Ax[b][slot(c)][i] = -2 kz[b][c] FGc[b][i][c] / klm[b][i] + (i == idx[c])
__global__ void ax_scatter_kernel(int n, int ninc, int batch, const int* idx, const int* slot, const double* kz,
                                  const double2* FGc, const double2* klm, int64_t ax_stride_b, double2* Ax) {
  const int64_t total = (int64_t)n * ninc * batch;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / ((int64_t)n * ninc));
    const int64_t r = e - (int64_t)b * n * ninc;
    const int c = (int)(r / n), i = (int)(r % n);
    const double2 fg = cdiv(FGc[(int64_t)b * n * ninc + (int64_t)i * ninc + c], klm[(int64_t)b * n + i]);
    double2 v = cscale(fg, -2.0 * kz[b * ninc + c]);
    if (i == idx[c]) v.x += 1.0;
    Ax[(int64_t)b * ax_stride_b + (int64_t)slot[c] * n + i] = v;
  }
}

int check_params(const euvl_em_params* p) {
  if (!p || p->ndivx <= 0 || p->ndivy <= 0 || p->n_layers < 2 || !p->h_layer_eps || !p->h_layer_d || p->k <= 0.0 ||
      p->dx <= 0.0 || p->dy <= 0.0) {
    euvl::set_error("invalid euvl_em_params");
    return EUVL_ERR_ARG;
  }
  return EUVL_OK;
}

}  // namespace

namespace euvl {
int fill(euvl_handle h, int64_t count, double2* out, double2 v) {
  if (count <= 0) return EUVL_OK;
  hipLaunchKernelGGL(fill_kernel, grid1(count), dim3(256), 0, h->stream, count, out, v);
  EUVL_LAUNCH_CHECK();
  return EUVL_OK;
}

// splits of the y range of mask_dft_y_kernel: enough workgroups to fill the GPU
int dft_splits(int nx, int ny, int mw, int n_clips) {
  const int64_t tiles = (int64_t)((nx + DFT_TILE - 1) / DFT_TILE) * ((2 * mw + 1 + DFT_TILE - 1) / DFT_TILE) * n_clips;
  int S = (int)std::min<int64_t>(16, std::max<int64_t>(1, 1024 / tiles));
  while (S > 1 && (int64_t)S * 64 > ny) --S;  // at least 64 rows per split
  return S;
}

// twiddle tables of exponential() (header.h:1637-1648): nx + 1 entries for x, then ny + 1 for y
void dft_twiddles(int nx, int ny, std::vector<double2>& t) {
  t.resize(nx + ny + 2);
  const std::complex<double> zi(0.0, 1.0);
  for (int i = 0; i <= nx; ++i) {
    const std::complex<double> e = std::exp(-2.0 * PI * zi * double(i) / double(nx));
    t[i] = make_double2(e.real(), e.imag());
  }
  for (int i = 0; i <= ny; ++i) {
    const std::complex<double> e = std::exp(-2.0 * PI * zi * double(i) / double(ny));
    t[nx + 1 + i] = make_double2(e.real(), e.imag());
  }
}

int mask_spectrum_enqueue(euvl_handle h, const uint8_t* bits, int nx, int ny, int n_clips, int lw, int mw,
                          double2* out, double2* scratch, double2* part, const double2* tw) {
  return mask_spectrum_splits(h, bits, nx, ny, n_clips, lw, mw, out, scratch, part, tw,
                              dft_splits(nx, ny, mw, n_clips));
}

// the same with the number S of y-range splits given (part holds S partial sums when S > 1)
int mask_spectrum_splits(euvl_handle h, const uint8_t* bits, int nx, int ny, int n_clips, int lw, int mw,
                         double2* out, double2* scratch, double2* part, const double2* tw, int S) {
  const int nm = 2 * mw + 1;
  const int ks = ((ny + S - 1) / S + 3) / 4 * 4;
  const dim3 g1((nx + DFT_TILE - 1) / DFT_TILE, (nm + DFT_TILE - 1) / DFT_TILE, S * n_clips);
  hipLaunchKernelGGL(mask_dft_y_kernel, g1, dim3(256), 0, h->stream, bits, nx, ny, mw, tw + nx + 1,
                     S > 1 ? part : scratch, n_clips, S, ks);
  EUVL_LAUNCH_CHECK();
  if (S > 1) {
    const int64_t per = (int64_t)n_clips * nx * nm;
    hipLaunchKernelGGL(mask_dft_reduce_kernel, grid1(per), dim3(256), 0, h->stream, part, S, per, scratch);
    EUVL_LAUNCH_CHECK();
  }
  const int64_t n2 = (int64_t)n_clips * (2 * lw + 1) * nm;
  hipLaunchKernelGGL(mask_dft_x_kernel, dim3((unsigned)n2), dim3(256), 0, h->stream, scratch, nx, ny, lw, mw, tw, out,
                     n_clips);
  EUVL_LAUNCH_CHECK();
  return EUVL_OK;
}

int mask_spectrum(euvl_handle h, const uint8_t* bits, int nx, int ny, int n_clips, int lw, int mw, double2* out,
                  double2* scratch /* n_clips*nx*(2mw+1) */, double2* part /* S*n_clips*nx*(2mw+1), S > 1 */,
                  double2* tw /* nx+1 + ny+1 */) {
  std::vector<double2> t;
  dft_twiddles(nx, ny, t);
  EUVL_TRY(upload_async(h, tw, t.data(), sizeof(double2) * t.size()));
  return mask_spectrum_enqueue(h, bits, nx, ny, n_clips, lw, mw, out, scratch, part, tw);
}
}  // namespace euvl

extern "C" int euvl_mask_spectrum(euvl_handle h, const uint8_t* d_bits, int ndivx, int ndivy, int n_clips, int lw,
                                  int mw, double* d_out) {
  if (!h || !d_bits || !d_out || ndivx <= 0 || ndivy <= 0 || n_clips <= 0 || n_clips > 65535 || lw < 0 || mw < 0 ||
      ((int64_t)ndivx * ndivy) % 8 != 0 || (int64_t)n_clips * (2 * lw + 1) * (2 * mw + 1) >= ((int64_t)1 << 31)) {
    euvl::set_error("euvl_mask_spectrum: bad argument");
    return EUVL_ERR_ARG;
  }
  const size_t scratch = sizeof(double2) * (size_t)n_clips * ndivx * (2 * mw + 1);
  const int S = euvl::dft_splits(ndivx, ndivy, mw, n_clips);
  const size_t part = S > 1 ? S * scratch : 0;
  const size_t tw = sizeof(double2) * (size_t)(ndivx + ndivy + 2);
  EUVL_TRY(euvl::ws_reserve(h, scratch + part + tw + 1024));
  double2* s = (double2*)euvl::ws_take(h, scratch);
  double2* pp = part ? (double2*)euvl::ws_take(h, part) : nullptr;
  double2* t = (double2*)euvl::ws_take(h, tw);
  return euvl::mask_spectrum(h, d_bits, ndivx, ndivy, n_clips, lw, mw, (double2*)d_out, s, pp, t);
}

// Scratch of the allocation-free mask spectrum: a fixed device array (64 MiB
// per GPU, held until the process ends).  Host code never uses the name as an
// address: the device address comes from hipGetSymbolAddress, once per device.
namespace {
constexpr int64_t NOALLOC_SCRATCH = (int64_t)1 << 22;  // double2 entries
static __device__ __align__(256) double2 scratch[NOALLOC_SCRATCH];
std::mutex g_scratch_mu;
double2* g_scratch_ptr[64] = {};

int scratch_for_device(int dev, double2** out) {
  if (dev < 0 || dev >= 64) {
    euvl::set_error("euvl_mask_spectrum_noalloc: device %d out of range", dev);
    return EUVL_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      euvl::set_error("euvl_mask_spectrum_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
      return EUVL_ERR_HIP;
    }
    g_scratch_ptr[dev] = (double2*)p;
  }
  *out = g_scratch_ptr[dev];
  return EUVL_OK;
}
}  // namespace

extern "C" int64_t euvl_mask_spectrum_noalloc_max(void) { return NOALLOC_SCRATCH * (int64_t)sizeof(double2); }

extern "C" int euvl_mask_spectrum_noalloc(euvl_handle h, const uint8_t* d_bits, int ndivx, int ndivy, int n_clips,
                                          int lw, int mw, double* d_out) {
  if (!h || !d_bits || !d_out || ndivx <= 0 || ndivy <= 0 || n_clips <= 0 || n_clips > 65535 || lw < 0 || mw < 0 ||
      ((int64_t)ndivx * ndivy) % 8 != 0 || (int64_t)n_clips * (2 * lw + 1) * (2 * mw + 1) >= ((int64_t)1 << 31)) {
    euvl::set_error("euvl_mask_spectrum_noalloc: bad argument");
    return EUVL_ERR_ARG;
  }
  int dev = -1;
  EUVL_HIP_CHECK(hipGetDevice(&dev));
  if (dev != h->device) {
    euvl::set_error("euvl_mask_spectrum_noalloc: current device %d is not the handle's device %d", dev, h->device);
    return EUVL_ERR_ARG;
  }
  double2* base = nullptr;
  EUVL_TRY(scratch_for_device(dev, &base));
  // layout in the scratch array: twiddles, the y sums, then the split partials;
  // every entry read is written first in the same call, so nothing needs clearing
  const int64_t tw = ndivx + ndivy + 2, tw_al = (tw + 15) / 16 * 16;
  const int64_t per = (int64_t)n_clips * ndivx * (2 * mw + 1), per_al = (per + 15) / 16 * 16;
  if (tw_al + per_al > NOALLOC_SCRATCH) {
    euvl::set_error("euvl_mask_spectrum_noalloc: %d clips of %d x %d px with %d orders need %lld bytes of scratch, "
                    "above the maximum of %lld", n_clips, ndivx, ndivy, 2 * mw + 1,
                    (long long)(16 * (tw_al + per_al)), (long long)(16 * NOALLOC_SCRATCH));
    return EUVL_ERR_ARG;
  }
  // as many y-range splits as the scratch holds (the same count as euvl_mask_spectrum when it fits)
  int S = euvl::dft_splits(ndivx, ndivy, mw, n_clips);
  while (S > 1 && tw_al + per_al * (1 + S) > NOALLOC_SCRATCH) --S;
  double2* t = base;
  double2* sc = base + tw_al;
  double2* part = S > 1 ? sc + per_al : nullptr;
  std::vector<double2> host;
  euvl::dft_twiddles(ndivx, ndivy, host);
  EUVL_HIP_CHECK(hipMemcpyAsync(t, host.data(), sizeof(double2) * host.size(), hipMemcpyHostToDevice, h->stream));
  EUVL_TRY(euvl::mask_spectrum_splits(h, d_bits, ndivx, ndivy, n_clips, lw, mw, (double2*)d_out, sc, part, t, S));
  // the host table must outlive the copy
  EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));
  return EUVL_OK;
}

extern "C" int euvl_multilayer(euvl_handle h, char polar, const euvl_em_params* p, int n, const double* d_kx,
                               const double* d_ky, double* d_uu, double* d_ub) {
  // Implemented through the mode-vector kernel with the order set given
  // directly as wave vectors: lidx = midx = 0 and per-mode offsets (kx, ky).
  EUVL_TRY(check_params(p));
  if (!h || n <= 0 || !d_kx || !d_ky || !d_uu || !d_ub || (polar != 'X' && polar != 'Y')) {
    euvl::set_error("euvl_multilayer: bad argument");
    return EUVL_ERR_ARG;
  }
  Upload up;
  std::vector<int> zeros(1, 0);
  size_t o_l = up.add(zeros.data(), sizeof(int));
  size_t o_eps = up.add(p->h_layer_eps, sizeof(double2) * p->n_layers);
  size_t o_d = up.add(p->h_layer_d, sizeof(double) * p->n_layers);
  const size_t vec = sizeof(double2) * (size_t)n;
  EUVL_TRY(euvl::ws_reserve(h, up.host.size() + 8 * vec + 4096));
  char* blk = (char*)euvl::ws_take(h, up.host.size());
  EUVL_HIP_CHECK(hipMemcpyAsync(blk, up.host.data(), up.host.size(), hipMemcpyHostToDevice, h->stream));
  ModeVecs v;
  v.kt = (double*)euvl::ws_take(h, vec);
  v.kxy2 = (double*)euvl::ws_take(h, vec);
  v.klm = (double2*)euvl::ws_take(h, vec);
  v.uu = (double2*)d_uu;
  v.ub = (double2*)d_ub;
  v.alru = (double2*)euvl::ws_take(h, vec);
  v.b2 = (double2*)euvl::ws_take(h, vec);
  // batch = n entries of one mode each: lidx = midx = 0 and sx0/sy0 = kx/ky per entry
  hipLaunchKernelGGL(mode_vectors_kernel, grid1(n), dim3(256), 0, h->stream, 1, n, polar, p->k, p->dx, p->dy,
                     (const int*)(blk + o_l), (const int*)(blk + o_l), d_kx, d_ky, p->n_layers,
                     (const double2*)(blk + o_eps), (const double*)(blk + o_d), v, (double*)nullptr,
                     (double*)nullptr);
  EUVL_LAUNCH_CHECK();
  EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));
  return EUVL_OK;
}

namespace euvl {

// Host tables of one EM solve (order set, incidence columns), replicated over
// the batch entries: entry b = clip * n_off + offset.
void build_solve_tables(const euvl_em_params* p, const int* lindex, const int* mindex, int N, int lmax, int mmax,
                        int n_off, const double* sx0, const double* sy0, int n_clips, SolveHost& t) {
  const int nsy = 2 * p->lsmax_y + 1;
  t.inc_idx.clear();
  t.inc_slot.clear();
  std::vector<int> inc_ls, inc_ms;
  std::vector<int> pos((2 * lmax + 1) * (2 * mmax + 1), -1);
  for (int i = 0; i < N; ++i) pos[(lindex[i] + lmax) * (2 * mmax + 1) + (mindex[i] + mmax)] = i;
  // lattice incidences inside the NA (header.h:1394-1398) and their modes
  for (int ls = -p->lsmax_x; ls <= p->lsmax_x; ++ls)
    for (int ms = -p->lsmax_y; ms <= p->lsmax_y; ++ms) {
      const double a = ls * p->mx / p->dx, b = ms * p->my / p->dy, c = p->na / p->wavelength;
      if (a * a + b * b > c * c) continue;
      if (std::abs(ls) > lmax || std::abs(ms) > mmax) continue;
      const int i = pos[(ls + lmax) * (2 * mmax + 1) + (ms + mmax)];
      if (i < 0) continue;  // no such mode: AS = 0 and Ax stays 0 (header.h:1409-1415)
      t.inc_idx.push_back(i);
      t.inc_slot.push_back((ls + p->lsmax_x) * nsy + (ms + p->lsmax_y));
      inc_ls.push_back(ls);
      inc_ms.push_back(ms);
    }
  const int ninc = (int)t.inc_idx.size(), E = n_off * n_clips;
  t.sx.resize(E);
  t.sy.resize(E);
  t.kz.assign((size_t)E * std::max(ninc, 1), 0.0);
  for (int b = 0; b < E; ++b) {
    const int o = b % n_off;
    t.sx[b] = sx0[o];
    t.sy[b] = sy0[o];
    for (int c = 0; c < ninc; ++c) {
      const double kx = sx0[o] + inc_ls[c] * 2.0 * PI / p->dx;
      const double ky = sy0[o] + inc_ms[c] * 2.0 * PI / p->dy;
      t.kz[(size_t)b * ninc + c] = std::sqrt(p->k * p->k - kx * kx - ky * ky);
    }
  }
}

// Uploads the tables into `blk` (device) and fills the device pointers.
size_t solve_tables_bytes(const euvl_em_params* p, int N, const SolveHost& t) {
  size_t top = 0;
  auto add = [&](size_t bytes) { top = ((top + 255) & ~size_t(255)) + bytes; };
  add(sizeof(int) * N);
  add(sizeof(int) * N);
  add(sizeof(double) * t.sx.size());
  add(sizeof(double) * t.sy.size());
  add(sizeof(double2) * p->n_layers);
  add(sizeof(double) * p->n_layers);
  add(sizeof(int) * std::max<size_t>(t.inc_idx.size(), 1));
  add(sizeof(int) * std::max<size_t>(t.inc_slot.size(), 1));
  add(sizeof(double) * t.kz.size());
  return top;
}

// `up` receives the packed host copy (uploaded through a pinned staging
// buffer, so it may be freed when this returns).
int upload_solve_tables(euvl_handle h, const euvl_em_params* p, const int* lindex, const int* mindex, int N,
                        const SolveHost& t, Upload& up, char* blk, SolveDev& d) {
  const size_t o_l = up.add(lindex, sizeof(int) * N);
  const size_t o_m = up.add(mindex, sizeof(int) * N);
  const size_t o_sx = up.add(t.sx.data(), sizeof(double) * t.sx.size());
  const size_t o_sy = up.add(t.sy.data(), sizeof(double) * t.sy.size());
  const size_t o_eps = up.add(p->h_layer_eps, sizeof(double2) * p->n_layers);
  const size_t o_d = up.add(p->h_layer_d, sizeof(double) * p->n_layers);
  const size_t o_idx = up.add(t.inc_idx.data(), sizeof(int) * std::max<size_t>(t.inc_idx.size(), 1));
  const size_t o_slot = up.add(t.inc_slot.data(), sizeof(int) * std::max<size_t>(t.inc_slot.size(), 1));
  const size_t o_kz = up.add(t.kz.data(), sizeof(double) * t.kz.size());
  EUVL_TRY(upload_async(h, blk, up.host.data(), up.host.size()));
  d.lidx = (const int*)(blk + o_l);
  d.midx = (const int*)(blk + o_m);
  d.sx0 = (const double*)(blk + o_sx);
  d.sy0 = (const double*)(blk + o_sy);
  d.leps = (const double2*)(blk + o_eps);
  d.ld = (const double*)(blk + o_d);
  d.inc_idx = (const int*)(blk + o_idx);
  d.inc_slot = (const int*)(blk + o_slot);
  d.kz = (const double*)(blk + o_kz);
  d.ninc = (int)t.inc_idx.size();
  return EUVL_OK;
}

size_t solve_workspace_bytes(int N, int E, int ninc) {
  const size_t mat = sizeof(double2) * (size_t)N * N * E;
  const size_t vec = sizeof(double2) * (size_t)N * E;
  const size_t rhs = sizeof(double2) * (size_t)N * std::max(ninc, 1) * E;
  return 13 * mat + 8 * vec + 3 * rhs + sizeof(int) * (size_t)N * E + sizeof(double) * 8 * (size_t)E +
         sizeof(int) * (size_t)E + coupling_fft_bytes(E) + 64 * 1024;
}

// The EM solve of E = n_clips x n_off batch entries (ampS, header.h:1208-1466)
// on tables already on the device.  ctl.fixed: the doubling count and the
// unpivoted top solve are fixed ahead of time and checked on the
// device (ctl.d_flags); nothing is read back, so the call can be captured in a
// graph.  Otherwise the decisions are taken from norms read back to the host,
// with the pivoted fallback of the top solve.  Takes its matrices from the
// workspace (reserved by the caller).
int amp_s_core(euvl_handle h, char polar, const euvl_em_params* p, const SolveDev& t, int N, int lmax, int mmax,
               const double2* spec, int per_clip, int64_t spec_stride, int E, double2* Ax, SolveCtl& ctl) {
  const int B = E;
  const double k = p->k;
  const int nsx = 2 * p->lsmax_x + 1, nsy = 2 * p->lsmax_y + 1;
  const int ninc = t.ninc;
  const size_t mat = sizeof(double2) * (size_t)N * N * B;
  const size_t vec = sizeof(double2) * (size_t)N * B;
  const size_t rhs = sizeof(double2) * (size_t)N * std::max(ninc, 1) * B;
  ModeVecs v;
  v.kt = (double*)ws_take(h, vec);
  v.kxy2 = (double*)ws_take(h, vec);
  v.klm = (double2*)ws_take(h, vec);
  v.uu = (double2*)ws_take(h, vec);
  v.ub = (double2*)ws_take(h, vec);
  v.alru = (double2*)ws_take(h, vec);
  v.b2 = (double2*)ws_take(h, vec);
  double2* Mx[7];
  for (int i = 0; i < 7; ++i) Mx[i] = (i == 5 || i == 3) ? nullptr : (double2*)ws_take(h, i == 2 ? 2 * mat : mat);
  // C and F (Mx[2], Mx[3]) contiguous: cos_sinc runs their products as pairs
  if (Mx[2]) Mx[3] = Mx[2] + (int64_t)N * N * B;
  // the top-solve matrix with its right-hand sides appended: [V | E], N x (N + ninc)
  const int64_t ldv = (int64_t)N + ninc, sv = (int64_t)N * ldv;
  Mx[5] = (double2*)ws_take(h, sizeof(double2) * (size_t)sv * B);
  double2* Brhs = (double2*)ws_take(h, rhs);
  double2* FGc = (double2*)ws_take(h, rhs);
  int* ipiv = (int*)ws_take(h, sizeof(int) * (size_t)N * B);
  double* red = (double*)ws_take(h, sizeof(double) * 8 * (size_t)B);  // device-side check slots
  int* info = (int*)ws_take(h, sizeof(int) * (size_t)B);
  for (int i = 0; i < 7; ++i)
    if (!Mx[i] || !Brhs || !FGc || !ipiv || !red || !info || !v.b2) {
      set_error("euvl_amp_s: workspace carve failed");
      return EUVL_ERR_NOMEM;
    }
  if (ctl.fixed && !ctl.d_flags) {
    set_error("euvl_amp_s: a fixed-decision solve needs a flag array");
    return EUVL_ERR_ARG;
  }
  const int64_t sm = (int64_t)N * N;

  hipLaunchKernelGGL(mode_vectors_kernel, grid1((int64_t)N * B), dim3(256), 0, h->stream, N, B, polar, k, p->dx,
                     p->dy, t.lidx, t.midx, t.sx0, t.sy0, p->n_layers, t.leps, t.ld, v, (double*)nullptr,
                     (double*)nullptr);
  EUVL_LAUNCH_CHECK();
  const std::complex<double> ea(p->eabs_re, p->eabs_im);
  const std::complex<double> eai = 1.0 / ea, lea = std::log(ea);
  double2* D = Mx[0];
  double2* G = Mx[1];
  hipLaunchKernelGGL(toeplitz_kernel, grid1(sm * B), dim3(256), 0, h->stream, N, B, polar, k, p->dx, p->dy, t.lidx,
                     t.midx, lmax, mmax, spec, per_clip, spec_stride, make_double2(ea.real(), ea.imag()),
                     make_double2(eai.real(), eai.imag()), make_double2(lea.real(), lea.imag()), v, D, G);
  EUVL_LAUNCH_CHECK();

  // Where the order grid fits a workgroup's LDS, the products by D (D is
  // Toeplitz in the order differences up to diagonal scalings) and the
  // coupling product H G run as FFT correlations with the clip's spectrum;
  // one spectrum buffer serves both (the coupling product refills it).
  FftGrid fg;
  const bool grid_fits = fft_grid_for(lmax, mmax, fg);
  double2* Ssp = grid_fits ? (double2*)ws_take(h, coupling_fft_bytes(B)) : nullptr;
  if (grid_fits && !Ssp) {
    set_error("euvl_amp_s: workspace carve failed");
    return EUVL_ERR_NOMEM;
  }
  XdOp xd;
  const bool use_xd = grid_fits && xd_fft_fits(N) && N >= fft_min_n(false);
  if (use_xd) {
    const double c = 2.0 * PI / (polar == 'X' ? p->dx : p->dy);
    const std::complex<double> a_p = k * k * (1.0 - ea), b_u = c * lea, a_x = k * k * ea;
    xd.g = fg;
    xd.N = N;
    xd.lmax = lmax;
    xd.lidx = t.lidx;
    xd.midx = t.midx;
    xd.uidx = polar == 'X' ? t.lidx : t.midx;
    xd.kt = v.kt;
    xd.kxy2 = v.kxy2;
    xd.a_p = make_double2(a_p.real(), a_p.imag());
    xd.b_u = make_double2(b_u.real(), b_u.imag());
    xd.a_x = make_double2(a_x.real(), a_x.imag());
    xd.S = Ssp;
    EUVL_TRY(xd_fft_spectra(h, fg, lmax, mmax, spec, per_clip, spec_stride, B, Ssp));
  }

  // C = cos(d sqrt D), F = sin(d sqrt D)/sqrt D, G2 = D F (6 scratch matrices from the workspace)
  double2* C = Mx[2];
  double2* Fs = Mx[3];
  double2* G2c = Mx[4];
  int ndbl = 0;
  EUVL_TRY(cos_sinc(h, N, B, D, p->dabs, C, Fs, G2c, &ndbl, ctl.fixed ? ctl.doublings : -1, red, ctl.d_flags,
                    &ctl.alpha, use_xd ? &xd : nullptr));

  // The top solve.  With M = I - G (G = diag(kt) sigma^ diag(kt) / k^2),
  // S = diag(b2 (uu + ub) / (i k)) and L2 = diag(alru (uu - ub)), the layer
  // coupling gives W1 = M^-1 S, U0 = P W1 + Q and R = C L2 - i G2 W1, where
  // P = diag(klm) C - i G2 and Q = (C - i diag(klm) F) L2; the amplitudes need
  // only R U0^-1 E (E: the ninc incidence columns).  With Z = W1 X the system
  // U0 X = E becomes V Z = E with V = P + H M, H = Q S^-1 = (C - i diag(klm) F)
  // diag(lam), lam = L2 S^-1, and R X = C diag(lam) M Z - i G2 Z.  So neither
  // M^-1 nor W1 is formed: one N x N product (H G) and three N x ninc products
  // replace a series for M^-1 and the two N x N products of the reference's
  // T1 / T2 blocks (header.h:1053-1144, 1355-1391).
  PrecScope scope(h, EUVL_PREC_FP32_COUPLING);
  double2* H = Mx[6];
  double2* V = Mx[5];  // [V | E] with leading dimension N + ninc
  hipLaunchKernelGGL(top_hv_kernel, grid1(sm * B), dim3(256), 0, h->stream, N, B, k, C, Fs, G2c, H, V, ldv, sv, v);
  EUVL_LAUNCH_CHECK();
  // V = P + H - H G: by FFTs over the order grid when the grid fits a
  // workgroup's LDS (G is sigma^ Toeplitz in the order differences), else a GEMM
  if (grid_fits && N >= fft_min_n(true)) {
    EUVL_TRY(coupling_fft(h, fg, N, B, t.lidx, t.midx, lmax, mmax, spec, per_clip, spec_stride,
                          make_double2(eai.real(), eai.imag()), v.kt, k, H, V, ldv, sv, Ssp));
  } else {
    EUVL_TRY(zgemm(h, N, N, N, MONE, H, N, sm, G, N, sm, ONE, V, ldv, sv, ZERO, B));
  }
  double2* Vc = Mx[0];    // D is no longer needed: a copy of V for the residual check
  double2* Bres = Mx[6];  // nor H: the right-hand sides kept for the check
  const int64_t nsxy = (int64_t)nsx * nsy;
  EUVL_TRY(fill(h, nsxy * N * B, Ax, ZERO));
  int lu_path = 0;
  if (ninc > 0) {
    const int64_t sr = (int64_t)N * ninc;
    const double tol = (h->prec_mask & EUVL_PREC_FP32_LU) ? 1e-5 : 1e-12;
    auto singular = [&](bool report) -> int {
      int* hi = (int*)h->h_pinned;
      EUVL_HIP_CHECK(hipMemcpyAsync(hi, info, sizeof(int) * B, hipMemcpyDeviceToHost, h->stream));
      EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));
      for (int b = 0; b < B; ++b)
        if (hi[b] != 0) {
          if (report) set_error("euvl_amp_s: top-solve matrix singular (offset %d, column %d)", b, hi[b] - 1);
          return EUVL_ERR_SINGULAR;
        }
      return EUVL_OK;
    };
    auto copy2d = [&](int rows, int cols, const double2* src, int64_t lds, int64_t ss, double2* dst, int64_t ldd,
                      int64_t sd) -> int {
      hipLaunchKernelGGL(copy2d_kernel, grid1((int64_t)rows * cols * B), dim3(256), 0, h->stream, rows, cols, B, src,
                         lds, ss, dst, ldd, sd);
      EUVL_LAUNCH_CHECK();
      return EUVL_OK;
    };
    // V is well conditioned for these stacks and its elimination without row
    // interchanges is stable, so the block Gauss-Jordan solve (no pivot
    // searches) comes first; it is verified by the residual
    // ||V Z - E|| / (||V|| ||Z||) and redone with a partial-pivoting LU if that
    // is not at rounding level (adaptive), or flagged (fixed).
    EUVL_TRY(copy2d(N, N, V, ldv, sv, Vc, N, sm));
    hipLaunchKernelGGL(rhs_kernel, grid1(sr * B), dim3(256), 0, h->stream, N, ninc, B, t.inc_idx, V + N, ldv, sv);
    EUVL_LAUNCH_CHECK();
    EUVL_TRY(gj_solve(h, N, ninc, V, ldv, sv, B, info));
    bool ok = true;
    if (ctl.fixed)
      EUVL_TRY(flag_info(h, B, info, FLAG_ZERO_PIVOT, ctl.d_flags));
    else
      ok = singular(false) == EUVL_OK;
    if (ok) {
      EUVL_TRY(copy2d(N, ninc, V + N, ldv, sv, Brhs, ninc, sr));
      hipLaunchKernelGGL(rhs_kernel, grid1(sr * B), dim3(256), 0, h->stream, N, ninc, B, t.inc_idx, Bres,
                         (int64_t)ninc, sr);
      EUVL_LAUNCH_CHECK();
      {
        PrecScope fp64(h, 0);  // the check itself always in FP64
        EUVL_TRY(zgemm(h, N, ninc, N, ONE, Vc, N, sm, Brhs, ninc, sr, MONE, Bres, ninc, sr, ZERO, B));
      }
      double* dr = red + 3 * B;
      EUVL_TRY(maxabs(h, N, ninc, ninc, sr, B, Bres, dr));
      EUVL_TRY(maxabs(h, N, ninc, ninc, sr, B, Brhs, dr + B));
      EUVL_TRY(reduce(h, 1, N, B, Vc, dr + 2 * B));
      if (ctl.fixed) {
        EUVL_TRY(flag_check(h, 2, B, dr, dr + B, dr + 2 * B, 0.0, tol, FLAG_LU_RESIDUAL, ctl.d_flags));
      } else {
        EUVL_HIP_CHECK(hipMemcpyAsync(h->h_pinned, dr, sizeof(double) * 3 * B, hipMemcpyDeviceToHost, h->stream));
        EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));
        for (int b = 0; b < B; ++b) {
          const double res = h->h_pinned[b], xm = h->h_pinned[B + b], nrm = h->h_pinned[2 * B + b];
          if (!(res <= tol * nrm * xm)) ok = false;
        }
      }
    }
    if (!ok) {
      lu_path = 1;
      EUVL_TRY(copy2d(N, N, Vc, N, sm, V, ldv, sv));
      EUVL_TRY(zgetrf(h, N, V, ldv, sv, ipiv, B, info));
      EUVL_TRY(singular(true));
      hipLaunchKernelGGL(rhs_kernel, grid1(sr * B), dim3(256), 0, h->stream, N, ninc, B, t.inc_idx, Brhs,
                         (int64_t)ninc, sr);
      EUVL_LAUNCH_CHECK();
      EUVL_TRY(zgetrs(h, N, ninc, V, ldv, sv, ipiv, Brhs, ninc, sr, B));
    }
    // R X = C Y - i G2 Z with Y = diag(lam) (Z - G Z)  (Z in Brhs)
    EUVL_TRY(zgemm(h, N, ninc, N, ONE, G, N, sm, Brhs, ninc, sr, ZERO, FGc, ninc, sr, ZERO, B));
    hipLaunchKernelGGL(top_y_kernel, grid1(sr * B), dim3(256), 0, h->stream, N, ninc, B, k, Brhs, FGc, v);
    EUVL_LAUNCH_CHECK();
    EUVL_TRY(zgemm(h, N, ninc, N, ONE, C, N, sm, FGc, ninc, sr, ZERO, Bres, ninc, sr, ZERO, B));
    EUVL_TRY(zgemm(h, N, ninc, N, MI, G2c, N, sm, Brhs, ninc, sr, ONE, Bres, ninc, sr, ZERO, B));
    hipLaunchKernelGGL(ax_scatter_kernel, grid1(sr * B), dim3(256), 0, h->stream, N, ninc, B, t.inc_idx,
                       t.inc_slot, t.kz, Bres, v.klm, nsxy * N, Ax);
    EUVL_LAUNCH_CHECK();
  }
  ctl.doublings_used = ndbl;
  ctl.lu_path = lu_path;
  return EUVL_OK;
}

}  // namespace euvl

extern "C" int euvl_amp_s(euvl_handle h, char polar, const euvl_em_params* p, const int* h_lindex,
                          const int* h_mindex, int nrange, int lmax, int mmax, const double* d_spec, int n_off,
                          const double* h_sx0, const double* h_sy0, double* d_Ax, int* h_info) {
  using namespace euvl;
  EUVL_TRY(check_params(p));
  if (!h || !h_lindex || !h_mindex || nrange <= 0 || !d_spec || n_off <= 0 || n_off > 64 || !h_sx0 || !h_sy0 ||
      !d_Ax || (polar != 'X' && polar != 'Y')) {
    set_error("euvl_amp_s: bad argument");
    return EUVL_ERR_ARG;
  }
  for (int i = 0; i < nrange; ++i)
    if (std::abs(h_lindex[i]) > lmax || std::abs(h_mindex[i]) > mmax) {
      set_error("euvl_amp_s: order (%d,%d) outside lmax=%d mmax=%d", h_lindex[i], h_mindex[i], lmax, mmax);
      return EUVL_ERR_ARG;
    }
  const int N = nrange, B = n_off;
  SolveHost th;
  build_solve_tables(p, h_lindex, h_mindex, N, lmax, mmax, B, h_sx0, h_sy0, 1, th);
  const size_t tb = solve_tables_bytes(p, N, th);
  EUVL_TRY(ws_reserve(h, tb + solve_workspace_bytes(N, B, (int)th.inc_idx.size())));
  char* blk = (char*)ws_take(h, tb);
  SolveDev td;
  Upload up;
  EUVL_TRY(upload_solve_tables(h, p, h_lindex, h_mindex, N, th, up, blk, td));
  SolveCtl ctl;
  EUVL_TRY(amp_s_core(h, polar, p, td, N, lmax, mmax, (const double2*)d_spec, B, 0, B, (double2*)d_Ax, ctl));
  if (h_info) {
    for (int b = 0; b < B; ++b) {
      h_info[4 * b + 0] = ctl.lu_path;
      h_info[4 * b + 1] = ctl.doublings_used;
      h_info[4 * b + 2] = 0;
      h_info[4 * b + 3] = td.ninc;
    }
  }
  return EUVL_OK;
}
