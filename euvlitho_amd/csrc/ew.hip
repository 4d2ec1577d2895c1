// Element-wise and reduction kernels on batches of n x n complex matrices
// (contiguous, stride n*n per batch entry) used by the EM solve.
#include <algorithm>

#include "common.h"
#include "ew.h"

namespace {

__device__ inline int64_t gidx() { return (int64_t)blockIdx.x * blockDim.x + threadIdx.x; }

// The element-wise kernels run on grid (blocks, batch) with 32-bit indices
// inside one n x n matrix (n*n < 2^31, checked by the callers): the row and
// column of an element come from a 32-bit division, not a 64-bit one.

// The six Paterson-Stockmeyer blocks of the cos and sinc polynomials from one
// read of the powers X1..X4 (D, D^2, D^3, D^4), for coefficients c (cos) and
// f (sinc), 13 each:
//   H = c11 X3 + c10 X2 + c9 X1 + c8 I + c12 X4,  Q = c7 X3 + c6 X2 + c5 X1 + c4 I,
//   O = c3 X3 + c2 X2 + c1 X1 + c0 I   (and Hf, Qf, Of with f).
// Each element is read before any output is written at its index, so an
// output may share a buffer with an input (Qf with X2).
struct PsCoeffs {
  double c[13], f[13];
};
__global__ void ps_blocks_kernel(int n, const double2* X1, const double2* X2, const double2* X3, const double2* X4,
                                 PsCoeffs k, double2* Hc, double2* Qc, double2* Oc, double2* Hf, double2* Qf,
                                 double2* Of) {
  const uint32_t per = (uint32_t)n * (uint32_t)n;
  const int64_t base = (int64_t)blockIdx.y * per;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < per; r += gridDim.x * blockDim.x) {
    const int64_t e = base + r;
    const double2 x1 = X1[e], x2 = X2[e], x3 = X3[e], x4 = X4[e];
    const double id = (r % ((uint32_t)n + 1) == 0) ? 1.0 : 0.0;
    auto comb = [&](const double* c, int j) {  // c[j+3] X3 + c[j+2] X2 + c[j+1] X1 + c[j] I
      return make_double2(c[j + 3] * x3.x + c[j + 2] * x2.x + c[j + 1] * x1.x + c[j] * id,
                          c[j + 3] * x3.y + c[j + 2] * x2.y + c[j + 1] * x1.y);
    };
    double2 hc = comb(k.c, 8), hf = comb(k.f, 8);
    hc = make_double2(hc.x + k.c[12] * x4.x, hc.y + k.c[12] * x4.y);
    hf = make_double2(hf.x + k.f[12] * x4.x, hf.y + k.f[12] * x4.y);
    const double2 qc = comb(k.c, 4), qf = comb(k.f, 4), oc = comb(k.c, 0), of = comb(k.f, 0);
    Hc[e] = hc;
    Qc[e] = qc;
    Oc[e] = oc;
    Hf[e] = hf;
    Qf[e] = qf;
    Of[e] = of;
  }
}

// The two blocks of each polynomial in powers of X6 = D^6, from one read of
// X1..X6 (D .. D^6): O = c5 X5 + .. + c1 X1 + c0 I (written to Oc, Of) and
// T = c12 X6 + c11 X5 + .. + c7 X1 + c6 I (written to Tc, Tf), so that
// q = O + X6 T.  Each element is read before any output is written at its
// index, so Tc and Tf may be the buffers of X2 and X3.
__global__ void ps6_blocks_kernel(int n, const double2* X1, const double2* X2, const double2* X3, const double2* X4,
                                  const double2* X5, const double2* X6, PsCoeffs k, double2* Oc, double2* Tc,
                                  double2* Of, double2* Tf) {
  const uint32_t per = (uint32_t)n * (uint32_t)n;
  const int64_t base = (int64_t)blockIdx.y * per;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < per; r += gridDim.x * blockDim.x) {
    const int64_t e = base + r;
    const double2 x[6] = {X1[e], X2[e], X3[e], X4[e], X5[e], X6[e]};
    const double id = (r % ((uint32_t)n + 1) == 0) ? 1.0 : 0.0;
    auto comb = [&](const double* c, int j, int top) {  // c[j] I + sum_{t=1..top} c[j+t] X_t
      double re = c[j] * id, im = 0.0;
#pragma unroll
      for (int t = 1; t <= 6; ++t)
        if (t <= top) {
          re += c[j + t] * x[t - 1].x;
          im += c[j + t] * x[t - 1].y;
        }
      return make_double2(re, im);
    };
    const double2 oc = comb(k.c, 0, 5), tc = comb(k.c, 6, 6), of = comb(k.f, 0, 5), tf = comb(k.f, 6, 6);
    Oc[e] = oc;
    Tc[e] = tc;
    Of[e] = of;
    Tf[e] = tf;
  }
}

__global__ void diag_add_kernel(int n, int batch, double2* X, double2 g) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * batch) return;
  const int64_t b = e / n, i = e - b * n;
  double2* x = X + b * n * n + i * (n + 1);
  *x = cadd(*x, g);
}

// Per-batch reductions into red[b] (red zeroed beforehand for kinds 0 and 1).
// Non-negative doubles order like their bit patterns, so the maxima use an
// integer atomicMax; NaN maps above +inf and propagates as "too large".
__device__ inline void atomic_max_nonneg(double* addr, double v) {
  atomicMax((unsigned long long*)addr, (unsigned long long)__double_as_longlong(v));
}

__device__ inline double wave_max(double v) {  // NaN wins
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(v, off);
    v = (o > v || o != o) ? o : v;
  }
  return v;
}

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// max over a 256-thread block, then one atomic per block: thousands of
// atomics on one address serialise (~40 ns each)
__device__ inline void block_max_to(double* addr, double v) {
  __shared__ double sh[4];
  v = wave_max(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = sh[0];
    for (int w = 1; w < 4; ++w) m = (sh[w] > m || sh[w] != sh[w]) ? sh[w] : m;
    atomic_max_nonneg(addr, m);
  }
}

// kind 0: max |X - I| (elementwise |re| + |im|); grid (blocks, batch)
__global__ __launch_bounds__(256) void reduce_maxdev_kernel(int n, const double2* X, double* red) {
  const int b = blockIdx.y;
  const uint32_t per = (uint32_t)n * (uint32_t)n;
  X += (int64_t)b * per;
  double acc = 0.0;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < per; e += gridDim.x * blockDim.x) {
    double2 v = X[e];
    if (e % ((uint32_t)n + 1) == 0) v.x -= 1.0;
    const double a = fabs(v.x) + fabs(v.y);
    acc = (a > acc || a != a) ? a : acc;
  }
  block_max_to(red + b, acc);
}

// kind 1: infinity norm (max row sum of |x|), one wave per row; grid (ceil(n/4), batch)
__global__ __launch_bounds__(256) void reduce_infnorm_kernel(int n, const double2* X, double* red) {
  const int b = blockIdx.y;
  const int lane = threadIdx.x & 63;
  double best = 0.0;
  // one wave per row, rows strided over the grid's waves
  for (int row = blockIdx.x * 4 + (threadIdx.x >> 6); row < n; row += gridDim.x * 4) {
    const double2* x = X + (int64_t)b * n * n + (int64_t)row * n;
    double s = 0.0;
    // |x| as sqrt(re^2 + im^2): the norm only sizes the scaling steps, and
    // the entries are far from the range where hypot's rescaling matters
    for (int j = lane; j < n; j += 64) s += sqrt(fma(x[j].x, x[j].x, x[j].y * x[j].y));
    s = wave_sum(s);
    best = (s > best || s != s) ? s : best;
  }
  block_max_to(red + b, best);
}

// kind 2: mean of Re diag, one block per batch entry
__global__ __launch_bounds__(256) void reduce_meandiag_kernel(int n, const double2* X, double* red) {
  const int b = blockIdx.x;
  X += (int64_t)b * n * n;
  __shared__ double sh[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += X[(int64_t)i * n + i].x;
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) red[b] = sh[0] / n;
}

// max over an m x ncols block (leading dim ld, batch stride sb) of |re| + |im|; grid (blocks, batch)
__global__ __launch_bounds__(256) void reduce_maxabs_kernel(int m, int ncols, int64_t ld, int64_t sb,
                                                            const double2* X, double* red) {
  const int b = blockIdx.y;
  X += (int64_t)b * sb;
  double acc = 0.0;
  // rows over the grid, columns over the threads: no per-element division
  for (int r = blockIdx.x; r < m; r += gridDim.x)
    for (int c = threadIdx.x; c < ncols; c += blockDim.x) {
      const double2 v = X[(int64_t)r * ld + c];
      const double a = fabs(v.x) + fabs(v.y);
      acc = (a > acc || a != a) ? a : acc;
    }
  block_max_to(red + b, acc);
}

__global__ void zero_kernel(double* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0.0;
}

__global__ void copy_kernel(int64_t total, double2* out, const double2* X) {
  for (int64_t e = gidx(); e < total; e += (int64_t)gridDim.x * blockDim.x) out[e] = X[e];
}

inline dim3 ew_grid(int64_t total) {
  int64_t b = (total + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}

// (blocks over one n x n matrix, batch); about 4096 workgroups in all
inline dim3 ew_grid2(int64_t per, int batch) {
  int64_t b = (per + 255) / 256, cap = std::max<int64_t>(1, 4096 / batch);
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return dim3((unsigned)b, (unsigned)batch);
}

int check_ew(int n, int batch) {
  if (n <= 0 || (int64_t)n * n >= ((int64_t)1 << 31) || batch <= 0 || batch > 65535) {
    euvl::set_error("element-wise kernel: n = %d, batch = %d out of range", n, batch);
    return EUVL_ERR_ARG;
  }
  return EUVL_OK;
}

// flags[b] |= bit where a device-side check of a fixed decision fails:
//   kind 0 (doubling count): s * max(cbrt(a[b]), c[b]^(1/4)) > theta, or non-finite
//   kind 2 (LU residual):    !(a[b] <= theta * c[b] * d[b])
__global__ void flag_kernel(int kind, int batch, const double* a, const double* c, const double* d, double s,
                            double theta, int bit, int* flags) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  bool bad;
  if (kind == 0) {
    const double al = fmax(cbrt(a[b]), sqrt(sqrt(c[b])));
    bad = !(s * al <= theta);
  } else {
    bad = !(a[b] <= theta * c[b] * d[b]);
  }
  if (bad) flags[b] |= bit;
}

// flags[b] |= bit where info[b] != 0 (a zero pivot)
__global__ void flag_info_kernel(int batch, const int* info, int bit, int* flags) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < batch && info[b] != 0) flags[b] |= bit;
}

__global__ void zero_int_kernel2(int* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

}  // namespace

namespace euvl {

int ew_ps_blocks(euvl_handle h, int n, int batch, const double2* X1, const double2* X2, const double2* X3,
                 const double2* X4, const double* c, const double* f, double2* Hc, double2* Qc, double2* Oc,
                 double2* Hf, double2* Qf, double2* Of) {
  EUVL_TRY(check_ew(n, batch));
  PsCoeffs k;
  for (int j = 0; j < 13; ++j) {
    k.c[j] = c[j];
    k.f[j] = f[j];
  }
  hipLaunchKernelGGL(ps_blocks_kernel, ew_grid2((int64_t)n * n, batch), dim3(256), 0, h->stream, n, X1, X2, X3, X4,
                     k, Hc, Qc, Oc, Hf, Qf, Of);
  EUVL_LAUNCH_CHECK();
  return EUVL_OK;
}

int ew_ps6_blocks(euvl_handle h, int n, int batch, const double2* const* X, const double* c, const double* f,
                  double2* Oc, double2* Tc, double2* Of, double2* Tf) {
  EUVL_TRY(check_ew(n, batch));
  PsCoeffs k;
  for (int j = 0; j < 13; ++j) {
    k.c[j] = c[j];
    k.f[j] = f[j];
  }
  hipLaunchKernelGGL(ps6_blocks_kernel, ew_grid2((int64_t)n * n, batch), dim3(256), 0, h->stream, n, X[0], X[1],
                     X[2], X[3], X[4], X[5], k, Oc, Tc, Of, Tf);
  EUVL_LAUNCH_CHECK();
  return EUVL_OK;
}

int ew_diag_add(euvl_handle h, int n, int batch, double2* X, double2 g) {
  EUVL_TRY(check_ew(n, batch));
  const int64_t total = (int64_t)n * batch;
  hipLaunchKernelGGL(diag_add_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, h->stream, n, batch, X, g);
  EUVL_LAUNCH_CHECK();
  return EUVL_OK;
}

int ew_copy(euvl_handle h, int64_t count, double2* out, const double2* X) {
  hipLaunchKernelGGL(copy_kernel, ew_grid(count), dim3(256), 0, h->stream, count, out, X);
  EUVL_LAUNCH_CHECK();
  return EUVL_OK;
}

int reduce(euvl_handle h, int kind, int n, int batch, const double2* X, double* d_red) {
  if (kind == 2) {
    hipLaunchKernelGGL(reduce_meandiag_kernel, dim3(batch), dim3(256), 0, h->stream, n, X, d_red);
    EUVL_LAUNCH_CHECK();
    return EUVL_OK;
  }
  hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(256), 0, h->stream, d_red, batch);
  EUVL_LAUNCH_CHECK();
  if (kind == 0) {
    EUVL_TRY(check_ew(n, batch));
    int64_t blocks = ((int64_t)n * n + 255) / 256;
    if (blocks > 256) blocks = 256;
    hipLaunchKernelGGL(reduce_maxdev_kernel, dim3((unsigned)blocks, batch), dim3(256), 0, h->stream, n, X, d_red);
  } else {
    hipLaunchKernelGGL(reduce_infnorm_kernel, dim3((unsigned)std::min(256, (n + 3) / 4), batch), dim3(256), 0, h->stream, n, X, d_red);
  }
  EUVL_LAUNCH_CHECK();
  return EUVL_OK;
}

int flag_check(euvl_handle h, int kind, int batch, const double* a, const double* c, const double* d, double s,
               double theta, int bit, int* flags) {
  hipLaunchKernelGGL(flag_kernel, dim3((batch + 255) / 256), dim3(256), 0, h->stream, kind, batch, a, c, d, s, theta,
                     bit, flags);
  EUVL_LAUNCH_CHECK();
  return EUVL_OK;
}

int flag_info(euvl_handle h, int batch, const int* info, int bit, int* flags) {
  hipLaunchKernelGGL(flag_info_kernel, dim3((batch + 255) / 256), dim3(256), 0, h->stream, batch, info, bit, flags);
  EUVL_LAUNCH_CHECK();
  return EUVL_OK;
}

int zero_ints(euvl_handle h, int count, int* p) {
  if (count <= 0) return EUVL_OK;
  hipLaunchKernelGGL(zero_int_kernel2, dim3(std::min(1024, (count + 255) / 256)), dim3(256), 0, h->stream, p, count);
  EUVL_LAUNCH_CHECK();
  return EUVL_OK;
}

int maxabs(euvl_handle h, int m, int ncols, int64_t ld, int64_t sb, int batch, const double2* X, double* d_slot) {
  hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(256), 0, h->stream, d_slot, batch);
  EUVL_LAUNCH_CHECK();
  int64_t blocks = m;
  if (blocks > 256) blocks = 256;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(reduce_maxabs_kernel, dim3((unsigned)blocks, batch), dim3(256), 0, h->stream, m, ncols, ld, sb, X,
                     d_slot);
  EUVL_LAUNCH_CHECK();
  return EUVL_OK;
}

}  // namespace euvl
