// Batched complex LU with partial pivoting (right-looking, panel width 32)
// and the matching triangular solves.  Panels are factored by one workgroup
// per matrix straight in global memory (all waves of a workgroup share one
// CU, so __syncthreads orders their global reads and writes); the trailing
// update is the MFMA GEMM of zgemm.hip.
#include "ew.h"

namespace {

constexpr int NB = 32;
constexpr int PT = 256;  // panel threads

__device__ inline double cabs1(double2 a) { return fabs(a.x) + fabs(a.y); }

__global__ void zero_int_kernel(int* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// Factor columns [kb, kb+nbc) of A[b] (rows kb..n-1) in place; writes ipiv.
__global__ __launch_bounds__(PT) void getrf_panel_kernel(int n, double2* A, int64_t lda, int64_t sa, int* ipiv,
                                                         int kb, int nbc, int* info) {
  __shared__ double s_val[PT / 64];
  __shared__ int s_idx[PT / 64];
  __shared__ int s_piv;
  const int b = blockIdx.x;
  A += (int64_t)b * sa;
  ipiv += (int64_t)b * n;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;

  for (int j = 0; j < nbc; ++j) {
    const int c = kb + j;
    // pivot search: first row of largest |re|+|im| (LAPACK izamax rule)
    double best = -1.0;
    int bidx = n;
    for (int r = c + tid; r < n; r += PT) {
      double v = cabs1(A[(int64_t)r * lda + c]);
      if (v > best) { best = v; bidx = r; }
    }
    for (int off = 32; off > 0; off >>= 1) {
      double ov = __shfl_xor(best, off);
      int oi = __shfl_xor(bidx, off);
      if (ov > best || (ov == best && oi < bidx)) { best = ov; bidx = oi; }
    }
    if (lane == 0) { s_val[wave] = best; s_idx[wave] = bidx; }
    __syncthreads();
    if (tid == 0) {
      double bv = s_val[0];
      int bi = s_idx[0];
      for (int w = 1; w < PT / 64; ++w)
        if (s_val[w] > bv || (s_val[w] == bv && s_idx[w] < bi)) { bv = s_val[w]; bi = s_idx[w]; }
      if (bi >= n) bi = c;
      s_piv = bi;
      ipiv[c] = bi;
      if (bv == 0.0 && info[b] == 0) info[b] = c + 1;
    }
    __syncthreads();
    const int piv = s_piv;
    if (piv != c) {
      for (int q = kb + tid; q < kb + nbc; q += PT) {
        double2 t = A[(int64_t)c * lda + q];
        A[(int64_t)c * lda + q] = A[(int64_t)piv * lda + q];
        A[(int64_t)piv * lda + q] = t;
      }
    }
    __syncthreads();
    const double2 d = A[(int64_t)c * lda + c];
    const bool nz = (d.x != 0.0 || d.y != 0.0);
    const double2 dinv = nz ? cdiv(make_double2(1.0, 0.0), d) : make_double2(0.0, 0.0);
    for (int r = c + 1 + tid; r < n; r += PT) {
      double2* row = A + (int64_t)r * lda;
      double2 l = nz ? cmul(row[c], dinv) : row[c];
      row[c] = l;
      if (nz) {
        const double2* prow = A + (int64_t)c * lda;
        for (int q = c + 1; q < kb + nbc; ++q) row[q] = csub(row[q], cmul(l, prow[q]));
      }
    }
    __syncthreads();
  }
}

// Apply the interchanges ipiv[k1..k2) to every column outside [skip0, skip1).
__global__ void laswp_kernel(int ncols, double2* A, int64_t lda, int64_t sa, const int* ipiv, int64_t spiv,
                             int k1, int k2, int skip0, int skip1) {
  const int b = blockIdx.y;
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ncols) return;
  if (j >= skip0 && j < skip1) return;
  A += (int64_t)b * sa;
  ipiv += (int64_t)b * spiv;
  for (int i = k1; i < k2; ++i) {
    int p = ipiv[i];
    if (p != i) {
      double2 t = A[(int64_t)i * lda + j];
      A[(int64_t)i * lda + j] = A[(int64_t)p * lda + j];
      A[(int64_t)p * lda + j] = t;
    }
  }
}

// X = T^-1 X for an nbc x nbc triangular block T (unit lower, or non-unit
// upper when UPPER), X nbc x ncols; one thread per column.
template <bool UPPER>
__global__ __launch_bounds__(256) void trsm_kernel(int nbc, const double2* T, int64_t ldt, int64_t st, double2* X,
                                                   int64_t ldx, int64_t sx, int ncols) {
  __shared__ double2 sT[NB][NB + 1];
  const int b = blockIdx.y;
  T += (int64_t)b * st;
  X += (int64_t)b * sx;
  for (int e = threadIdx.x; e < nbc * nbc; e += blockDim.x) {
    int r = e / nbc, c = e % nbc;
    sT[r][c] = T[(int64_t)r * ldt + c];
  }
  __syncthreads();
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ncols) return;
  double2 x[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) x[i] = (i < nbc) ? X[(int64_t)i * ldx + j] : make_double2(0.0, 0.0);
  if (!UPPER) {
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      if (i < nbc) {
        double2 s = x[i];
#pragma unroll
        for (int p = 0; p < i; ++p) s = csub(s, cmul(sT[i][p], x[p]));
        x[i] = s;
      }
    }
  } else {
#pragma unroll
    for (int i = NB - 1; i >= 0; --i) {
      if (i < nbc) {
        double2 s = x[i];
#pragma unroll
        for (int p = i + 1; p < NB; ++p)
          if (p < nbc) s = csub(s, cmul(sT[i][p], x[p]));
        x[i] = cdiv(s, sT[i][i]);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < NB; ++i)
    if (i < nbc) X[(int64_t)i * ldx + j] = x[i];
}

}  // namespace

namespace euvl {

int zgetrf(euvl_handle h, int n, double2* A, int64_t lda, int64_t sa, int* ipiv, int batch, int* d_info) {
  const double2 one = make_double2(1.0, 0.0), mone = make_double2(-1.0, 0.0), zero = make_double2(0.0, 0.0);
  hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(256), 0, h->stream, d_info, batch);
  EUVL_LAUNCH_CHECK();
  for (int kb = 0; kb < n; kb += NB) {
    const int nbc = (n - kb < NB) ? n - kb : NB;
    hipLaunchKernelGGL(getrf_panel_kernel, dim3(batch), dim3(PT), 0, h->stream, n, A, lda, sa, ipiv, kb, nbc,
                       d_info);
    EUVL_LAUNCH_CHECK();
    hipLaunchKernelGGL(laswp_kernel, dim3((n + 255) / 256, batch), dim3(256), 0, h->stream, n, A, lda, sa, ipiv,
                       (int64_t)n, kb, kb + nbc, kb, kb + nbc);
    EUVL_LAUNCH_CHECK();
    const int rest = n - kb - nbc;
    if (rest > 0) {
      double2* A11 = A + (int64_t)kb * lda + kb;
      double2* A12 = A11 + nbc;
      double2* A21 = A11 + (int64_t)nbc * lda;
      double2* A22 = A21 + nbc;
      hipLaunchKernelGGL(trsm_kernel<false>, dim3((rest + 255) / 256, batch), dim3(256), 0, h->stream, nbc, A11,
                         lda, sa, A12, lda, sa, rest);
      EUVL_LAUNCH_CHECK();
      EUVL_TRY(zgemm(h, rest, rest, nbc, mone, A21, lda, sa, A12, lda, sa, one, A22, lda, sa, zero, batch));
    }
  }
  return EUVL_OK;
}

int zgetrs(euvl_handle h, int n, int nrhs, const double2* LU, int64_t lda, int64_t sa, const int* ipiv,
           double2* B, int64_t ldb, int64_t sb, int batch) {
  const double2 one = make_double2(1.0, 0.0), mone = make_double2(-1.0, 0.0), zero = make_double2(0.0, 0.0);
  hipLaunchKernelGGL(laswp_kernel, dim3((nrhs + 255) / 256, batch), dim3(256), 0, h->stream, nrhs, B, ldb, sb,
                     ipiv, (int64_t)n, 0, n, 0, 0);
  EUVL_LAUNCH_CHECK();
  for (int kb = 0; kb < n; kb += NB) {
    const int nbc = (n - kb < NB) ? n - kb : NB;
    const double2* L11 = LU + (int64_t)kb * lda + kb;
    double2* Bk = B + (int64_t)kb * ldb;
    hipLaunchKernelGGL(trsm_kernel<false>, dim3((nrhs + 255) / 256, batch), dim3(256), 0, h->stream, nbc, L11,
                       lda, sa, Bk, ldb, sb, nrhs);
    EUVL_LAUNCH_CHECK();
    const int rest = n - kb - nbc;
    if (rest > 0)
      EUVL_TRY(zgemm(h, rest, nrhs, nbc, mone, L11 + (int64_t)nbc * lda, lda, sa, Bk, ldb, sb, one,
                     Bk + (int64_t)nbc * ldb, ldb, sb, zero, batch));
  }
  const int last = ((n - 1) / NB) * NB;
  for (int kb = last; kb >= 0; kb -= NB) {
    const int nbc = (n - kb < NB) ? n - kb : NB;
    const double2* U11 = LU + (int64_t)kb * lda + kb;
    double2* Bk = B + (int64_t)kb * ldb;
    hipLaunchKernelGGL(trsm_kernel<true>, dim3((nrhs + 255) / 256, batch), dim3(256), 0, h->stream, nbc, U11,
                       lda, sa, Bk, ldb, sb, nrhs);
    EUVL_LAUNCH_CHECK();
    if (kb > 0)
      EUVL_TRY(zgemm(h, kb, nrhs, nbc, mone, LU + kb, lda, sa, Bk, ldb, sb, one, B, ldb, sb, zero, batch));
  }
  return EUVL_OK;
}

}  // namespace euvl

extern "C" int euvl_zgetrf_batched(euvl_handle h, int n, double* d_A, int64_t lda, int64_t sa, int* d_ipiv,
                                   int batch) {
  if (!h || n < 0 || batch < 0 || lda < n) {
    euvl::set_error("euvl_zgetrf_batched: bad argument");
    return EUVL_ERR_ARG;
  }
  if (n == 0 || batch == 0) return EUVL_OK;
  int* d_info = (int*)h->d_red;
  if (batch > 1024) {
    euvl::set_error("euvl_zgetrf_batched: batch > 1024");
    return EUVL_ERR_ARG;
  }
  EUVL_TRY(euvl::zgetrf(h, n, (double2*)d_A, lda, sa, d_ipiv, batch, d_info));
  int* hi = (int*)h->h_pinned;
  EUVL_HIP_CHECK(hipMemcpyAsync(hi, d_info, sizeof(int) * batch, hipMemcpyDeviceToHost, h->stream));
  EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));
  for (int b = 0; b < batch; ++b)
    if (hi[b] != 0) {
      euvl::set_error("euvl_zgetrf_batched: matrix %d has a zero pivot at column %d", b, hi[b] - 1);
      return EUVL_ERR_SINGULAR;
    }
  return EUVL_OK;
}

extern "C" int euvl_zgetrs_batched(euvl_handle h, int n, int nrhs, const double* d_LU, int64_t lda, int64_t sa,
                                   const int* d_ipiv, double* d_B, int64_t ldb, int64_t sb, int batch) {
  if (!h || n < 0 || nrhs < 0 || batch < 0 || lda < n || ldb < nrhs) {
    euvl::set_error("euvl_zgetrs_batched: bad argument");
    return EUVL_ERR_ARG;
  }
  if (n == 0 || nrhs == 0 || batch == 0) return EUVL_OK;
  return euvl::zgetrs(h, n, nrhs, (const double2*)d_LU, lda, sa, d_ipiv, (double2*)d_B, ldb, sb, batch);
}
