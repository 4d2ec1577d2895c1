// Batched complex LU with partial pivoting (right-looking, panel width 32)
// and the matching triangular solves.  Panels are factored by one workgroup
// per matrix straight in global memory (all waves of a workgroup share one
// CU, so __syncthreads orders their global reads and writes); the trailing
// update is the MFMA GEMM of zgemm.hip.
#include "ew.h"

namespace {

constexpr int NB = 32;
constexpr int PT = 1024;  // panel threads

__device__ inline double cabs1(double2 a) { return fabs(a.x) + fabs(a.y); }

__global__ void zero_int_kernel(int* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// Copy the panel A[kb:n, kb:kb+nbc] of every batch entry to Pt[b][q][r]
// (column q contiguous over the m = n-kb panel rows) and back.
__global__ void panel_in_kernel(int n, const double2* A, int64_t lda, int64_t sa, int kb, int nbc, double2* Pt) {
  const int b = blockIdx.y;
  const int m = n - kb;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  const double2* row = A + (int64_t)b * sa + (int64_t)(kb + r) * lda + kb;
  double2* p = Pt + (int64_t)b * NB * n;
  for (int q = 0; q < nbc; ++q) p[(int64_t)q * n + r] = row[q];
}

__global__ void panel_out_kernel(int n, double2* A, int64_t lda, int64_t sa, int kb, int nbc, const double2* Pt) {
  const int b = blockIdx.y;
  const int m = n - kb;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  double2* row = A + (int64_t)b * sa + (int64_t)(kb + r) * lda + kb;
  const double2* p = Pt + (int64_t)b * NB * n;
  for (int q = 0; q < nbc; ++q) row[q] = p[(int64_t)q * n + r];
}

// Factor the transposed panel Pt[b] (nbc columns of m = n-kb rows, column
// stride n) with partial pivoting; one workgroup per batch entry.  Column
// accesses are contiguous across threads.  Writes ipiv[kb + j] (absolute rows).
__global__ __launch_bounds__(PT) void getrf_panel_kernel(int n, double2* Pt, int* ipiv, int kb, int nbc, int* info) {
  __shared__ double s_val[PT / 64];
  __shared__ int s_idx[PT / 64];
  __shared__ int s_piv;
  __shared__ double2 s_prow[NB];
  const int b = blockIdx.x;
  const int m = n - kb;
  double2* P = Pt + (int64_t)b * NB * n;
  ipiv += (int64_t)b * n;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;

  for (int j = 0; j < nbc; ++j) {
    double2* col = P + (int64_t)j * n;
    // pivot search: first row of largest |re|+|im| (LAPACK izamax rule)
    double best = -1.0;
    int bidx = m;
    for (int r = j + tid; r < m; r += PT) {
      const double v = cabs1(col[r]);
      if (v > best) { best = v; bidx = r; }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(best, off);
      const int oi = __shfl_xor(bidx, off);
      if (ov > best || (ov == best && oi < bidx)) { best = ov; bidx = oi; }
    }
    if (lane == 0) { s_val[wave] = best; s_idx[wave] = bidx; }
    __syncthreads();
    if (tid < 64) {
      double bv = (tid < PT / 64) ? s_val[tid] : -2.0;
      int bi = (tid < PT / 64) ? s_idx[tid] : m;
      for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(bv, off);
        const int oi = __shfl_xor(bi, off);
        if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
      }
      if (tid == 0) {
        if (bi >= m) bi = j;
        s_piv = bi;
        ipiv[kb + j] = kb + bi;
        if (bv == 0.0 && info[b] == 0) info[b] = kb + j + 1;
      }
    }
    __syncthreads();
    const int piv = s_piv;
    // swap rows j and piv across the panel columns and stage the pivot row
    if (tid < nbc) {
      double2* cq = P + (int64_t)tid * n;
      const double2 t = cq[j];
      const double2 u = cq[piv];
      cq[j] = u;
      if (piv != j) cq[piv] = t;
      s_prow[tid] = u;
    }
    __syncthreads();
    const double2 d = s_prow[j];
    const bool nz = (d.x != 0.0 || d.y != 0.0);
    const double2 dinv = nz ? cdiv(make_double2(1.0, 0.0), d) : make_double2(0.0, 0.0);
    for (int r = j + 1 + tid; r < m; r += PT) {
      const double2 l = nz ? cmul(col[r], dinv) : col[r];
      col[r] = l;
      if (nz)
        for (int q = j + 1; q < nbc; ++q) {
          double2* cq = P + (int64_t)q * n;
          cq[r] = csub(cq[r], cmul(l, s_prow[q]));
        }
    }
    __syncthreads();
  }
}

// Register-resident panel factorisation for panels of at most RPT * RT rows.
// The nbc <= 32 columns are processed in leaves of LW: a leaf lives in
// registers (row r of the panel is held by thread r % 512, slot r / 512), is
// factored column by column with partial pivoting (pivot rows exchanged
// through LDS, the swap mirrored on the panel's other columns in global
// memory), and then the columns to its right are updated in one pass:
// U12 = L11^-1 A12 on the 8 pivot rows and A22 -= L21 U12 on the rest.
constexpr int RT = 512;  // threads of the register-resident panel kernel
constexpr int RPT = 4;   // rows per thread (panel height <= 2048)
constexpr int LW = 4;    // leaf width (RPT x LW complex per thread stay in registers)

__global__ __launch_bounds__(RT) void getrf_panel_reg_kernel(int n, double2* Pt, int* ipiv, int kb, int nbc,
                                                             int* info) {
  __shared__ double s_val[RT / 64];
  __shared__ int s_idx[RT / 64];
  __shared__ int s_piv;
  __shared__ double2 s_row[2][LW];
  __shared__ double2 s_l11[LW][LW];
  __shared__ double2 s_u[LW][NB];
  __shared__ int s_pivs[LW];
  const int b = blockIdx.x;
  const int m = n - kb;
  double2* P = Pt + (int64_t)b * NB * n;
  ipiv += (int64_t)b * n;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;

  for (int c0 = 0; c0 < nbc; c0 += LW) {
    const int lw = (nbc - c0 < LW) ? nbc - c0 : LW;
    double2 x[RPT][LW];
#pragma unroll
    for (int s = 0; s < RPT; ++s) {
      const int r = tid + s * RT;
#pragma unroll
      for (int q = 0; q < LW; ++q)
        x[s][q] = (r < m && q < lw) ? P[(int64_t)(c0 + q) * n + r] : make_double2(0.0, 0.0);
    }
    for (int jj = 0; jj < lw; ++jj) {
      const int j = c0 + jj;
      // pivot search over rows >= j (first of the largest |re|+|im|)
      double best = -1.0;
      int bidx = m;
#pragma unroll
      for (int s = 0; s < RPT; ++s) {
        const int r = tid + s * RT;
        double v = -1.0;
#pragma unroll
        for (int q = 0; q < LW; ++q)
          if (q == jj) v = cabs1(x[s][q]);
        if (r >= j && r < m && v > best) { best = v; bidx = r; }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(best, off);
        const int oi = __shfl_xor(bidx, off);
        if (ov > best || (ov == best && oi < bidx)) { best = ov; bidx = oi; }
      }
      if (lane == 0) { s_val[wave] = best; s_idx[wave] = bidx; }
      __syncthreads();
      if (tid < 64) {
        double bv = (tid < RT / 64) ? s_val[tid] : -2.0;
        int bi = (tid < RT / 64) ? s_idx[tid] : m;
        for (int off = 32; off > 0; off >>= 1) {
          const double ov = __shfl_xor(bv, off);
          const int oi = __shfl_xor(bi, off);
          if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
        }
        if (tid == 0) {
          if (bi >= m) bi = j;
          s_piv = bi;
          ipiv[kb + j] = kb + bi;
          if (bv == 0.0 && info[b] == 0) info[b] = kb + j + 1;
        }
      }
      __syncthreads();
      const int piv = s_piv;
      // exchange leaf rows j and piv through LDS
      const int oj = j % RT, sj = j / RT, op = piv % RT, sp = piv / RT;
#pragma unroll
      for (int s = 0; s < RPT; ++s) {
        if (tid == oj && s == sj)
#pragma unroll
          for (int q = 0; q < LW; ++q) s_row[0][q] = x[s][q];
        if (tid == op && s == sp)
#pragma unroll
          for (int q = 0; q < LW; ++q) s_row[1][q] = x[s][q];
      }
      if (tid == 0) s_pivs[jj] = piv;
      __syncthreads();
#pragma unroll
      for (int s = 0; s < RPT; ++s) {
        if (tid == oj && s == sj)
#pragma unroll
          for (int q = 0; q < LW; ++q) x[s][q] = s_row[1][q];
        if (piv != j && tid == op && s == sp)
#pragma unroll
          for (int q = 0; q < LW; ++q) x[s][q] = s_row[0][q];
      }
      // pivot row = old row piv (s_row[1]); eliminate below it
      const double2 d = s_row[1][jj];
      const bool nz = (d.x != 0.0 || d.y != 0.0);
      const double2 dinv = nz ? cdiv(make_double2(1.0, 0.0), d) : make_double2(0.0, 0.0);
#pragma unroll
      for (int s = 0; s < RPT; ++s) {
        const int r = tid + s * RT;
        if (r > j && r < m && nz) {
          double2 l = make_double2(0.0, 0.0);
#pragma unroll
          for (int q = 0; q < LW; ++q)
            if (q == jj) {
              l = cmul(x[s][q], dinv);
              x[s][q] = l;
            }
#pragma unroll
          for (int q = 0; q < LW; ++q)
            if (q > jj && q < lw) x[s][q] = csub(x[s][q], cmul(l, s_row[1][q]));
        }
      }
      __syncthreads();
    }
    // write the leaf back and stage L11 (rows c0..c0+lw-1 of the leaf)
#pragma unroll
    for (int s = 0; s < RPT; ++s) {
      const int r = tid + s * RT;
      if (r < m)
#pragma unroll
        for (int q = 0; q < LW; ++q)
          if (q < lw) P[(int64_t)(c0 + q) * n + r] = x[s][q];
      if (r >= c0 && r < c0 + lw)
#pragma unroll
        for (int q = 0; q < LW; ++q) s_l11[r - c0][q] = x[s][q];
    }
    // apply this leaf's interchanges, in order, to the panel's other columns
    if (tid < nbc && (tid < c0 || tid >= c0 + lw)) {
      double2* cq = P + (int64_t)tid * n;
      for (int jj = 0; jj < lw; ++jj) {
        const int j = c0 + jj, piv = s_pivs[jj];
        if (piv != j) {
          const double2 t = cq[j];
          cq[j] = cq[piv];
          cq[piv] = t;
        }
      }
    }
    const int c1 = c0 + lw;
    const int nr = nbc - c1;  // columns right of the leaf
    if (nr <= 0) break;
    __syncthreads();
    // U12 = L11^-1 A12 on the pivot rows, one thread per right column
    if (tid < nr) {
      double2* cq = P + (int64_t)(c1 + tid) * n;
      double2 u[LW];
#pragma unroll
      for (int i = 0; i < LW; ++i) {
        if (i < lw) {
          double2 v = cq[c0 + i];
#pragma unroll
          for (int p = 0; p < LW; ++p)
            if (p < i) v = csub(v, cmul(s_l11[i][p], u[p]));
          u[i] = v;
          cq[c0 + i] = v;
          s_u[i][tid] = v;
        } else {
          u[i] = make_double2(0.0, 0.0);
        }
      }
    }
    __syncthreads();
    // A22 -= L21 U12 for the rows below the leaf's pivot rows; columns in
    // chunks of 8 so that the thread's RPT x 8 loads are in flight together
    for (int q0 = 0; q0 < nr; q0 += 8) {
      double2 v[RPT][8];
#pragma unroll
      for (int s = 0; s < RPT; ++s) {
        const int r = tid + s * RT;
#pragma unroll
        for (int qq = 0; qq < 8; ++qq)
          v[s][qq] = (r >= c1 && r < m && q0 + qq < nr) ? P[(int64_t)(c1 + q0 + qq) * n + r]
                                                         : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int s = 0; s < RPT; ++s) {
        const int r = tid + s * RT;
        if (r >= c1 && r < m) {
#pragma unroll
          for (int qq = 0; qq < 8; ++qq) {
            if (q0 + qq < nr) {
              double2 t = v[s][qq];
#pragma unroll
              for (int p = 0; p < LW; ++p)
                if (p < lw) t = csub(t, cmul(x[s][p], s_u[p][q0 + qq]));
              P[(int64_t)(c1 + q0 + qq) * n + r] = t;
            }
          }
        }
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
  PrecScope scope(h, EUVL_PREC_FP32_LU);
  const double2 one = make_double2(1.0, 0.0), mone = make_double2(-1.0, 0.0), zero = make_double2(0.0, 0.0);
  double2* Pt = (double2*)aux_reserve(h, sizeof(double2) * (size_t)batch * NB * n);
  if (!Pt) {
    set_error("zgetrf: panel scratch allocation failed");
    return EUVL_ERR_NOMEM;
  }
  hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(256), 0, h->stream, d_info, batch);
  EUVL_LAUNCH_CHECK();
  for (int kb = 0; kb < n; kb += NB) {
    const int nbc = (n - kb < NB) ? n - kb : NB;
    const int m = n - kb;
    hipLaunchKernelGGL(panel_in_kernel, dim3((m + 255) / 256, batch), dim3(256), 0, h->stream, n, A, lda, sa, kb, nbc,
                       Pt);
    EUVL_LAUNCH_CHECK();
    if (m <= RT * RPT)
      hipLaunchKernelGGL(getrf_panel_reg_kernel, dim3(batch), dim3(RT), 0, h->stream, n, Pt, ipiv, kb, nbc, d_info);
    else
      hipLaunchKernelGGL(getrf_panel_kernel, dim3(batch), dim3(PT), 0, h->stream, n, Pt, ipiv, kb, nbc, d_info);
    EUVL_LAUNCH_CHECK();
    hipLaunchKernelGGL(panel_out_kernel, dim3((m + 255) / 256, batch), dim3(256), 0, h->stream, n, A, lda, sa, kb,
                       nbc, Pt);
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
  PrecScope scope(h, EUVL_PREC_FP32_LU);
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
