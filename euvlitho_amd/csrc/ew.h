// Internal interfaces shared between the libeuvl translation units.
#pragma once
#include "common.h"

namespace euvl {

int zgemm(euvl_handle h, int m, int n, int k, double2 alpha, const double2* A, int64_t lda, int64_t sa,
          const double2* B, int64_t ldb, int64_t sb, double2 beta, double2* C, int64_t ldc, int64_t sc,
          double2 gamma, int batch);
int zgetrf(euvl_handle h, int n, double2* A, int64_t lda, int64_t sa, int* ipiv, int batch, int* d_info);
int zgetrs(euvl_handle h, int n, int nrhs, const double2* LU, int64_t lda, int64_t sa, const int* ipiv,
           double2* B, int64_t ldb, int64_t sb, int batch);

int ew_lincomb(euvl_handle h, int n, int batch, double2* out, double2 a1, const double2* X1, double2 a2,
               const double2* X2, double2 a3, const double2* X3, double2 g);
int ew_bscale(euvl_handle h, int n, int batch, double2* out, const double2* X, const double2* d_s, double2 g);
int ew_diagscale(euvl_handle h, int n, int batch, double2* out, const double2* X, const double2* row,
                 const double2* col);
int ew_copy(euvl_handle h, int64_t count, double2* out, const double2* X);
// out[0:count) = v (a fill kernel: safe inside graph capture, unlike hipMemsetAsync)
int fill(euvl_handle h, int64_t count, double2* out, double2 v);
// kind 0: max|X-I| (|re|+|im|), 1: infinity norm (max row sum of |x|), 2: mean Re diag
int reduce(euvl_handle h, int kind, int n, int batch, const double2* X, double* d_red);
int reduce_to_host(euvl_handle h, int kind, int n, int batch, const double2* X, double* out);
// max |re| + |im| over an m x ncols block of each batch entry (leading dim ld, batch stride sb) into
// d_slot[0:batch) (maxabs), or to the host, synchronising (maxabs_to_host)
int maxabs(euvl_handle h, int m, int ncols, int64_t ld, int64_t sb, int batch, const double2* X, double* d_slot);
int maxabs_to_host(euvl_handle h, int m, int ncols, int64_t ld, int64_t sb, int batch, const double2* X,
                   double* out);
// LU without row interchanges (blocked, GEMM panels); ipiv is written as the identity.  Only for matrices
// whose unpivoted elimination is safe; callers verify what they solve with it.
int zgetrf_nopiv(euvl_handle h, int n, double2* A, int64_t lda, int64_t sa, int* ipiv, int batch, int* d_info);

// matrix functions (matfun.hip); buffers are batch x n x n contiguous.
// C = cos(d sqrt D), F = sin(d sqrt D)/sqrt D, G2 = D F (G2 may be null); takes 6 workspace matrices.
int cos_sinc(euvl_handle h, int n, int batch, const double2* D, double d, double2* C, double2* F, double2* G2,
             int* doublings);

}  // namespace euvl
