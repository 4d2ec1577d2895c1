// Batched complex FP64 GEMM on CDNA4 FP64 MFMA (v_mfma_f64_16x16x4_f64).
//
// C = alpha*A*B + beta*C + gamma*I, row-major, interleaved complex.
// Workgroup tile 64x64, 4 waves in 2x2, each wave 32x32 = 2x2 MFMA tiles of
// 16x16 with separate real/imag accumulators (4 real MFMAs per complex MFMA
// step: C_re += A_re B_re - A_im B_im, C_im += A_re B_im + A_im B_re).
// K is staged through LDS in slabs of 16 with a register prefetch of the next
// slab.  Operand layout of the f64 16x16x4 MFMA: lane l holds A[l%16][l/16]
// and B[l/16][l%16]; the 4 accumulator doubles of lane l are
// D[4*r + l/16][l%16], r = 0..3 (measured on gfx950).
#include "common.h"

namespace {

constexpr int BM = 64, BN = 64, BK = 16;
constexpr int PADA = 1, PADB = 1;

__global__ __launch_bounds__(256) void zgemm_mfma_kernel(
    int M, int N, int K, double2 alpha, const double2* __restrict__ A, int64_t lda, int64_t sa,
    const double2* __restrict__ B, int64_t ldb, int64_t sb, double2 beta, double2* __restrict__ C,
    int64_t ldc, int64_t sc, double2 gamma, int tiles_n) {
  __shared__ double2 As[BK][BM + PADA];
  __shared__ double2 Bs[BK][BN + PADB];

  const int b = blockIdx.z;
  A += (int64_t)b * sa;
  B += (int64_t)b * sb;
  C += (int64_t)b * sc;

  const int tile = blockIdx.x;
  const int row0 = (tile / tiles_n) * BM;
  const int col0 = (tile % tiles_n) * BN;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int li = lane & 15, lk = lane >> 4;

  // global->LDS assignment: A slab 64 rows x 16 k: thread -> row tid/4, k 4*(tid%4)..+3
  const int a_r = tid >> 2, a_k = (tid & 3) * 4;
  // B slab 16 k x 64 cols: thread -> k tid/16, cols 4*(tid%16)..+3
  const int b_k = tid >> 4, b_c = (tid & 15) * 4;

  double2 ra[4], rb[4];
  auto load_slab = [&](int k0) {
    const int gr = row0 + a_r;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int gk = k0 + a_k + q;
      ra[q] = (gr < M && gk < K) ? A[(int64_t)gr * lda + gk] : make_double2(0.0, 0.0);
    }
    const int gk = k0 + b_k;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int gc = col0 + b_c + q;
      rb[q] = (gk < K && gc < N) ? B[(int64_t)gk * ldb + gc] : make_double2(0.0, 0.0);
    }
  };

  double4_t acc_re[2][2], acc_im[2][2];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      acc_re[t][u] = (double4_t){0.0, 0.0, 0.0, 0.0};
      acc_im[t][u] = (double4_t){0.0, 0.0, 0.0, 0.0};
    }

  load_slab(0);
  for (int k0 = 0; k0 < K; k0 += BK) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) As[a_k + q][a_r] = ra[q];
#pragma unroll
    for (int q = 0; q < 4; ++q) Bs[b_k][b_c + q] = rb[q];
    __syncthreads();
    if (k0 + BK < K) load_slab(k0 + BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      double2 fa[2], fb[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) fa[t] = As[ks + lk][wm * 32 + t * 16 + li];
#pragma unroll
      for (int u = 0; u < 2; ++u) fb[u] = Bs[ks + lk][wn * 32 + u * 16 + li];
      // four passes over the 2x2 tiles so that successive updates of one
      // accumulator are 7 independent MFMAs apart
      double nai[2] = {-fa[0].y, -fa[1].y};
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u)
          acc_re[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[t].x, fb[u].x, acc_re[t][u], 0, 0, 0);
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u)
          acc_im[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[t].x, fb[u].y, acc_im[t][u], 0, 0, 0);
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u)
          acc_re[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[t], fb[u].y, acc_re[t][u], 0, 0, 0);
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u)
          acc_im[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[t].y, fb[u].x, acc_im[t][u], 0, 0, 0);
    }
  }

  const bool use_beta = (beta.x != 0.0 || beta.y != 0.0);
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gr = row0 + wm * 32 + t * 16 + r * 4 + lk;
        const int gc = col0 + wn * 32 + u * 16 + li;
        if (gr < M && gc < N) {
          double2 v = cmul(alpha, make_double2(acc_re[t][u][r], acc_im[t][u][r]));
          double2* cp = C + (int64_t)gr * ldc + gc;
          if (use_beta) v = cadd(v, cmul(beta, *cp));
          if (gr == gc) v = cadd(v, gamma);
          *cp = v;
        }
      }
}

}  // namespace

namespace euvl {

int zgemm(euvl_handle h, int m, int n, int k, double2 alpha, const double2* A, int64_t lda, int64_t sa,
          const double2* B, int64_t ldb, int64_t sb, double2 beta, double2* C, int64_t ldc, int64_t sc,
          double2 gamma, int batch) {
  if (m <= 0 || n <= 0 || batch <= 0) return EUVL_OK;
  if (k <= 0) {
    // C = beta*C + gamma*I: run with a zero-length K (the kernel's slab loop is empty).
  }
  const int tiles_m = (m + BM - 1) / BM, tiles_n = (n + BN - 1) / BN;
  dim3 grid(tiles_m * tiles_n, 1, batch);
  int ev = -1;
  if (h->prof) {
    if (h->ev_next + 2 > h->ev_pool.size()) {
      for (int i = 0; i < 256; ++i) {
        hipEvent_t e;
        EUVL_HIP_CHECK(hipEventCreate(&e));
        h->ev_pool.push_back(e);
      }
    }
    ev = (int)h->ev_next;
    h->ev_next += 2;
    EUVL_HIP_CHECK(hipEventRecord(h->ev_pool[ev], h->stream));
  }
  hipLaunchKernelGGL(zgemm_mfma_kernel, grid, dim3(256), 0, h->stream, m, n, k, alpha, A, lda, sa, B, ldb,
                     sb, beta, C, ldc, sc, gamma, tiles_n);
  EUVL_LAUNCH_CHECK();
  if (ev >= 0) {
    EUVL_HIP_CHECK(hipEventRecord(h->ev_pool[ev + 1], h->stream));
    h->ev_used.push_back({ev, 8.0 * (double)m * (double)n * (double)(k > 0 ? k : 0) * batch});
  }
  return EUVL_OK;
}

}  // namespace euvl

extern "C" int euvl_zgemm_batched(euvl_handle h, int m, int n, int k, const double* alpha, const double* d_A,
                                  int64_t lda, int64_t sa, const double* d_B, int64_t ldb, int64_t sb,
                                  const double* beta, double* d_C, int64_t ldc, int64_t sc,
                                  const double* gamma, int batch) {
  if (!h || m < 0 || n < 0 || k < 0 || batch < 0 || !alpha || !beta) {
    euvl::set_error("euvl_zgemm_batched: bad argument");
    return EUVL_ERR_ARG;
  }
  if ((k > 0 && (lda < k || ldb < n)) || ldc < n) {
    euvl::set_error("euvl_zgemm_batched: leading dimension too small");
    return EUVL_ERR_ARG;
  }
  double2 g = gamma ? make_double2(gamma[0], gamma[1]) : make_double2(0.0, 0.0);
  return euvl::zgemm(h, m, n, k, make_double2(alpha[0], alpha[1]), (const double2*)d_A, lda, sa,
                     (const double2*)d_B, ldb, sb, make_double2(beta[0], beta[1]), (double2*)d_C, ldc, sc, g,
                     batch);
}

extern "C" int euvl_gemm_stats(euvl_handle h, int enable, double* out) {
  if (!h) return EUVL_ERR_ARG;
  if (out) {
    EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));
    double ms = 0.0, flops = 0.0;
    for (auto& u : h->ev_used) {
      float t = 0.0f;
      EUVL_HIP_CHECK(hipEventElapsedTime(&t, h->ev_pool[u.first], h->ev_pool[u.first + 1]));
      ms += t;
      flops += u.second;
    }
    out[0] = (double)h->ev_used.size();
    out[1] = ms;
    out[2] = flops;
  }
  h->ev_used.clear();
  h->ev_next = 0;
  h->prof = enable != 0;
  return EUVL_OK;
}
