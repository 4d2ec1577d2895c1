// Batched complex FP64 GEMM on CDNA4 FP64 MFMA (v_mfma_f64_16x16x4_f64).
//
// C = alpha*A*B + beta*C + gamma*I, row-major, interleaved complex.
// A workgroup of 4 waves (2 x 2) computes a (32 TM) x (32 TN) tile; each wave
// owns TM x TN MFMA tiles of 16 x 16 with separate real/imag accumulators
// (4 real MFMAs per complex MFMA step: C_re += A_re B_re - A_im B_im,
// C_im += A_re B_im + A_im B_re).  K is staged through LDS in slabs of 16 with
// a register prefetch of the next slab.  Operand layout of the f64 16x16x4
// MFMA: lane l holds A[l%16][l/16] and B[l/16][l%16]; the 4 accumulator
// doubles of lane l are D[4*r + l/16][l%16], r = 0..3 (measured on gfx950).
//
// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2);
// the linear workgroup id is remapped so that every XCD works on one
// contiguous range of tiles (shared A row panels stay in one L2).
#include <cstdlib>

#include "common.h"

namespace {

template <int TM, int TN, int BK, int MINB>
__global__ __launch_bounds__(256, MINB) void zgemm_mfma_kernel(
    int M, int N, int K, double2 alpha, const double2* __restrict__ A, int64_t lda, int64_t sa,
    const double2* __restrict__ B, int64_t ldb, int64_t sb, double2 beta, double2* __restrict__ C,
    int64_t ldc, int64_t sc, double2 gamma, int tiles_n, int tiles) {
  constexpr int BM = 32 * TM, BN = 32 * TN;
  constexpr int QA = BM * BK / 256, QB = BN * BK / 256;  // complex values per thread per slab
  __shared__ double2 As[BK][BM + 1];
  __shared__ double2 Bs[BK][BN + 1];

  // XCD-aware remap of the linear workgroup id (bijective on [0, total))
  const int total = gridDim.x;
  const int g = blockIdx.x;
  const int q8 = total / 8;
  const int ng = (g < 8 * q8) ? (g % 8) * q8 + g / 8 : g;
  const int b = ng / tiles;
  const int tile = ng - b * tiles;
  A += (int64_t)b * sa;
  B += (int64_t)b * sb;
  C += (int64_t)b * sc;
  const int row0 = (tile / tiles_n) * BM;
  const int col0 = (tile % tiles_n) * BN;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int li = lane & 15, lk = lane >> 4;

  // global->LDS assignment: A slab BM x 16 (row-major, QA consecutive k per thread),
  // B slab 16 x BN (QB consecutive columns per thread)
  const int a_r = tid / (BK / QA), a_k = (tid % (BK / QA)) * QA;
  const int b_k = tid / (BN / QB), b_c = (tid % (BN / QB)) * QB;

  double2 ra[QA], rb[QB];
  auto load_slab = [&](int k0) {
    const int gr = row0 + a_r;
#pragma unroll
    for (int q = 0; q < QA; ++q) {
      const int gk = k0 + a_k + q;
      ra[q] = (gr < M && gk < K) ? A[(int64_t)gr * lda + gk] : make_double2(0.0, 0.0);
    }
    const int gk = k0 + b_k;
#pragma unroll
    for (int q = 0; q < QB; ++q) {
      const int gc = col0 + b_c + q;
      rb[q] = (gk < K && gc < N) ? B[(int64_t)gk * ldb + gc] : make_double2(0.0, 0.0);
    }
  };

  double4_t acc_re[TM][TN], acc_im[TM][TN];
#pragma unroll
  for (int t = 0; t < TM; ++t)
#pragma unroll
    for (int u = 0; u < TN; ++u) {
      acc_re[t][u] = (double4_t){0.0, 0.0, 0.0, 0.0};
      acc_im[t][u] = (double4_t){0.0, 0.0, 0.0, 0.0};
    }

  load_slab(0);
  for (int k0 = 0; k0 < K; k0 += BK) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < QA; ++q) As[a_k + q][a_r] = ra[q];
#pragma unroll
    for (int q = 0; q < QB; ++q) Bs[b_k][b_c + q] = rb[q];
    __syncthreads();
    if (k0 + BK < K) load_slab(k0 + BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      double2 fa[TM], fb[TN];
      double nai[TM];
#pragma unroll
      for (int t = 0; t < TM; ++t) {
        fa[t] = As[ks + lk][wm * 16 * TM + t * 16 + li];
        nai[t] = -fa[t].y;
      }
#pragma unroll
      for (int u = 0; u < TN; ++u) fb[u] = Bs[ks + lk][wn * 16 * TN + u * 16 + li];
      // four passes over the tiles so that successive updates of one
      // accumulator are separated by independent MFMAs
#pragma unroll
      for (int t = 0; t < TM; ++t)
#pragma unroll
        for (int u = 0; u < TN; ++u)
          acc_re[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[t].x, fb[u].x, acc_re[t][u], 0, 0, 0);
#pragma unroll
      for (int t = 0; t < TM; ++t)
#pragma unroll
        for (int u = 0; u < TN; ++u)
          acc_im[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[t].x, fb[u].y, acc_im[t][u], 0, 0, 0);
#pragma unroll
      for (int t = 0; t < TM; ++t)
#pragma unroll
        for (int u = 0; u < TN; ++u)
          acc_re[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[t], fb[u].y, acc_re[t][u], 0, 0, 0);
#pragma unroll
      for (int t = 0; t < TM; ++t)
#pragma unroll
        for (int u = 0; u < TN; ++u)
          acc_im[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[t].y, fb[u].x, acc_im[t][u], 0, 0, 0);
    }
  }

  const bool use_beta = (beta.x != 0.0 || beta.y != 0.0);
#pragma unroll
  for (int t = 0; t < TM; ++t)
#pragma unroll
    for (int u = 0; u < TN; ++u)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gr = row0 + wm * 16 * TM + t * 16 + r * 4 + lk;
        const int gc = col0 + wn * 16 * TN + u * 16 + li;
        if (gr < M && gc < N) {
          double2 v = cmul(alpha, make_double2(acc_re[t][u][r], acc_im[t][u][r]));
          double2* cp = C + (int64_t)gr * ldc + gc;
          if (use_beta) v = cadd(v, cmul(beta, *cp));
          if (gr == gc) v = cadd(v, gamma);
          *cp = v;
        }
      }
}

// kernel variants (64x64 workgroup tile, K slab 16): 0 = uncapped registers
// (2 waves/SIMD), 1 = registers capped for >= 3 waves/SIMD (the compiler
// reaches 4 with 126 VGPRs).  Measured on MI355X (N = 1681 x 4 / 8705):
// 50 / 59 and 56 / 64 TFLOP/s; 128x64, 64x128 and 128x128 tiles (1 wave/SIMD)
// reached 32-40, and a K slab of 32 48-55.  1 is the default and
// EUVL_GEMM_TILE overrides it for tuning runs.
int g_variant = -1;

int pick_variant(int, int) {
  if (g_variant < 0) {
    const char* e = getenv("EUVL_GEMM_TILE");
    g_variant = e ? atoi(e) : 1;
    if (g_variant < 0 || g_variant > 1) g_variant = 1;
  }
  return g_variant;
}

template <int TM, int TN, int BK, int MINB>
void launch(euvl_handle h, int m, int n, int k, double2 alpha, const double2* A, int64_t lda, int64_t sa,
            const double2* B, int64_t ldb, int64_t sb, double2 beta, double2* C, int64_t ldc, int64_t sc,
            double2 gamma, int batch) {
  constexpr int BM = 32 * TM, BN = 32 * TN;
  const int tiles_m = (m + BM - 1) / BM, tiles_n = (n + BN - 1) / BN;
  const int tiles = tiles_m * tiles_n;
  hipLaunchKernelGGL((zgemm_mfma_kernel<TM, TN, BK, MINB>), dim3(tiles * batch), dim3(256), 0, h->stream, m, n, k,
                     alpha, A, lda, sa, B, ldb, sb, beta, C, ldc, sc, gamma, tiles_n, tiles);
}

}  // namespace

namespace euvl {

int zgemm(euvl_handle h, int m, int n, int k, double2 alpha, const double2* A, int64_t lda, int64_t sa,
          const double2* B, int64_t ldb, int64_t sb, double2 beta, double2* C, int64_t ldc, int64_t sc,
          double2 gamma, int batch) {
  if (m <= 0 || n <= 0 || batch <= 0) return EUVL_OK;
  if ((int64_t)((m + 31) / 32) * ((n + 31) / 32) * batch > (int64_t)1 << 30) {
    set_error("zgemm: grid too large");
    return EUVL_ERR_ARG;
  }
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
  switch (pick_variant(m, n)) {
    case 0: launch<2, 2, 16, 1>(h, m, n, k, alpha, A, lda, sa, B, ldb, sb, beta, C, ldc, sc, gamma, batch); break;
    default: launch<2, 2, 16, 3>(h, m, n, k, alpha, A, lda, sa, B, ldb, sb, beta, C, ldc, sc, gamma, batch); break;
  }
  EUVL_LAUNCH_CHECK();
  if (ev >= 0) {
    EUVL_HIP_CHECK(hipEventRecord(h->ev_pool[ev + 1], h->stream));
    const double kk = k > 0 ? k : 0, nb = batch;
    const bool rd_c = beta.x != 0.0 || beta.y != 0.0;
    const double bytes = 16.0 * (m * kk * (sa ? nb : 1.0) + kk * n * (sb ? nb : 1.0) + (double)m * n * nb * (rd_c ? 2 : 1));
    h->ev_used.push_back({ev, 8.0 * (double)m * (double)n * kk * nb, bytes});
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
    double ms = 0.0, flops = 0.0, bytes = 0.0;
    for (auto& u : h->ev_used) {
      float t = 0.0f;
      EUVL_HIP_CHECK(hipEventElapsedTime(&t, h->ev_pool[u.ev], h->ev_pool[u.ev + 1]));
      ms += t;
      flops += u.flops;
      bytes += u.bytes;
    }
    out[0] = (double)h->ev_used.size();
    out[1] = ms;
    out[2] = flops;
    out[3] = bytes;
  }
  h->ev_used.clear();
  h->ev_next = 0;
  h->prof = enable != 0;
  return EUVL_OK;
}
