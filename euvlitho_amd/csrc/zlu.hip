// Batched complex LU (right-looking, panel width 32) and the matching
// triangular solves.  Panels are factored by tournament pivoting over
// several workgroups (below); the trailing update is the MFMA GEMM of
// zgemm.hip.
#include <algorithm>
#include <climits>

#include "ew.h"

namespace {

constexpr int NB = 32;

__device__ inline double cabs1(double2 a) { return fabs(a.x) + fabs(a.y); }

__global__ void zero_int_kernel(int* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// Panel factorisation by tournament pivoting (TSLU, the panel step of
// communication-avoiding LU).  The m x nbc panel is cut into groups of up to
// TS rows; every group is factored with partial pivoting in registers (one
// row per thread, the pivot row broadcast through LDS) only to choose its nbc
// candidate pivot rows; the candidates of TS / nbc groups form the groups of
// the next round, until one group is left.  The partial-pivoting LU of that
// last group gives the panel's pivot rows and its L11\U11 block.  Each group
// is a workgroup of its own, so a tall panel spreads over ceil(m / TS) CUs
// with its rows in registers, instead of one CU streaming the panel through
// global memory once per column block.  A last parallel pass moves the rows
// and forms L21 = A21 U11^-1.  Every round reads the original panel values:
// nothing is written to A before that pass.
// rows (= threads) per group.  A column step costs about the same with one
// wave (no cross-wave exchange) as with eight, so larger groups (fewer
// rounds) win: measured per 1681 x 1681 x 4 LU, 24 ms with 512, 31 ms with 64.
constexpr int TSZ = 512;

__device__ inline double2 czero() { return make_double2(0.0, 0.0); }

// Round over the groups of one panel.  Group g of batch entry b factors the
// panel rows rows_in[b][g*TS ...] (panel positions; the identity when rows_in
// is null) and writes its min(nbc, rows) pivot rows, in order, to
// rows_out[b][g*nbc ...].  The last round (one group) instead writes the
// LU11 block (row j = [L11(j, :j), U11(j, j:)]), the original values of the
// panel's top nbc rows, the row interchanges ipiv (LAPACK order, absolute
// rows) and the list of moved rows, and flags the first zero pivot in info.
//
// The column loop is not unrolled: a 32-step unrolled body is larger than the
// instruction cache and every step would refetch it.  Instead each thread
// keeps a window of its row that shifts left by one column per step, so that
// the current column is always x[0] and every register index is static; the
// multipliers that leave the window go to lscr (last round only), from where
// the pivot rows' L11 parts are gathered at the end.
template <int TS>
__global__ __launch_bounds__(TS) void tslu_round_kernel(const double2* __restrict__ A, int64_t lda, int64_t sa,
                                                        int n, int kb, int nbc, const int* __restrict__ rows_in,
                                                        int n_in, int* __restrict__ rows_out, int last,
                                                        double2* __restrict__ lu11, double2* __restrict__ top,
                                                        double2* __restrict__ lscr, int gstride,
                                                        int* __restrict__ ipiv, int* __restrict__ moved,
                                                        int* __restrict__ info) {
  __shared__ double s_val[TS / 64];
  __shared__ int s_idx[TS / 64];
  __shared__ double2 s_prow[NB];
  __shared__ int s_pick[NB], s_pickt[NB];
  __shared__ int s_tp[2 * NB], s_tr[2 * NB];
  const int b = blockIdx.y, g = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double2* P = A + (int64_t)b * sa + (int64_t)kb * lda + kb;  // panel origin
  const int first = g * TS;
  const int rows = min(TS, n_in - first);
  const int steps = min(nbc, rows);
  int pos = -1;
  if (tid < rows) pos = rows_in ? rows_in[(int64_t)b * n + first + tid] : first + tid;
  double2 x[NB];
  {
    // unconditional loads from a valid row, masked afterwards
    const double2* row = P + (int64_t)max(pos, 0) * lda;
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const double2 v = row[min(q, nbc - 1)];
      const bool ok = pos >= 0 && q < nbc;
      x[q] = make_double2(ok ? v.x : 0.0, ok ? v.y : 0.0);  // component selects: no stack temporary
    }
  }
  double2* my_l = lscr + (((int64_t)b * gstride + g) * TS + tid) * NB;
  bool active = pos >= 0;
#pragma unroll 1
  for (int j = 0; j < steps; ++j) {
    // pivot: largest |re|+|im| of the current column over the active rows (first in input order on ties)
    double best = active ? cabs1(x[0]) : -1.0;
    int bidx = active ? tid : TS;
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(best, off);
      const int oi = __shfl_xor(bidx, off);
      if (ov > best || (ov == best && oi < bidx)) { best = ov; bidx = oi; }
    }
    int bi = bidx;  // one wave: the butterfly left the winner in every lane
    if constexpr (TS > 64) {
      if (lane == 0) { s_val[wave] = best; s_idx[wave] = bidx; }
      __syncthreads();
      double bv = -2.0;
      bi = TS;
#pragma unroll
      for (int w = 0; w < TS / 64; ++w) {
        const double v = s_val[w];
        const int i = s_idx[w];
        if (v > bv || (v == bv && i < bi)) { bv = v; bi = i; }
      }
    }
    if (tid == bi) {
#pragma unroll
      for (int q = 0; q < NB; ++q) s_prow[q] = x[q];
      s_pick[j] = pos;
      s_pickt[j] = tid;
      active = false;
      if (last) {
        double2* out = lu11 + ((int64_t)b * NB + j) * NB + j;  // U11(j, j:)
#pragma unroll
        for (int q = 0; q < NB; ++q)
          if (q < nbc - j) out[q] = x[q];
        if (x[0].x == 0.0 && x[0].y == 0.0 && info[b] == 0) info[b] = kb + j + 1;
      }
    }
    __syncthreads();
    const double2 d = s_prow[0];
    const bool elim = active && (d.x != 0.0 || d.y != 0.0);
    const double2 x0 = x[0];
    double2 l = czero();
    if (elim) l = cdiv(x0, d);
    // shift the window left by one column while eliminating
#pragma unroll
    for (int q = 0; q < NB - 1; ++q) x[q] = csub(x[q + 1], cmul(l, s_prow[q + 1]));
    x[NB - 1] = czero();
    // the multiplier leaving the window (read back for the pivot rows of the last round)
    my_l[j] = make_double2(elim ? l.x : x0.x, elim ? l.y : x0.y);
  }
  __syncthreads();
  if (!last) {
    if (tid < steps) rows_out[(int64_t)b * n + g * nbc + tid] = s_pick[tid];
    return;
  }
  // L11(j, k), k < j: the multipliers the j-th pivot row got before it was picked
  for (int e = tid; e < NB * NB; e += TS) {
    const int jr = e / NB, k = e % NB;
    if (jr < nbc && k < jr)
      lu11[((int64_t)b * NB + jr) * NB + k] = lscr[((int64_t)b * gstride * TS + s_pickt[jr]) * NB + k];
  }
  // original values of the panel's top nbc rows (they move to the pivots' places)
  if (tid < nbc) {
    const double2* row = P + (int64_t)tid * lda;
    double2* out = top + ((int64_t)b * NB + tid) * NB;
    for (int q = 0; q < nbc; ++q) out[q] = row[q];
  }
  // the pivot rows as LAPACK's sequence of interchanges: position j <-> the
  // position the j-th pivot row holds at that point (positions >= nbc that
  // are touched end up holding rows that started above nbc)
  if (tid == 0) {
    int nt = 0;
    for (int j = 0; j < nbc; ++j) {
      const int r = s_pick[j];
      int p = r;  // where row r is now
      for (int t = 0; t < nt; ++t)
        if (s_tr[t] == r) p = s_tp[t];
      int rj = j;  // which row is at position j now
      for (int t = 0; t < nt; ++t)
        if (s_tp[t] == j) rj = s_tr[t];
      ipiv[(int64_t)b * n + kb + j] = kb + p;
      bool fj = false, fp = false;
      for (int t = 0; t < nt; ++t) {
        if (s_tp[t] == j) { s_tr[t] = r; fj = true; }
        else if (s_tp[t] == p) { s_tr[t] = rj; fp = true; }
      }
      if (!fj) { s_tp[nt] = j; s_tr[nt] = r; ++nt; }
      if (!fp && p != j) { s_tp[nt] = p; s_tr[nt] = rj; ++nt; }
    }
    int* mv = moved + (int64_t)b * (1 + 4 * NB);
    mv[0] = nt;
    for (int t = 0; t < nt; ++t) {
      mv[1 + 2 * t] = s_tp[t];
      mv[2 + 2 * t] = s_tr[t];
    }
  }
}

// Final pass over the panel rows: position q receives L11\U11 row q (q < nbc)
// or the row that ends up there, turned into L21 by solving l U11 = a column
// by column (l_c = a_c / U(c,c), then a -= l_c U(c, :)) with the same
// shifting window as the rounds.
__global__ __launch_bounds__(256) void tslu_apply_kernel(double2* __restrict__ A, int64_t lda, int64_t sa, int kb,
                                                         int m, int nbc, const double2* __restrict__ lu11,
                                                         const double2* __restrict__ top,
                                                         const int* __restrict__ moved) {
  __shared__ double2 s_u[NB][NB + 1];
  __shared__ int s_tp[2 * NB], s_tr[2 * NB];
  __shared__ int s_nt;
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  for (int e = tid; e < NB * NB; e += blockDim.x) {
    const int r = e / NB, c = e % NB;
    s_u[r][c] = lu11[((int64_t)b * NB + r) * NB + c];
  }
  const int* mv = moved + (int64_t)b * (1 + 4 * NB);
  if (tid == 0) s_nt = mv[0];
  if (tid < 2 * NB) {
    s_tp[tid] = mv[1 + 2 * tid];
    s_tr[tid] = mv[2 + 2 * tid];
  }
  __syncthreads();
  const int q = blockIdx.x * blockDim.x + tid;
  if (q >= m) return;
  double2* Pq = A + (int64_t)b * sa + (int64_t)(kb + q) * lda + kb;
  if (q < nbc) {
    for (int c = 0; c < nbc; ++c) Pq[c] = s_u[q][c];
    return;
  }
  int src = q;
  const int nt = s_nt;
  for (int t = 0; t < nt; ++t)
    if (s_tp[t] == q) src = s_tr[t];
  const double2* row = (src < nbc) ? top + ((int64_t)b * NB + src) * NB
                                   : A + (int64_t)b * sa + (int64_t)(kb + src) * lda + kb;
  double2 a[NB];
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    const double2 v = row[min(c, nbc - 1)];
    a[c] = make_double2(c < nbc ? v.x : 0.0, c < nbc ? v.y : 0.0);
  }
#pragma unroll 1
  for (int c = 0; c < nbc; ++c) {
    const double2 d = s_u[c][c];
    double2 l = a[0];
    if (d.x != 0.0 || d.y != 0.0) l = cdiv(l, d);
    Pq[c] = l;
#pragma unroll
    for (int k = 0; k < NB - 1; ++k) a[k] = csub(a[k + 1], cmul(l, s_u[c][min(c + 1 + k, NB - 1)]));
    a[NB - 1] = czero();
  }
}

// ---------------------------------------------------------------------------
// LU without row interchanges.  Each 32-wide panel starts with its 32 x 32
// diagonal block, factored by one workgroup per batch entry, which also
// forms U11^-1 and L11^-1; the panel's column and row blocks then become
// GEMMs (L21 = A21 U11^-1, U12 = L11^-1 A12), as does the trailing update.
// No column step of the factorisation runs over the panel's full height, so
// a panel costs a few short launches instead of 32 dependent pivot searches.
// The block is small and factored once per launch, so the kernel keeps it
// in registers of 512 threads (thread (r, c) holds rows r and r+16 of column
// c) with short loops: a fully unrolled single-lane-per-row form runs to
// ~20k instructions and is bound by instruction fetch.  The inverses are
// row-parallel substitutions (lane i of a 32-lane group holds row i of four
// columns, the finished row is broadcast at each step); four waves form
// U11^-1 while the other four form L11^-1.  512 threads measured 34 us per
// block, 256 threads 44 us and 1024 threads 37 us.
constexpr int DLU_THREADS = 512;
__global__ __launch_bounds__(DLU_THREADS) void diag_lu_kernel(double2* __restrict__ A, int64_t lda, int64_t sa,
                                                              int kb, int nbc, double2* __restrict__ uinv,
                                                              double2* __restrict__ linv, int* __restrict__ ipiv,
                                                              int n, int* __restrict__ info) {
  constexpr int RG = DLU_THREADS / NB;  // row groups (16)
  constexpr int RS = NB / RG;           // rows per thread (2)
  __shared__ double2 a[NB][NB + 1];
  // step j publishes row j and column j into buffers of parity j & 1, so one
  // barrier per step orders them (step j+1 writes the other buffers)
  __shared__ double2 rowb[2][NB], colb[2][NB];
  const int b = blockIdx.x, t = threadIdx.x;
  const int tq = t & (NB - 1), tr = t / NB;
  double2* A11 = A + (int64_t)b * sa + (int64_t)kb * lda + kb;
  double2 x[RS];
#pragma unroll
  for (int u = 0; u < RS; ++u) {
    const int r = tr + RG * u;
    const bool in = r < nbc && tq < nbc;
    const double pad = (r == tq) ? 1.0 : 0.0;  // identity outside the nbc x nbc block
    x[u] = in ? A11[(int64_t)r * lda + tq] : make_double2(pad, 0.0);
  }
  // right-looking elimination: row j and column j are final once step j-1 is
  // done; every thread forms the multipliers of its own rows
#pragma unroll 1
  for (int j = 0; j < NB - 1; ++j) {
    const int pb = j & 1;
#pragma unroll
    for (int u = 0; u < RS; ++u) {
      const int i = tr + RG * u;
      if (i == j) rowb[pb][tq] = x[u];
      if (tq == j && i >= j) colb[pb][i] = x[u];
    }
    __syncthreads();
    // the same pivot for every thread: a uniform reciprocal, and updates as
    // selects (no divergent branches, so the LDS reads are issued together)
    const double2 d = rowb[pb][j];
    if (d.x != 0.0 || d.y != 0.0) {  // a zero pivot leaves the block as it is
      const double2 rd = cdiv(make_double2(1.0, 0.0), d);
      const double2 uj = rowb[pb][tq];
      double2 c[RS];
#pragma unroll
      for (int u = 0; u < RS; ++u) c[u] = colb[pb][min(tr + RG * u, NB - 1)];
#pragma unroll
      for (int u = 0; u < RS; ++u) {
        const int i = tr + RG * u;
        const double2 l = cmul(c[u], rd);
        const double2 upd = (tq == j) ? l : csub(x[u], cmul(l, uj));
        const bool w = i > j && tq >= j;  // selects per component keep x in registers
        x[u].x = w ? upd.x : x[u].x;
        x[u].y = w ? upd.y : x[u].y;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < RS; ++u) a[tr + RG * u][tq] = x[u];
  __syncthreads();
#pragma unroll
  for (int u = 0; u < RS; ++u) {
    const int r = tr + RG * u;
    if (r < nbc && tq < nbc) A11[(int64_t)r * lda + tq] = x[u];
  }
  if (t < nbc) ipiv[(int64_t)b * n + kb + t] = kb + t;
  if (t == 0)
#pragma unroll 1
    for (int j = 0; j < nbc; ++j)
      if (a[j][j].x == 0.0 && a[j][j].y == 0.0 && info[b] == 0) info[b] = kb + j + 1;
  // U11^-1 (row groups 0-3) and L11^-1 (4-7): lane i holds row i of CB columns
  constexpr int CB = NB / (RG / 2);  // columns per group (4)
  const int i = tq, g = tr & (RG / 2 - 1);
  double2 y[CB];
#pragma unroll
  for (int v = 0; v < CB; ++v) y[v] = make_double2(i == g * CB + v ? 1.0 : 0.0, 0.0);
  if (tr < RG / 2) {
    // back substitution: y(p) *= 1/U(p,p), then rows i < p subtract U(i,p) y(p)
    const double2 rii = cdiv(make_double2(1.0, 0.0), a[i][i]);
#pragma unroll 1
    for (int p = NB - 1; p >= 0; --p) {
      const double2 uip = a[i][p];
      double2 yp[CB];
#pragma unroll
      for (int v = 0; v < CB; ++v) {
        const double2 sc = cmul(y[v], rii);
        y[v].x = (i == p) ? sc.x : y[v].x;
        y[v].y = (i == p) ? sc.y : y[v].y;
        yp[v] = make_double2(__shfl(y[v].x, p, NB), __shfl(y[v].y, p, NB));
      }
#pragma unroll
      for (int v = 0; v < CB; ++v) {
        const double2 upd = csub(y[v], cmul(uip, yp[v]));
        y[v].x = (i < p) ? upd.x : y[v].x;
        y[v].y = (i < p) ? upd.y : y[v].y;
      }
    }
#pragma unroll
    for (int v = 0; v < CB; ++v) {
      const int c = g * CB + v;
      uinv[((int64_t)b * NB + i) * NB + c] = make_double2(i <= c ? y[v].x : 0.0, i <= c ? y[v].y : 0.0);
    }
  } else {
    // forward substitution with the unit lower factor
#pragma unroll 1
    for (int p = 0; p < NB - 1; ++p) {
      const double2 lip = a[i][p];
      double2 yp[CB];
#pragma unroll
      for (int v = 0; v < CB; ++v) yp[v] = make_double2(__shfl(y[v].x, p, NB), __shfl(y[v].y, p, NB));
#pragma unroll
      for (int v = 0; v < CB; ++v) {
        const double2 upd = csub(y[v], cmul(lip, yp[v]));
        y[v].x = (i > p) ? upd.x : y[v].x;
        y[v].y = (i > p) ? upd.y : y[v].y;
      }
    }
#pragma unroll
    for (int v = 0; v < CB; ++v) {
      const int c = g * CB + v;
      linv[((int64_t)b * NB + i) * NB + c] = make_double2(i >= c ? y[v].x : 0.0, i >= c ? y[v].y : 0.0);
    }
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
// upper when UPPER), X nbc x ncols.  Each column is solved by 32 lanes, lane
// i holding x(i): at step p the finished x(p) is broadcast to its column's
// lanes and the rows it feeds subtract T(i, p) x(p) in parallel (a column
// per thread would leave a 171-column right-hand side on a handful of CUs
// behind a 500-step dependent chain).
constexpr int TRSM_COLS = 8;  // columns per 256-thread workgroup
template <bool UPPER>
__global__ __launch_bounds__(256) void trsm_kernel(int nbc, const double2* T, int64_t ldt, int64_t st, double2* X,
                                                   int64_t ldx, int64_t sx, int ncols) {
  __shared__ double2 sT[NB][NB + 1];
  const int b = blockIdx.y;
  T += (int64_t)b * st;
  X += (int64_t)b * sx;
  for (int e = threadIdx.x; e < nbc * nbc; e += blockDim.x) {
    const int r = e / nbc, c = e % nbc;
    sT[r][c] = T[(int64_t)r * ldt + c];
  }
  __syncthreads();
  const int i = threadIdx.x & (NB - 1);
  const int j = blockIdx.x * TRSM_COLS + threadIdx.x / NB;
  const bool live = j < ncols && i < nbc;  // every lane takes part in the broadcasts
  double2 x = live ? X[(int64_t)i * ldx + j] : make_double2(0.0, 0.0);
  if (!UPPER) {
#pragma unroll 1
    for (int p = 0; p < nbc - 1; ++p) {
      const double2 xp = make_double2(__shfl(x.x, p, NB), __shfl(x.y, p, NB));
      if (i > p) x = csub(x, cmul(sT[i][p], xp));
    }
  } else {
    // each lane's diagonal reciprocal up front: no division on the
    // substitution's dependent chain
    const double2 rii = (i < nbc) ? cdiv(make_double2(1.0, 0.0), sT[i][i]) : make_double2(0.0, 0.0);
#pragma unroll 1
    for (int p = nbc - 1; p >= 0; --p) {
      if (i == p) x = cmul(x, rii);
      const double2 xp = make_double2(__shfl(x.x, p, NB), __shfl(x.y, p, NB));
      if (i < p) x = csub(x, cmul(sT[i][p], xp));
    }
  }
  if (live) X[(int64_t)i * ldx + j] = x;
}

// Block Gauss-Jordan solve without row interchanges: the inverse of one
// GJ_NB x GJ_NB diagonal block per batch entry, by one workgroup.  Thread
// (r0, c) keeps elements (r0 + q RS, c) of the block in registers for the
// whole elimination; step p turns column p into the unit vector.  With the
// scaled pivot row s(c) = a(p,c) / a(p,p), and s(p) = 1 + 1 / a(p,p), every
// row takes a(i,c) - a(i,p) s(c), where the pivot row uses a(p,p) - 1 in place
// of a(i,p): that gives s(c) in the pivot row, 1 / a(p,p) at the pivot and
// -a(i,p) / a(p,p) in column p, with no branch per element.  Only row p and
// column p cross threads.  The wave holding row p (one wave in RS) reads
// a(p,p) from its lane p, takes the reciprocal and publishes s; lane p of
// every wave publishes its elements of column p.  Both go to LDS buffers of
// parity p & 1, so one barrier per step orders them, and the other waves do
// nothing per step but their RPT complex updates.  After GJ_NB steps the
// block holds its inverse.  Outside the nbc x nbc block the identity is used;
// a zero pivot inside it sets info (1-based column, the first one kept) and
// 1 / a(p,p) is taken as 0.
constexpr int GJ_NB = 64;

__device__ inline double readlane_f64(double v, int lane) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)bits, lane);
  const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int NT>
__global__ __launch_bounds__(NT) void gj_inv_kernel(const double2* __restrict__ A, int64_t lda, int64_t sa, int kb,
                                                    int nbc, double2* __restrict__ D, int* __restrict__ info) {
  static_assert(NT % GJ_NB == 0 && GJ_NB == 64, "one row of the block per wave");
  constexpr int RS = NT / GJ_NB;       // row stride between a thread's elements (= waves)
  constexpr int RPT = GJ_NB / RS;      // elements per thread
  __shared__ double2 rowb[2][GJ_NB], colb[2][GJ_NB];
  __shared__ int first_zero;  // smallest 1-based column of a zero pivot, INT_MAX if none
  const int b = blockIdx.x, t = threadIdx.x;
  const int c = t & (GJ_NB - 1), r0 = t / GJ_NB;
  const double2* A11 = A + (int64_t)b * sa + (int64_t)kb * lda + kb;
  double2 x[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int r = r0 + RS * q;
    const bool in = r < nbc && c < nbc;
    x[q] = A11[in ? (int64_t)r * lda + c : 0];
    if (!in) x[q] = make_double2(r == c ? 1.0 : 0.0, 0.0);
  }
  if (t == 0) first_zero = INT_MAX;
  __syncthreads();
#pragma unroll 1
  for (int p = 0; p < GJ_NB; ++p) {
    const int pb = p & 1;
    if ((p % RS) == r0) {  // this wave holds row p, as element p / RS
      const int qp = p / RS;
      double2 v = x[0];
#pragma unroll
      for (int q = 1; q < RPT; ++q)
        if (q == qp) v = x[q];
      const double apx = readlane_f64(v.x, p), apy = readlane_f64(v.y, p);
      const bool zero = apx == 0.0 && apy == 0.0;
      if (zero && c == 0 && p < nbc) atomicMin(&first_zero, kb + p + 1);
      // 1 / a(p,p) = conj(a(p,p)) / |a(p,p)|^2: one real division
      const double inv2 = zero ? 0.0 : 1.0 / (apx * apx + apy * apy);
      const double2 rp = make_double2(apx * inv2, -apy * inv2);
      rowb[pb][c] = (c == p) ? make_double2(1.0 + rp.x, rp.y) : cmul(v, rp);
    }
    if (c == p) {
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        const int i = r0 + RS * q;
        colb[pb][i] = make_double2(x[q].x - (i == p ? 1.0 : 0.0), x[q].y);
      }
    }
    __syncthreads();
    const double2 sc = rowb[pb][c];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const double2 aip = colb[pb][r0 + RS * q];
      x[q].x = fma(-aip.x, sc.x, fma(aip.y, sc.y, x[q].x));
      x[q].y = fma(-aip.x, sc.y, fma(-aip.y, sc.x, x[q].y));
    }
  }
#pragma unroll
  for (int q = 0; q < RPT; ++q) D[((int64_t)b * GJ_NB + r0 + RS * q) * GJ_NB + c] = x[q];
  if (t == 0 && first_zero != INT_MAX && info[b] == 0) info[b] = first_zero;
}

}  // namespace

namespace euvl {

int zgetrf(euvl_handle h, int n, double2* A, int64_t lda, int64_t sa, int* ipiv, int batch, int* d_info) {
  PrecScope scope(h, EUVL_PREC_FP32_LU);
  const double2 one = make_double2(1.0, 0.0), mone = make_double2(-1.0, 0.0), zero = make_double2(0.0, 0.0);
  // scratch: two candidate-row lists (batch x n ints), LU11 and top blocks
  // (batch x NB x NB complex each), the moved-row lists and the last round's
  // multipliers (batch x groups x TS x NB complex)
  const size_t b_rows = sizeof(int) * (size_t)batch * n;
  const size_t b_blk = sizeof(double2) * (size_t)batch * NB * NB;
  const size_t b_mv = sizeof(int) * (size_t)batch * (1 + 4 * NB);
  const int gstride = (n + TSZ - 1) / TSZ;  // groups of the first round
  const size_t b_l = sizeof(double2) * (size_t)batch * gstride * TSZ * NB;
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  char* scr = (char*)aux_reserve(h, 2 * al(b_rows) + 2 * al(b_blk) + al(b_mv) + al(b_l));
  if (!scr) {
    set_error("zgetrf: panel scratch allocation failed");
    return EUVL_ERR_NOMEM;
  }
  int* rows[2] = {(int*)scr, (int*)(scr + al(b_rows))};
  double2* lu11 = (double2*)(scr + 2 * al(b_rows));
  double2* top = (double2*)(scr + 2 * al(b_rows) + al(b_blk));
  int* moved = (int*)(scr + 2 * al(b_rows) + 2 * al(b_blk));
  double2* lscr = (double2*)(scr + 2 * al(b_rows) + 2 * al(b_blk) + al(b_mv));
  hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(256), 0, h->stream, d_info, batch);
  EUVL_LAUNCH_CHECK();
  for (int kb = 0; kb < n; kb += NB) {
    const int nbc = (n - kb < NB) ? n - kb : NB;
    const int m = n - kb;
    // tournament rounds until one group is left
    const int* in = nullptr;
    int n_in = m, cur = 0;
    for (;;) {
      const int groups = (n_in + TSZ - 1) / TSZ;
      const int last = groups == 1;
      hipLaunchKernelGGL(tslu_round_kernel<TSZ>, dim3(groups, batch), dim3(TSZ), 0, h->stream, (const double2*)A, lda,
                         sa,
                         n, kb, nbc, in, n_in, rows[cur], last, lu11, top, lscr, gstride, ipiv, moved, d_info);
      EUVL_LAUNCH_CHECK();
      if (last) break;
      const int tail = n_in - (groups - 1) * TSZ;
      n_in = (groups - 1) * nbc + (tail < nbc ? tail : nbc);
      in = rows[cur];
      cur ^= 1;
    }
    hipLaunchKernelGGL(tslu_apply_kernel, dim3((m + 255) / 256, batch), dim3(256), 0, h->stream, A, lda, sa, kb, m,
                       nbc, (const double2*)lu11, (const double2*)top, (const int*)moved);
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
      hipLaunchKernelGGL(trsm_kernel<false>, dim3((rest + TRSM_COLS - 1) / TRSM_COLS, batch), dim3(256), 0,
                         h->stream, nbc, A11, lda, sa, A12, lda, sa, rest);
      EUVL_LAUNCH_CHECK();
      EUVL_TRY(zgemm(h, rest, rest, nbc, mone, A21, lda, sa, A12, lda, sa, one, A22, lda, sa, zero, batch));
    }
  }
  return EUVL_OK;
}

int zgetrf_nopiv(euvl_handle h, int n, double2* A, int64_t lda, int64_t sa, int* ipiv, int batch, int* d_info) {
  PrecScope scope(h, EUVL_PREC_FP32_LU);
  const double2 one = make_double2(1.0, 0.0), mone = make_double2(-1.0, 0.0), zero = make_double2(0.0, 0.0);
  // scratch: U11^-1 and L11^-1 (batch x NB x NB each)
  const size_t b_blk = sizeof(double2) * (size_t)batch * NB * NB;
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  char* scr = (char*)aux_reserve(h, 2 * al(b_blk));
  if (!scr) {
    set_error("zgetrf_nopiv: scratch allocation failed");
    return EUVL_ERR_NOMEM;
  }
  double2* uinv = (double2*)scr;
  double2* linv = (double2*)(scr + al(b_blk));
  hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(256), 0, h->stream, d_info, batch);
  EUVL_LAUNCH_CHECK();
  for (int kb = 0; kb < n; kb += NB) {
    const int nbc = (n - kb < NB) ? n - kb : NB;
    hipLaunchKernelGGL(diag_lu_kernel, dim3(batch), dim3(DLU_THREADS), 0, h->stream, A, lda, sa, kb, nbc, uinv, linv, ipiv,
                       n, d_info);
    EUVL_LAUNCH_CHECK();
    const int rest = n - kb - nbc;
    if (rest <= 0) continue;
    double2* A11 = A + (int64_t)kb * lda + kb;
    double2* A12 = A11 + nbc;
    double2* A21 = A11 + (int64_t)nbc * lda;
    double2* A22 = A21 + nbc;
    // L21 = A21 U11^-1 and U12 = L11^-1 A12 in place: with nbc <= 32 columns
    // (rows) every GEMM tile spans all of them, so each workgroup reads the
    // rows (columns) it overwrites and no other, and it has read all of them
    // (K = nbc, at most two slabs) before its epilogue writes
    EUVL_TRY(zgemm(h, rest, nbc, nbc, one, A21, lda, sa, uinv, NB, (int64_t)NB * NB, zero, A21, lda, sa, zero, batch));
    EUVL_TRY(zgemm(h, nbc, rest, nbc, one, linv, NB, (int64_t)NB * NB, A12, lda, sa, zero, A12, lda, sa, zero, batch));
    EUVL_TRY(zgemm(h, rest, rest, nbc, mone, A21, lda, sa, A12, lda, sa, one, A22, lda, sa, zero, batch));
  }
  return EUVL_OK;
}

// A X = B by block Gauss-Jordan elimination without row interchanges, on the
// augmented matrix [A | B] (n x (n + nrhs), leading dimension lda >= n + nrhs,
// batch stride sa): block row k (GJ_NB rows) is multiplied by the inverse of
// its diagonal block, then eliminated from the rows below (one GEMM with
// K = GJ_NB over the rest of the matrix and the right-hand sides), and a block
// back substitution leaves X in columns n .. n + nrhs.  Compared with a
// factorisation followed by two triangular solves it has one dependent chain
// of n / GJ_NB short steps (inverse, scaling, update) instead of a 32-wide
// panel chain per factor and per solve, and every update is one GEMM.  The
// left n columns are overwritten (no factors are kept).  info[b] receives the
// first zero pivot of a diagonal block (1-based column), 0 if none.
size_t gj_solve_scratch_bytes(int batch) { return sizeof(double2) * (size_t)batch * GJ_NB * GJ_NB; }

int gj_solve(euvl_handle h, int n, int nrhs, double2* A, int64_t lda, int64_t sa, int batch, int* d_info) {
  constexpr int GJ_THREADS = 1024;
  PrecScope scope(h, EUVL_PREC_FP32_LU);
  const double2 one = make_double2(1.0, 0.0), mone = make_double2(-1.0, 0.0), zero = make_double2(0.0, 0.0);
  if (lda < (int64_t)n + nrhs) {
    set_error("gj_solve: leading dimension %lld below n + nrhs = %d", (long long)lda, n + nrhs);
    return EUVL_ERR_ARG;
  }
  double2* dinv = (double2*)aux_reserve(h, gj_solve_scratch_bytes(batch));
  if (!dinv) {
    set_error("gj_solve: scratch allocation failed");
    return EUVL_ERR_NOMEM;
  }
  hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(256), 0, h->stream, d_info, batch);
  EUVL_LAUNCH_CHECK();
  const int ncols = n + nrhs;
  for (int kb = 0; kb < n; kb += GJ_NB) {
    const int nbc = std::min(GJ_NB, n - kb);
    hipLaunchKernelGGL(gj_inv_kernel<GJ_THREADS>, dim3(batch), dim3(GJ_THREADS), 0, h->stream, (const double2*)A, lda,
                       sa, kb, nbc, dinv, d_info);
    EUVL_LAUNCH_CHECK();
    // block row k <- D_k [A_k,(k+1:) | B_k], in place: every tile of the
    // product spans its nbc <= GJ_NB rows
    double2* rowk = A + (int64_t)kb * lda + kb + nbc;
    const int right = ncols - kb - nbc;
    EUVL_TRY(zgemm(h, nbc, right, nbc, one, dinv, GJ_NB, (int64_t)GJ_NB * GJ_NB, rowk, lda, sa, zero, rowk, lda, sa,
                   zero, batch));
    const int below = n - kb - nbc;
    if (below > 0)
      EUVL_TRY(zgemm(h, below, right, nbc, mone, A + (int64_t)(kb + nbc) * lda + kb, lda, sa, rowk, lda, sa, one,
                     rowk + (int64_t)nbc * lda, lda, sa, zero, batch));
  }
  // back substitution by columns, last block first: X_k = B'_k is final once
  // the blocks after it are done, and the rows above subtract U'_(0:k, k) X_k
  // (one GEMM with K = GJ_NB over all of them, so each step fills the chip)
  const int last = ((n - 1) / GJ_NB) * GJ_NB;
  for (int kb = last; kb > 0; kb -= GJ_NB) {
    const int nbc = std::min(GJ_NB, n - kb);
    EUVL_TRY(zgemm(h, kb, nrhs, nbc, mone, A + kb, lda, sa, A + (int64_t)kb * lda + n, lda, sa, one, A + n, lda, sa,
                   zero, batch));
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
    hipLaunchKernelGGL(trsm_kernel<false>, dim3((nrhs + TRSM_COLS - 1) / TRSM_COLS, batch), dim3(256), 0, h->stream,
                       nbc, L11, lda, sa, Bk, ldb, sb, nrhs);
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
    hipLaunchKernelGGL(trsm_kernel<true>, dim3((nrhs + TRSM_COLS - 1) / TRSM_COLS, batch), dim3(256), 0, h->stream,
                       nbc, U11, lda, sa, Bk, ldb, sb, nrhs);
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

extern "C" int euvl_zgetrf_nopiv_batched(euvl_handle h, int n, double* d_A, int64_t lda, int64_t sa, int* d_ipiv,
                                         int batch) {
  if (!h || n < 0 || batch < 0 || batch > 1024 || lda < n) {
    euvl::set_error("euvl_zgetrf_nopiv_batched: bad argument");
    return EUVL_ERR_ARG;
  }
  if (n == 0 || batch == 0) return EUVL_OK;
  int* d_info = (int*)h->d_red;
  EUVL_TRY(euvl::zgetrf_nopiv(h, n, (double2*)d_A, lda, sa, d_ipiv, batch, d_info));
  int* hi = (int*)h->h_pinned;
  EUVL_HIP_CHECK(hipMemcpyAsync(hi, d_info, sizeof(int) * batch, hipMemcpyDeviceToHost, h->stream));
  EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));
  for (int b = 0; b < batch; ++b)
    if (hi[b] != 0) {
      euvl::set_error("euvl_zgetrf_nopiv_batched: matrix %d has a zero pivot at column %d", b, hi[b] - 1);
      return EUVL_ERR_SINGULAR;
    }
  return EUVL_OK;
}

extern "C" int euvl_zgesv_nopiv_batched(euvl_handle h, int n, int nrhs, double* d_A, int64_t lda, int64_t sa,
                                       int batch) {
  if (!h || n <= 0 || nrhs <= 0 || !d_A || lda < (int64_t)n + nrhs || batch <= 0 || batch > 256 ||
      (batch > 1 && sa < (int64_t)n * lda)) {
    euvl::set_error("euvl_zgesv_nopiv_batched: bad argument");
    return EUVL_ERR_ARG;
  }
  int* info = (int*)h->d_red;  // batch <= 256 ints of the handle's reduction slots
  EUVL_TRY(euvl::gj_solve(h, n, nrhs, (double2*)d_A, lda, sa, batch, info));
  int* hi = (int*)h->h_pinned;
  EUVL_HIP_CHECK(hipMemcpyAsync(hi, info, sizeof(int) * batch, hipMemcpyDeviceToHost, h->stream));
  EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));
  for (int b = 0; b < batch; ++b)
    if (hi[b] != 0) {
      euvl::set_error("euvl_zgesv_nopiv_batched: matrix %d has a zero pivot at column %d", b, hi[b] - 1);
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
