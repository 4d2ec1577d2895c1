// Internal interfaces shared between the libeuvl translation units.
#pragma once
#include "common.h"
#include "optics.h"

namespace euvl {

int zgemm(euvl_handle h, int m, int n, int k, double2 alpha, const double2* A, int64_t lda, int64_t sa,
          const double2* B, int64_t ldb, int64_t sb, double2 beta, double2* C, int64_t ldc, int64_t sc,
          double2 gamma, int batch);
// zgemm whose batch entry b reads A entry b mod per_a and B entry b mod per_b
// (0: no period), so that products sharing an operand run as one launch
int zgemm_periodic(euvl_handle h, int m, int n, int k, double2 alpha, const double2* A, int64_t lda, int64_t sa,
                   const double2* B, int64_t ldb, int64_t sb, double2 beta, double2* C, int64_t ldc, int64_t sc,
                   double2 gamma, int batch, int per_a, int per_b);
int zgetrf(euvl_handle h, int n, double2* A, int64_t lda, int64_t sa, int* ipiv, int batch, int* d_info);
int zgetrs(euvl_handle h, int n, int nrhs, const double2* LU, int64_t lda, int64_t sa, const int* ipiv,
           double2* B, int64_t ldb, int64_t sb, int batch);

// X[b] += g I for the batch entries b < batch (the diagonal only)
int ew_diag_add(euvl_handle h, int n, int batch, double2* X, double2 g);
// the six Paterson-Stockmeyer blocks of cos (c) and sinc (f) from D..D^4 (ps_blocks_kernel)
int ew_ps_blocks(euvl_handle h, int n, int batch, const double2* X1, const double2* X2, const double2* X3,
                 const double2* X4, const double* c, const double* f, double2* Hc, double2* Qc, double2* Oc,
                 double2* Hf, double2* Qf, double2* Of);
// the two blocks per polynomial in powers of D^6 from X[0..5] = D..D^6 (ps6_blocks_kernel):
// q = O + D^6 T; Tc and Tf may be the buffers of X[1] and X[2]
int ew_ps6_blocks(euvl_handle h, int n, int batch, const double2* const* X, const double* c, const double* f,
                  double2* Oc, double2* Tc, double2* Of, double2* Tf);
int ew_copy(euvl_handle h, int64_t count, double2* out, const double2* X);
// out[0:count) = v (a fill kernel: safe inside graph capture, unlike hipMemsetAsync)
int fill(euvl_handle h, int64_t count, double2* out, double2 v);
// kind 0: max|X-I| (|re|+|im|), 1: infinity norm (max row sum of |x|), 2: mean Re diag
int reduce(euvl_handle h, int kind, int n, int batch, const double2* X, double* d_red);
// max |re| + |im| over an m x ncols block of each batch entry (leading dim ld, batch stride sb) into
// d_slot[0:batch)
int maxabs(euvl_handle h, int m, int ncols, int64_t ld, int64_t sb, int batch, const double2* X, double* d_slot);
// LU without row interchanges (blocked, GEMM panels); ipiv is written as the identity.  Only for matrices
// whose unpivoted elimination is safe; callers verify what they solve with it.
int zgetrf_nopiv(euvl_handle h, int n, double2* A, int64_t lda, int64_t sa, int* ipiv, int batch, int* d_info);
// A X = B on the augmented [A | B] (n x (n + nrhs), lda >= n + nrhs) by block
// Gauss-Jordan elimination without row interchanges; X replaces B, the left
// n columns are overwritten; info[b] = first zero pivot (1-based) or 0.
// device scratch (handle's aux buffer) of gj_solve for `batch` systems
size_t gj_solve_scratch_bytes(int batch);
int gj_solve(euvl_handle h, int n, int nrhs, double2* A, int64_t lda, int64_t sa, int batch, int* d_info);

// Device-side checks of decisions fixed ahead of time (graph-capturable runs):
// flags[b] |= bit where the check fails (see flag_kernel in ew.hip).
int flag_check(euvl_handle h, int kind, int batch, const double* a, const double* c, const double* d, double s,
               double theta, int bit, int* flags);
int flag_info(euvl_handle h, int batch, const int* info, int bit, int* flags);
int zero_ints(euvl_handle h, int count, int* p);

// flag bits of the fixed-decision EM solve
constexpr int FLAG_DOUBLINGS = 1;  // the fixed doubling count is below what ||D^3||, ||D^4|| ask for
// bit 2 is not used (the solve has no series whose length is fixed)
constexpr int FLAG_LU_RESIDUAL = 4;  // the unpivoted top solve failed its residual check
constexpr int FLAG_ZERO_PIVOT = 8;   // the unpivoted LU met a zero pivot

// matrix functions (matfun.hip); buffers are batch x n x n contiguous.
// C = cos(d sqrt D), F = sin(d sqrt D)/sqrt D, G2 = D F (G2 may be null); takes 6 workspace matrices.
// fixed_s < 0: the doubling count is chosen from ||D^3||, ||D^4|| read back to the host (*doublings
// receives it).  fixed_s >= 0: that count is used without a host read-back; the norms go to
// d_red[0, 2 batch) and entries whose bound needs more doublings get FLAG_DOUBLINGS in d_flags.
// xd (optional): D is the EM solve's absorber operator and its right products
// (the powers D^2..D^4 and G2 = F D, which equals D F) run by FFTs (xd_fft).
struct XdOp;
int cos_sinc(euvl_handle h, int n, int batch, const double2* D, double d, double2* C, double2* F, double2* G2,
             int* doublings, int fixed_s = -1, double* d_red = nullptr, int* d_flags = nullptr,
             double* alpha_out = nullptr, const XdOp* xd = nullptr);
// the theta of the doubling bound in the current precision scope
double cos_sinc_theta(euvl_handle h);

// The EM solve core (em.hip).
struct SolveHost {  // host tables, entry b = clip * n_off + offset
  std::vector<int> inc_idx, inc_slot;
  std::vector<double> sx, sy, kz;
};
struct SolveDev {  // the same tables on the device
  const int* lidx = nullptr;
  const int* midx = nullptr;
  const double* sx0 = nullptr;
  const double* sy0 = nullptr;
  const double2* leps = nullptr;
  const double* ld = nullptr;
  const int* inc_idx = nullptr;
  const int* inc_slot = nullptr;
  const double* kz = nullptr;
  int ninc = 0;
};
struct SolveCtl {
  bool fixed = false;  // decisions fixed ahead of time, checked on the device
  int doublings = 0;
  int* d_flags = nullptr;  // per batch entry, FLAG_* bits (fixed mode)
  int doublings_used = 0, lu_path = 0;  // what ran
  double alpha = 0.0;  // max(||D^3||^(1/3), ||D^4||^(1/4)) over the batch (adaptive runs)
};
void build_solve_tables(const euvl_em_params* p, const int* lindex, const int* mindex, int N, int lmax, int mmax,
                        int n_off, const double* sx0, const double* sy0, int n_clips, SolveHost& t);
size_t solve_tables_bytes(const euvl_em_params* p, int N, const SolveHost& t);
int upload_solve_tables(euvl_handle h, const euvl_em_params* p, const int* lindex, const int* mindex, int N,
                        const SolveHost& t, Upload& up, char* blk, SolveDev& d);
size_t solve_workspace_bytes(int N, int E, int ninc);
// The coupling product V -= H G of the top solve by 2D FFTs over the order grid
// (fft.hip).  FftGrid: points along l and m (no prime factor above 3, at least
// 4 lmax + 1 and 4 mmax + 1) and the radices of their Stockham passes;
// fft_grid_for is false when the grid does not fit the workgroup's LDS (the
// product then stays a GEMM).  S: coupling_fft_bytes(E) of scratch.
constexpr int FFT_MAX_STAGES = 8;
struct FftGrid {
  int P[2];
  int nst[2];
  int rad[2][FFT_MAX_STAGES];
};
bool fft_grid_for(int lmax, int mmax, FftGrid& g);
size_t coupling_fft_bytes(int E);
int coupling_fft(euvl_handle h, const FftGrid& g, int N, int E, const int* lidx, const int* midx, int lmax, int mmax,
                 const double2* W, int per_clip, int64_t w_stride, double2 ea_inv, const double* kt, double k,
                 const double2* H, double2* V, int64_t ldv, int64_t sv, double2* S);
// Right products Y = X D by the absorber operator D of the EM solve (the
// matrix of toeplitz_kernel), by 2D FFTs over the order grid (fft.hip).  With
// m^ the clip's indicator spectrum, u the order index along the
// polarisation's axis (l for X, m for Y) and c = 2 pi / dx (dy),
//   D[i][j] = k^2 (1 - ea) m^(i - j) + k^2 ea delta_ij
//             - c log(ea) (u_i - u_j) m^(i - j) kt_j - delta_ij kxy2_j,
// so row r of X D needs the correlations of x = X[r][:] and of u x with m^.
// a_p = k^2 (1 - ea), b_u = c log(ea), a_x = k^2 ea; S: the spectra of m^
// (xd_fft_spectra, coupling_fft_bytes(E)); kt, kxy2: batch x N.
struct XdOp {
  FftGrid g;
  int N = 0, lmax = 0;
  const int* lidx = nullptr;
  const int* midx = nullptr;
  const int* uidx = nullptr;
  const double* kt = nullptr;
  const double* kxy2 = nullptr;
  double2 a_p, b_u, a_x;
  const double2* S = nullptr;
};
bool xd_fft_fits(int N);
// smallest N for the FFT form of the products by D (coupling = false) or of H G (true)
int fft_min_n(bool coupling);
int xd_fft_spectra(euvl_handle h, const FftGrid& g, int lmax, int mmax, const double2* W, int per_clip,
                   int64_t w_stride, int E, double2* S);
int xd_fft(euvl_handle h, const XdOp& op, int E, const double2* X, double2* Y);
int amp_s_core(euvl_handle h, char polar, const euvl_em_params* p, const SolveDev& t, int N, int lmax, int mmax,
               const double2* spec, int per_clip, int64_t spec_stride, int E, double2* Ax, SolveCtl& ctl);
// indicator spectrum of n_clips packed masks on a window |l| <= lw, |m| <= mw, with the
// twiddle tables already on the device (tw: nx + 1 then ny + 1 entries); no synchronisation
int mask_spectrum_enqueue(euvl_handle h, const uint8_t* bits, int nx, int ny, int n_clips, int lw, int mw,
                          double2* out, double2* scratch, double2* part, const double2* tw);
int mask_spectrum_splits(euvl_handle h, const uint8_t* bits, int nx, int ny, int n_clips, int lw, int mw,
                         double2* out, double2* scratch, double2* part, const double2* tw, int S);
int dft_splits(int nx, int ny, int mw, int n_clips);
void dft_twiddles(int nx, int ny, std::vector<double2>& t);

}  // namespace euvl
