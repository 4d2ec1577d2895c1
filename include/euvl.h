/*
 * euvl.h -- C ABI of libeuvl.so, the MI355X (gfx950) implementation of the
 * EUVlitho mask -> rigorous EM -> M3D -> image path.
 *
 * Conventions
 *   - Every entry point returns an int status: EUVL_OK (0) or a negative code.
 *     The message of the last failure on the calling thread is returned by
 *     euvl_last_error().  Nothing aborts the process.
 *   - "d_" pointers are device memory owned by the caller; "h_" pointers are
 *     host memory read during the call only.  Complex numbers are interleaved
 *     (re, im) doubles; matrices are row-major.
 *   - Work is enqueued on the stream set with euvl_set_stream (default: the
 *     null stream).  Entry points that need a result on the host to continue
 *     (convergence checks, the expm scaling choice) synchronise that stream.
 *   - One handle must not be used by two host threads at once; handles own
 *     their device workspaces.
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the takahashi-edalab/EUVlitho tree).
 */
#ifndef EUVL_H
#define EUVL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EUVL_OK 0
#define EUVL_ERR_ARG -1          /* bad argument (size, pointer, polarisation) */
#define EUVL_ERR_HIP -2          /* HIP runtime failure (launch, allocation) */
#define EUVL_ERR_SINGULAR -3     /* zero pivot in an LU factorisation */
#define EUVL_ERR_NOT_CONVERGED -4 /* a series or iteration bound failed (cos/sinc norms, Jacobi sweeps) */
#define EUVL_ERR_NOMEM -5        /* workspace allocation failed */
#define EUVL_ERR_STALE -6        /* a graph's captured workspace was reallocated (capture it again) */

typedef struct euvl_handle_s* euvl_handle;

/* Handle on the current HIP device.  No reference counterpart: the reference
 * creates a cuSOLVER/cuBLAS handle, stream and workspace inside every call
 * (cnn/include/header.h:796-804, 1501-1508, 1590-1593). */
int euvl_create(euvl_handle* out);
int euvl_destroy(euvl_handle h);
/* Entry points return with their work queued on this stream (host tables go
 * through pinned staging buffers of the handle).  Changing the stream
 * synchronises the previous one first, since the next call reuses the
 * handle's workspace: the previous stream must still exist then. */
int euvl_set_stream(euvl_handle h, void* hip_stream);
const char* euvl_last_error(void);
/* Device bytes currently held by the handle's workspace. */
int64_t euvl_workspace_bytes(euvl_handle h);
/* Grow the handle's workspace to at least `bytes` now (so that later calls
 * do not reallocate), or free it (and the LU panel buffer) until the next
 * call needs it.  Both synchronise the handle's stream.  No reference
 * counterpart: they let a serving loop admit work by device-memory footprint. */
int euvl_reserve_workspace(euvl_handle h, int64_t bytes);
int euvl_release_workspace(euvl_handle h);
/* Per-launch timing of the GEMM kernel with HIP events on the handle's
 * stream.  If out is non-null, synchronises and writes {launches, total
 * milliseconds, algorithmic flops (8 m n k per complex GEMM), algorithmic
 * bytes (A and B read once, C written, and read too when beta != 0)} of the
 * launches recorded since the last call (out holds 4 doubles); then clears
 * them and enables (enable & 1) or disables recording.  With enable & 2, out
 * holds 8 doubles: the same four figures follow for the launches whose
 * products ran on FP32 MFMA.  No reference counterpart (instrumentation). */
int euvl_gemm_stats(euvl_handle h, int enable, double* out);

/* Compute precision of the products inside the EM solve (the precision
 * form).  mask = 0 (default) keeps every product in FP64 MFMA, the plain
 * drop-in.  Each set bit moves one group of products to FP32 MFMA with FP32
 * accumulation over K; operands are rounded to FP32 as they are staged, and
 * storage, the epilogue (alpha, beta, gamma) and every other kernel stay FP64:
 *   EUVL_PREC_FP32_MATFUN    the cos / sinc polynomial and its doublings
 *                            (the counterpart of the absorber eigenproblem)
 *   EUVL_PREC_FP32_COUPLING  the absorber / cap coupling and top-solve products
 *   EUVL_PREC_FP32_LU        trailing updates of the LU and its solves
 *   EUVL_PREC_FP32_GEMM      direct euvl_zgemm_batched calls
 * No reference counterpart (the reference is FP64 throughout). */
#define EUVL_PREC_FP32_MATFUN 1
#define EUVL_PREC_FP32_COUPLING 2
#define EUVL_PREC_FP32_LU 4
#define EUVL_PREC_FP32_GEMM 8
/* Complex products, FP64 or FP32-compute, are formed with 3 real MFMAs by
 * default (the zgemm3m form: Re = ArBr - AiBi,
 * Im = (Ar + Ai)(Br + Bi) - ArBr - AiBi, the differences taken in FP64), a
 * quarter fewer MFMAs than the textbook 4 with a rounding error bounded like
 * theirs in norm ((|Ar| + |Ai|)(|Br| + |Bi|) in place of |A||B|).
 * EUVL_PREC_GEMM_4M selects the 4-multiplication form in both precisions. */
#define EUVL_PREC_GEMM_4M 16
/* Image formation in FP32: the Abbe transforms (the y pass in FP32, the x pass
 * on FP32 MFMAs with Kahan-compensated sums of |E|^2 over the source points of
 * each split, the split sums in FP64) and the SOCS field transforms and kernel
 * sums (FP32, compensated).  Pupil fields and their propagation phases are
 * formed in FP64 and rounded. */
#define EUVL_PREC_FP32_IMAGING 32
int euvl_set_precision(euvl_handle h, int mask);

/* GEMM autotuner: times every kernel variant (tile shape, waves per
 * workgroup) of the m x n x k x batch product in FP64 (fp32 = 0) or
 * FP32-compute (fp32 = 1) on scratch operands, and keeps the fastest (of
 * those within 1 % of the fastest, the one with the largest tile) for that
 * size bucket (m, n, k rounded up to 64 below 256 and to 128 above; batch
 * exact) for every later GEMM of the process.  *out_variant receives the
 * choice; out_ms (optional, 7 doubles, indexed by variant number) the time
 * per variant, -1 for a number the product form has no variant for (5 in
 * all but the FP64 3M form, and there when 64 x 64 tiles would give 512
 * workgroups or more; 6 in the 4M forms).  Allocates and
 * frees its scratch; synchronises the handle's stream.  No reference
 * counterpart. */
int euvl_autotune_gemm(euvl_handle h, int m, int n, int k, int batch, int fp32, int* out_variant,
                       double* out_ms);

/* Smallest order count N at which the EM solve runs its products by the
 * absorber operator D (D^2..D^6 and D F; default 1200) and the coupling
 * product H G (default 800) as 2D FFT correlations over the order grid in
 * place of GEMMs, when the grid fits a workgroup's LDS; a negative value
 * leaves that threshold as it is.  Process-wide; the defaults can also be
 * set by EUVL_FFT_MIN_N and EUVL_COUPLING_FFT_MIN_N.  Both paths give the
 * same results to rounding.  No reference counterpart (tuning). */
int euvl_set_fft_min_n(int products, int coupling);

/* ---------------------------------------------------------------------------
 * Dense complex linear algebra (batched, row-major, interleaved complex).
 * ------------------------------------------------------------------------- */

/* C[b] = alpha * A[b] * B[b] + beta * C[b] + gamma * I, b < batch.
 * A: m x k (leading dim lda, batch stride sa), B: k x n, C: m x n.
 * Replaces matproduct (cnn/include/header.h:1559-1619: cublasZgemm OP_N/OP_N
 * above N = 5000, Eigen A*B below). alpha/beta/gamma are (re, im) pairs. */
int euvl_zgemm_batched(euvl_handle h, int m, int n, int k,
                       const double* alpha,
                       const double* d_A, int64_t lda, int64_t sa,
                       const double* d_B, int64_t ldb, int64_t sb,
                       const double* beta,
                       double* d_C, int64_t ldc, int64_t sc,
                       const double* gamma, int batch);

/* In-place LU with partial pivoting, A[b] = P L U (n x n, leading dim lda,
 * batch stride sa); d_ipiv[b*n + i] = row swapped with row i (0-based).
 * Replaces the factorisation half of matinv (cnn/include/header.h:1468-1557:
 * cusolverDnXgetrf above N = 5000, Eigen colPivHouseholderQr below).
 * Returns EUVL_ERR_SINGULAR if a pivot is exactly zero. */
int euvl_zgetrf_batched(euvl_handle h, int n, double* d_A, int64_t lda, int64_t sa,
                        int* d_ipiv, int batch);

/* In-place LU without row interchanges (d_ipiv receives the identity), for
 * matrices whose unpivoted elimination is stable (diagonally dominant, or
 * well conditioned with small growth, as the EM top-solve operator): each
 * 32-wide panel is its diagonal block factored in LDS plus GEMMs, with no
 * pivot search.  Returns EUVL_ERR_SINGULAR on an exactly zero pivot; it does
 * not detect growth, so callers verify what they solve with it.  No reference
 * counterpart (the reference always pivots). */
int euvl_zgetrf_nopiv_batched(euvl_handle h, int n, double* d_A, int64_t lda, int64_t sa, int* d_ipiv,
                              int batch);

/* A[b] X = B[b] without row interchanges, on the augmented matrix [A | B]
 * (n x (n + nrhs), row-major, leading dim lda >= n + nrhs, batch stride sa):
 * X replaces B in columns n .. n + nrhs - 1, the left n columns are
 * overwritten.  Block Gauss-Jordan elimination (the inverse of each 64 x 64
 * diagonal block, GEMM updates with K = 64); for matrices whose elimination
 * without pivoting is stable (diagonally dominant, or verified by the caller's
 * residual).  Returns EUVL_ERR_SINGULAR on a zero pivot of a diagonal block.
 * No reference counterpart as a call: the top solve of ampS (matinv,
 * cnn/include/header.h:1468-1557 with the products of :1355-1391) uses it on
 * its reduced system.  batch <= 256. */
int euvl_zgesv_nopiv_batched(euvl_handle h, int n, int nrhs, double* d_A, int64_t lda, int64_t sa, int batch);

/* Solve A[b] X = B[b] in place of B (n x nrhs, leading dim ldb) using the
 * factors of euvl_zgetrf_batched.  Replaces cusolverDnXgetrs in matinv
 * (cnn/include/header.h:1530-1535). */
int euvl_zgetrs_batched(euvl_handle h, int n, int nrhs, const double* d_LU, int64_t lda, int64_t sa,
                        const int* d_ipiv, double* d_B, int64_t ldb, int64_t sb, int batch);

/* Layer propagation functions of the absorber operator D[b] (n x n, batch
 * entries contiguous with stride n*n):
 *   C = cos(d sqrt(D)),  F = sin(d sqrt(D)) / sqrt(D),  G2 = D F
 * (d = thickness; d_G2 may be null).  Even in sqrt(D), so they are power
 * series in D: evaluated by a degree-9 Taylor series in (d/2^s)^2 D and s
 * angle doublings (*h_doublings receives s).  Replaces the eigen-pairs,
 * al1 = sqrt(lambda) and gamma = exp(i al1 dabs) of absorberS0
 * (cnn/include/header.h:834-921, 1085-1140): every use of (V, al1, gamma) in
 * ampS reduces to these three matrices, with no square-root branch choice. */
int euvl_zcos_sinc_batched(euvl_handle h, int n, const double* d_D, double thickness, double* d_C,
                           double* d_F, double* d_G2, int batch, int* h_doublings);

/* ---------------------------------------------------------------------------
 * Rigorous EM solve.
 * ------------------------------------------------------------------------- */

/* Optics and stack shared by the EM entry points (mirrors the scalar
 * arguments of ampS, cnn/include/header.h:1208-1213, and its material table
 * header.h:1217-1230 given as an explicit layer list). */
typedef struct {
  int ndivx, ndivy;        /* mask pixels (1 nm pixels: FDIVX = NDIVX) */
  int mx, my;              /* magnification (MX, MY) */
  double na, wavelength, k;/* NA, lambda (nm), k = 2 pi / lambda */
  double dx, dy;           /* pitch (nm) */
  int lsmax_x, lsmax_y;    /* incidence lattice half-widths (lsmaxX, lsmaxY) */
  double eabs_re, eabs_im; /* absorber permittivity (NABS = 1) */
  double dabs;             /* absorber thickness (nm) */
  int n_layers;            /* multilayer layers, bottom (substrate) to top (cap) */
  const double* h_layer_eps; /* 2*n_layers: (re, im) permittivity per layer */
  const double* h_layer_d;   /* n_layers thicknesses (layer 0 unused) */
} euvl_em_params;

/* Fourier coefficients of the binary mask indicator m (1 = vacuum) on the
 * window |l| <= lw, |m| <= mw:  out[(l+lw)*(2mw+1) + (m+mw)] =
 * (1/(NDIVX NDIVY)) sum_{i,j} m(i,j) exp(-2 pi i (i l/NDIVX + j m/NDIVY)),
 * with x = i and y = j counted from the bottom row (the y flip of
 * header.h:1157-1168).  d_bits: packed mask.bin bytes of n_clips clips.
 * out is n_clips x (2lw+1) x (2mw+1) complex.
 * Replaces maskamp + fourier + exponential (cnn/include/header.h:1147-1206,
 * 1621-1648) and the thin-mask spectrum of m3d.cpp:287-327: eps, 1/eps and
 * log(eps) tables are affine in this one spectrum. */
int euvl_mask_spectrum(euvl_handle h, const uint8_t* d_bits, int ndivx, int ndivy, int n_clips,
                       int lw, int mw, double* d_out);

/* euvl_mask_spectrum without any allocation: no reference counterpart; it is
 * for callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results (to rounding:
 * the y sums may be split into fewer partial sums).  Its scratch is one fixed
 * device array of euvl_mask_spectrum_noalloc_max() bytes (64 MiB) per GPU,
 * held until the process ends; requests that need more (twiddles plus
 * n_clips x ndivx x (2mw+1) complex) fail with EUVL_ERR_ARG.  Every call on a
 * GPU shares that array: calls that overlap on one GPU (several threads or
 * streams) race and must be serialised by the caller.  The current device
 * must be the handle's. */
int euvl_mask_spectrum_noalloc(euvl_handle h, const uint8_t* d_bits, int ndivx, int ndivy, int n_clips, int lw,
                               int mw, double* d_out);
int64_t euvl_mask_spectrum_noalloc_max(void);

/* Diagonal T-matrix cascade of the multilayer for n modes with transverse
 * wave vectors (kx, ky) (device arrays of n doubles), polarisation 'X' or 'Y'.
 * Writes d_uu, d_ub (n complex each).  Replaces multilayerS
 * (cnn/include/header.h:5-363). */
int euvl_multilayer(euvl_handle h, char polar, const euvl_em_params* p, int n,
                    const double* d_kx, const double* d_ky, double* d_uu, double* d_ub);

/* ampS for one clip and n_off incidence offsets (sx0[o], sy0[o]) at once
 * (batched over offsets).  Order set (h_lindex, h_mindex) of length nrange
 * with |l| <= lmax, |m| <= mmax.  d_spec: the clip's indicator spectrum from
 * euvl_mask_spectrum with lw = 2 lmax, mw = 2 mmax.
 * Output d_Ax: n_off x (2 lsmax_x+1) x (2 lsmax_y+1) x nrange complex;
 * the entry [o][ls+lsmax_x][ms+lsmax_y][:] is Ax of header.h:1394-1465 for
 * the lattice incidence (ls, ms) inside the NA and 0 elsewhere.
 * h_info (optional, 4 ints per offset): LU path of the top solve (0: without
 * row interchanges, verified by its residual; 1: the partial-pivoting
 * fallback), cos/sinc doublings, 0 (reserved), incidence columns solved.
 * Replaces ampS (cnn/include/header.h:1208-1466) with absorberS0
 * (header.h:754-1145) and the top solve (header.h:1355-1391). */
int euvl_amp_s(euvl_handle h, char polar, const euvl_em_params* p,
               const int* h_lindex, const int* h_mindex, int nrange, int lmax, int mmax,
               const double* d_spec, int n_off, const double* h_sx0, const double* h_sy0,
               double* d_Ax, int* h_info);

/* ---------------------------------------------------------------------------
 * Image formation.
 * ------------------------------------------------------------------------- */

/* Rigorous Abbe image accumulation for n_src source points of one
 * polarisation from EM amplitudes of that polarisation, the per-source-point
 * loop of intensity.cpp:276-388 (field, rotate(), defocus phase, unnormalised
 * inverse DFT, |Ex|^2+|Ey|^2+|Ez|^2).  Adds weight * intensity into d_img
 * (nx x ny doubles, index x*ny + y).
 *   polar: 'X' (emint.csv: Ax = amp/sqrt(k^2-kx^2), column 0 of rotate()) or
 *          'Y' (emintY.csv: Ay = amp/sqrt(k^2-ky^2), column 1)
 *   d_Ax: amplitudes, n_off x nsx x nsy x nrange complex (euvl_amp_s layout)
 *   h_src: n_src x 3 ints (offset index o, l0s, m0s); h_off: per offset
 *          (sx0, sy0, fx, fy) with fx = nsx/ndivs, fy = nsy/ndivs
 *   h_pupil: ncut x 2 ints (linput, minput)
 *   h_modemap: (2 lmax+1)(2 mmax+1) ints, mode index of (l, m) or -1
 *   kx0, ky0: incidence of rotate(); z0, z: defocus terms of intensity.cpp:353 */
int euvl_abbe_image(euvl_handle h, char polar, const euvl_em_params* p, const double* d_Ax, int nrange,
                    int lmax, int mmax, const int* h_modemap,
                    int n_off, const double* h_off, int n_src, const int* h_src,
                    int ncut, const int* h_pupil, double kx0, double ky0, double z0, double z,
                    int nx, int ny, double weight, double* d_img);

/* M3D parameters (a0, ax, ay) of one clip, m3d.cpp:329-467: thin-mask
 * spectrum d_fmask (noutx x nouty complex, euvl_mask_spectrum with lw =
 * lpmax_x, mw = lpmax_y), rigorous amplitudes d_Ax of offset 0
 * (nsx x nsy x nrange), blank amplitudes h_vc, h_ab (nsx x nsy complex, host),
 * pupil points h_pupil (ncut x 3 ints: linput, minput, xinput flag).
 * Outputs d_a0, d_ax, d_ay: ncut complex each (ax = ay = 0 where xinput = 0). */
int euvl_m3d_fit(euvl_handle h, const euvl_em_params* p, const double* d_fmask, int lpmax_x, int lpmax_y,
                 const double* d_Ax, int nrange, int lmax, int mmax, const int* h_modemap,
                 const double* h_vc, const double* h_ab, double kx0, double ky0, double z0,
                 int ncut, const int* h_pupil, double* d_a0, double* d_ax, double* d_ay);

/* Thin-mask and M3D-linear Abbe images on the source mesh (h_src: n_src x 2
 * doubles, the wave-number offsets dsx, dsy of linear.py:46-97).  Thin
 * amplitude from the indicator spectrum d_fmask (noutx x nouty, lw = lpmax_x,
 * mw = lpmax_y): fmask = ampab delta + (ampvc - ampab) m^, times the z0
 * phases, over phase0 (linear.py:120-143); pupil fields with
 * rotate() and the central obscuration co (electro_field.py:162-189); image
 * = mean over sources of |Ex|^2+|Ey|^2+|Ez|^2 on nx x ny (linear.py:10-43).
 * ampvc, ampab, phase0: host (re, im).  d_a0, d_ax, d_ay: ncut complex (ax,
 * ay zero where the pupil point has fewer than 8 sources).  Writes
 * d_intlinear (linear model) and d_intft (thin mask), nx x ny, index x*ny+y.
 * Replaces int_linear + int_abbe (cnn/intensity/linear/linear.py:10-188). */
int euvl_linear_images(euvl_handle h, const euvl_em_params* p, const double* d_fmask, int lpmax_x, int lpmax_y,
                       const double* ampvc, const double* ampab, const double* phase0, const double* d_a0,
                       const double* d_ax, const double* d_ay, int ncut, const int* h_pupil, int n_src,
                       const double* h_src, double kx0, double ky0, double z0, double z, double co, int nx, int ny,
                       double* d_intlinear, double* d_intft);

/* STCC kernels: TCC_{S0,SX,SY} = sum_c P_c diag(w) P_c^H / S over the source
 * mesh (w = k^2/(k^2-(kx0+sx)^2) times 1, sx, sy), Hermitian eigenpairs by a
 * parallel Jacobi method, sorted by |alpha| descending; the leading nk of
 * each: d_alpha (3 x nk doubles), d_phi (3 x nk x ncut complex, kernel m of
 * set b at [b][m][:]).  For nk <= 24 and ncut >= 128 the leading pairs come
 * from block subspace iteration (64 columns, Rayleigh-Ritz by the same
 * Jacobi method on 64 x 64) once every leading residual |T f - alpha f| is
 * below 1e-13 |alpha_1|; otherwise, or when that does not converge, from the
 * full decomposition.  *h_sweeps receives the Jacobi sweep count of the full
 * decomposition, or minus the subspace iterations.
 * Replaces eigenfunction() (cnn/intensity/stccsocs/eigenfunction.py:5-179)
 * and its cupy.linalg.eigh calls (:130-132). */
int euvl_tcc_kernels(euvl_handle h, const euvl_em_params* p, int ncut, const int* h_pupil, int n_src,
                     const double* h_src, double kx0, double ky0, double z0, double z, int nk, double* d_alpha,
                     double* d_phi, int* h_sweeps);

/* STCC-SOCS image (socs.py:12-154): nsocs kernels for a0 and nsocsxy for
 * ax/ay on an nout x nout grid, Fourier-interpolated to nx x ny (d_img,
 * index x*ny+y).  Amplitude inputs as euvl_linear_images; kernels from
 * euvl_tcc_kernels.  Replaces intsocs (cnn/intensity/stccsocs/socs.py:12-154). */
int euvl_socs_image(euvl_handle h, const euvl_em_params* p, const double* d_fmask, int lpmax_x, int lpmax_y,
                    const double* ampvc, const double* ampab, const double* phase0, const double* d_a0,
                    const double* d_ax, const double* d_ay, int ncut, const int* h_pupil, const double* d_alpha,
                    const double* d_phi, int nk, int nsocs, int nsocsxy, int nout, double kx0, double ky0,
                    double z0, int nx, int ny, double* d_img);

/* ---- mask producers (cnn/intensity/mask/, elitho/mask_pattern.py) ----
 * Masks are written packed as mask.bin: n_masks clips of ndivx*ndivy/8 bytes,
 * MSB-first, bit index ndivx*j + i, 1 = vacuum.  ndivx must be a multiple
 * of 8. */

/* The program cnn/intensity/mask/mask.cpp:28-62 with maskgen/randmask of
 * mVBF.h:1-70: n_masks clips drawn in sequence from one C-library rand()
 * stream seeded with srand(seed) (mask.cpp never calls srand: seed 1), column
 * bands of width cd (56 in the reference) that alternate between vacuum and a
 * line/space pattern along y whose segments are multiples of gap (80).  The
 * draws run on the host, the raster on the device (d_bits: n_masks *
 * ndivx*ndivy/8 bytes).  Needs ndivx > cd and ndivy > gap. */
int euvl_mask_linespace(euvl_handle h, unsigned seed, int ndivx, int ndivy, int cd, int gap, int n_masks,
                        uint8_t* d_bits);

/* Host-only (no device work): the band layout euvl_mask_linespace
 * rasterises, in the form euvl_mask_raster takes with one layer per clip:
 * h_sel (n_masks x max(ndivx, ndivy) ints), h_pat_off (n_masks), the pattern
 * pool h_pool (pool_cap bytes; *h_pool_bytes receives the size needed, and
 * EUVL_ERR_ARG is returned if pool_cap is smaller).  No reference
 * counterpart (the reference draws and rasterises in one loop). */
int euvl_mask_linespace_plan(unsigned seed, int ndivx, int ndivy, int cd, int gap, int n_masks, int* h_sel,
                             int64_t* h_pat_off, uint8_t* h_pool, int64_t pool_cap, int64_t* h_pool_bytes);

/* Rasterises band layouts: pixel (i, j) of clip m is the AND over its
 * n_layers (<= EUVL_MASK_MAX_LAYERS) layers, then inverted if invert != 0.
 * Layer L = m*n_layers + l cuts the x axis (h_axis[L] = 0: column i selects a
 * pattern over y, of ndivy bytes) or the y axis (1: row j selects a pattern
 * over x, of ndivx bytes): h_sel[L*max(ndivx, ndivy) + position] is a pattern
 * index counted from byte h_pat_off[L] of the pool h_pool (pool_bytes bytes,
 * one 0/1 byte per pixel), or -1 (constant 1) or -2 (constant 0).  Every
 * index is checked against the pool on the host.  Drives the Python
 * generators LinePattern and RandomLineSpacePattern
 * (elitho/mask_pattern.py:18-196), whose draws run on the host. */
#define EUVL_MASK_MAX_LAYERS 2
int euvl_mask_raster(euvl_handle h, int ndivx, int ndivy, int n_masks, int n_layers, const int* h_axis,
                     const int* h_sel, const int64_t* h_pat_off, const uint8_t* h_pool, int64_t pool_bytes,
                     int invert, uint8_t* d_bits);

/* Device time in milliseconds of the last raster kernel launched by
 * euvl_mask_raster or euvl_mask_linespace on this handle (HIP events on the
 * handle's stream; waits for it).  No reference counterpart
 * (instrumentation). */
int euvl_mask_raster_ms(euvl_handle h, double* ms);

/* ---- CNN M3D predictor (cnn/cnn/predict/predict.py) ----
 * FP32 tensors, NCHW, caller-owned device memory. */

/* compmask (cnn/cnn/predict/compmask.py:17-41): n_clips mask.bin clips of
 * ndiv x ndiv bits -> d_out[c][r][x] (nout x nout floats), the mean of the
 * (ndiv/nout)^2 block of bits at block row r (mask.bin rows, top first) and
 * block column x.  ndiv must be a multiple of nout. */
int euvl_cnn_compmask(euvl_handle h, const uint8_t* d_bits, int ndiv, int n_clips, int nout, float* d_out);

/* One conv block of CNNmodel.forward (predict.py:34-45): x = norm(pool(
 * relu(conv(x)))) with conv = Conv2d(cin, cout, 3, padding=1,
 * padding_mode='circular') (d_w: [cout][cin][3][3], d_bias: [cout]), pool =
 * MaxPool2d(2, 2) and norm = BatchNorm2d in eval mode folded into a
 * per-channel d_scale = weight / sqrt(running_var + eps) and d_shift = bias -
 * running_mean * d_scale.  d_x: [batch][cin][height][width] -> d_y:
 * [batch][cout][height/2][width/2]; height and width even. */
int euvl_cnn_conv_block(euvl_handle h, int batch, int cin, int cout, int height, int width, const float* d_x,
                        const float* d_w, const float* d_bias, const float* d_scale, const float* d_shift,
                        float* d_y);

/* nn.Linear (predict.py:32-33,49-50): d_y[batch][out] = act(d_x[batch][in]
 * d_w[out][in]^T + d_bias) / d_div, act = ReLU if relu != 0, d_div (per
 * output feature) optional.  A split-K FP32 MFMA GEMM whose K slices are
 * summed in a fixed order (partials in the handle's workspace). */
int euvl_cnn_linear(euvl_handle h, int batch, int in_features, int out_features, const float* d_x,
                    const float* d_w, const float* d_bias, int relu, const float* d_div, float* d_y);

/* ---- the TCCEE image (cnn/intensity/stccsocs/elitho/stcc.py) ---- */
typedef struct {
  double k, kx0, ky0, z0, dx, dy, na; /* elitho const: k, kx0, ky0, z0, dx, dy, NA */
  int mx, my;                         /* MX, MY */
  int lpmax_x, lpmax_y;               /* pupil window half widths (lpmaxX, lpmaxY) */
  int xdiv;                           /* image grid XDIV (both axes, as stcc.py:300) */
} euvl_stcc_params;

/* intensity() of elitho/stcc.py:394-431: the image ifft2(TCCEE,
 * norm="forward").real on xdiv x xdiv (d_img, index x*xdiv+y), from the
 * forward-normalised FFT of the 0/1 mask pattern on the pupil window
 * (d_fmask, (2 lpmax_x+1) x (2 lpmax_y+1) complex, index [l+lpmax_x][m+lpmax_y]),
 * the zero-order amplitudes (ab = amp_absorber, vc = amp_vacuum, phase0 =
 * phasexx; host (re, im)), optional M3D windows d_a0, d_ax, d_ay (same
 * layout as d_fmask; all three or none; required for polar 'Y', whose branch
 * always reads them, stcc.py:239), the pupil points h_pupil (ncut (l, m)
 * pairs, pupil.py:24-44) and the uniform k-source h_src (n_src (sx, sy)
 * pairs, source.py:157-199).  The TCCs of the polarisation are three batched
 * ZGEMMs over the source points; TCCEE is gathered over the difference
 * window in a fixed order. */
int euvl_stcc_image(euvl_handle h, char polar, const euvl_stcc_params* p, const double* d_fmask, const double* ab,
                    const double* vc, const double* phase0, const double* d_a0, const double* d_ax,
                    const double* d_ay, int ncut, const int* h_pupil, int n_src, const double* h_src,
                    double* d_img);

/* ---------------------------------------------------------------------------
 * Batched plans and graphs.  The reference runs its drivers once per mask
 * (the ndata loop of m3d.cpp:197-558; intensity.cpp:219-408 per mask file).
 * A plan runs the same pipeline -- mask spectrum, EM solve (ampS,
 * header.h:1208-1466), M3D fit (m3d.cpp:329-467) and the rigorous Abbe image
 * (intensity.cpp:276-408) -- for n_clips masks of one problem in one pass of
 * batched kernels (batch = clips x source divisions).  No reference
 * counterpart as an interface: it batches the calls above.
 * ------------------------------------------------------------------------- */
typedef struct euvl_plan_s* euvl_plan;
typedef struct euvl_graph_s* euvl_graph;

typedef struct {
  char polar;                  /* 'X' or 'Y' */
  const euvl_em_params* p;     /* optics, stack, grid (copied) */
  const int* h_lindex;         /* EM order set (nrange), |l| <= lmax, |m| <= mmax */
  const int* h_mindex;
  int nrange, lmax, mmax;
  int n_off;                   /* source divisions */
  const double* h_off;         /* (sx0, sy0, fx, fy) per division; the first is the central one */
  int n_clips;                 /* masks per run */
  double kx0, ky0, z0;
  /* M3D fit (skipped when ncut = 0) */
  int ncut;
  const int* h_pupil3;         /* (linput, minput, xinput) per pupil point */
  int lpmax_x, lpmax_y;        /* thin-mask window */
  const double* h_vc;          /* blank amplitudes, nsx x nsy complex */
  const double* h_ab;
  /* Abbe image (skipped when n_src = 0; needs the pupil points above) */
  int n_src;
  const int* h_src;            /* (division, l0, m0) per source point */
  double z;                    /* defocus */
  int nx, ny;
  double weight;               /* 1 / n_src for the mean over source points */
  double alpha_margin;         /* >= 1: factor on the calibrated norm bound behind the doubling count */
} euvl_plan_desc;

/* Makes a plan: all tables go to device memory owned by the plan. */
int euvl_plan_create(euvl_handle h, const euvl_plan_desc* d, euvl_plan* out);
int euvl_plan_destroy(euvl_plan pl);
/* Workspace bytes a run needs. */
int64_t euvl_plan_workspace_bytes(euvl_plan pl);
/* Reserves the handle's workspace and LU scratch for the largest of the
 * plans, so that later runs and their graphs never reallocate. */
int euvl_plan_reserve(euvl_handle h, int n_plans, const euvl_plan* plans);
/* One run over n_clips packed masks (d_bits, clip after clip).  Outputs:
 * d_Ax [clip][division][nsx][nsy][nrange] complex; d_a0, d_ax, d_ay
 * [clip][ncut] complex; d_img [clip][nx][ny] float64 (overwritten).
 * fixed = 0: adaptive (host-read norms pick the doubling count, the top
 * solve falls back to partial pivoting), and the plan
 * keeps the largest decisions of all its adaptive runs (the doublings for
 * alpha_margin times the largest norm bound seen).  fixed = 1: the plan's decisions are used as they are, checked
 * on the device; d_flags [clip x division] receives 1 (doubling bound),
 * 4 (top-solve residual), 8 (zero pivot) where a check
 * failed -- those clips must be run again adaptively.  A fixed run only
 * enqueues work (no synchronisation, allocation or host copy) and may be
 * captured; the caller synchronises before reading anything. */
int euvl_plan_run(euvl_plan pl, const uint8_t* d_bits, double* d_Ax, double* d_a0, double* d_ax, double* d_ay,
                  double* d_img, int* d_flags, int fixed);
int euvl_plan_decisions(euvl_plan pl, int* doublings);
int euvl_plan_set_decisions(euvl_plan pl, int doublings);

/* HIP graph of the work enqueued on the handle's stream between begin and
 * end (fixed-mode plan runs; the stream must not be the null stream).  A
 * graph keeps the workspace addresses of the capture: reserve first
 * (euvl_plan_reserve) so that the handle's workspace does not grow while the
 * graph is in use.  euvl_graph_launch replays it on the handle's current
 * stream; it returns EUVL_ERR_ARG for another handle than the capture's and
 * EUVL_ERR_STALE, launching nothing, when that handle's workspace or LU
 * scratch has been reallocated or shrunk since the capture. */
int euvl_graph_begin(euvl_handle h);
int euvl_graph_end(euvl_handle h, euvl_graph* out);
int euvl_graph_launch(euvl_graph g, euvl_handle h);
int euvl_graph_nodes(euvl_graph g, int64_t* n);
int euvl_graph_destroy(euvl_graph g);

/* ---- training data (cnn/intensity/m3d/compress.py, the model scripts of cnn/cnn/model) ---- */

/* The CNN data loader's augmentation (MaskampDataset.__getitem__,
 * cnn/cnn/model/re0/re0.py:24-62) for n samples at once: h_shifts holds
 * (lshift, mshift) per sample.  Masks (float16 bit patterns, [n][nout][nout],
 * optional: both d_masks and d_out_masks or neither) are rolled as
 * np.roll(x, (mshift, -lshift), axis=(0, 1)); the amplitudes d_re, d_im
 * ([n][ncut] float32) become d_out_re = re cos - im sin and d_out_im = re sin
 * + im cos at index (l lshift + m mshift) mod nout of the float32 tables
 * h_cos, h_sin (nout entries), with h_lorder, h_morder the pupil orders (the
 * rshift / ishift of re0.py:24-36, float32, unfused). */
int euvl_augment_batch(euvl_handle h, int n, int nout, const uint16_t* d_masks, const int* h_shifts, int ncut,
                       const int* h_lorder, const int* h_morder, const float* d_re, const float* d_im,
                       const float* h_cos, const float* h_sin, uint16_t* d_out_masks, float* d_out_re,
                       float* d_out_im);

#ifdef __cplusplus
}
#endif
#endif /* EUVL_H */
