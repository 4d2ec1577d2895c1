// STCC-SOCS imaging: transmission cross coefficients from the source mesh,
// their Hermitian eigendecomposition (kernels), and the SOCS image.
//   eigenfunction.py:5-179  -> euvl_tcc_kernels
//   socs.py:12-154          -> euvl_socs_image
// TCC_w = sum_c P_c diag(w) P_c^H / S is one MFMA GEMM per weight (the three
// polarisation components stacked along K); the eigendecomposition is a
// parallel two-sided Jacobi method (round-robin pairs, n/2 rotations per
// step applied in one out-of-place launch over 2 x 2 blocks), or, when only
// a few leading kernels are asked for, block subspace iteration with
// Rayleigh-Ritz steps by the same Jacobi method on the small projected matrix.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <numeric>
#include <vector>

#include "ew.h"
#include "optics.h"

namespace {

struct TccSetup {
  double k, dx, dy, na, kx0, ky0, z0, z;
  int mx, my;
};

// P[i][c*S + s] = phase R_c (zero outside the pupil), and
// Q[w][c*S + s][i] = conj(P[i][c*S + s]) wt_w[s] / S for w = (1, sx, sy) * k^2/(k^2-(kx0+sx)^2)
// (eigenfunction.py:84-111)
__global__ void tcc_phase_kernel(TccSetup a, const int* pupil, int ncut, const double* src, int S, double2* P,
                                 double2* Q) {
  const int64_t total = (int64_t)ncut * S;
  const int64_t cols = 3 * (int64_t)S;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / S), s = (int)(e % S);
    const double sx = src[2 * s], sy = src[2 * s + 1];
    const double ksx = 2.0 * PI / a.dx * pupil[2 * i] + sx;
    const double ksy = 2.0 * PI / a.dy * pupil[2 * i + 1] + sy;
    const double p2 = a.mx * a.mx * ksx * ksx + a.my * a.my * ksy * ksy;
    double2 v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    if (p2 <= (a.k * a.na) * (a.k * a.na)) {
      double R[3];
      rotate_col(a.k, a.mx, a.my, ksx, ksy, a.kx0, a.ky0, 0, R);
      const double ph = ((ksx + a.kx0) * (ksx + a.kx0) + (ksy + a.ky0) * (ksy + a.ky0)) / 2.0 / a.k * a.z0 +
                        p2 / 2.0 / a.k * a.z;
      double sp, cp;
      sincos(ph, &sp, &cp);
      for (int c = 0; c < 3; ++c) v[c] = make_double2(cp * R[c], sp * R[c]);
    }
    const double w = a.k * a.k / (a.k * a.k - (a.kx0 + sx) * (a.kx0 + sx)) / S;
    const double wt[3] = {w, sx * w, sy * w};
    for (int c = 0; c < 3; ++c) {
      P[(int64_t)i * cols + (int64_t)c * S + s] = v[c];
      for (int q = 0; q < 3; ++q)
        Q[(int64_t)q * cols * ncut + ((int64_t)c * S + s) * ncut + i] = make_double2(v[c].x * wt[q], -v[c].y * wt[q]);
    }
  }
}

// The Jacobi rotation of pair (p, q) of the Hermitian matrix M: block
// [[a, b], [conj b, d]], b = |b| e^{i phi}: J = diag(1, e^{-i phi}) G, G the real
// Jacobi rotation of [[a, |b|], [|b|, d]] (c, s with t = tan theta).
// Returns (c, s, cos phi, sin phi); the identity for a bye (p or q < 0).
__device__ inline double4_t pair_rotation(const double2* M, int n, int p, int q) {
  double4_t r = {1.0, 0.0, 1.0, 0.0};
  if (p < 0 || q < 0) return r;
  const double a = M[(int64_t)p * n + p].x, d = M[(int64_t)q * n + q].x;
  const double2 bq = M[(int64_t)p * n + q];
  const double ab = hypot(bq.x, bq.y);
  if (ab > 1e-300 && ab > 1e-18 * (fabs(a) + fabs(d))) {
    const double tau = (d - a) / (2.0 * ab);
    const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
    const double c = 1.0 / sqrt(1.0 + t * t);
    r = (double4_t){c, t * c, bq.x / ab, bq.y / ab};
  }
  return r;
}

// Pair order.  The matrix (np x np, n padded to even np with a zero row and
// column that only ever pairs with itself as a no-op) is stored with the
// indices permuted so that the pairs of the current round-robin step are the
// positions (2k, 2k+1): the step then reads contiguous 2 x 2 blocks.  The
// tournament's ring (position 0 fixed, the rest rotating by one each step)
// moves the index at pair position p to pair_next(p) in the next step's order;
// every step writes its output in that order, and after np - 1 steps (a
// sweep) the order is the initial one again.
__device__ __host__ inline int pair_next(int p, int np) {
  const int r = (p & 1) ? np - 1 - (p - 1) / 2 : p / 2;  // ring position
  const int r2 = (r == 0) ? 0 : (r == np - 1 ? 1 : r + 1);
  return (r2 < np / 2) ? 2 * r2 : 2 * (np - 1 - r2) + 1;
}
__device__ __host__ inline int pair_index0(int p, int np) { return (p & 1) ? np - 1 - p / 2 : p / 2; }

// A (n x n, natural order) -> Ap (np x np, initial pair order, zero padded);
// Vp = the identity with its columns in pair order (rows in natural order)
__global__ void pair_order_init_kernel(int n, int np, const double2* A, double2* Ap, double2* Vp, int batch) {
  const int64_t per = (int64_t)np * np, total = per * batch;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / per);
    const int64_t r = e - (int64_t)b * per;
    const int i = (int)(r / np), j = (int)(r % np);
    const int oi = pair_index0(i, np), oj = pair_index0(j, np);
    Ap[e] = (oi < n && oj < n) ? A[(int64_t)b * n * n + (int64_t)oi * n + oj] : make_double2(0.0, 0.0);
    Vp[e] = (i == oj) ? make_double2(1.0, 0.0) : make_double2(0.0, 0.0);
  }
}

// One round-robin step in one launch, out of place, in pair order: thread
// (k, l) rotates the 2 x 2 block rows (2k, 2k+1) x columns (2l, 2l+1) from both
// sides (rows by J_k^H, then columns by J_l) and writes it at the positions of
// the next step's order; the pairs partition the positions, so every element
// is written exactly once, and reading only the input copy makes the
// parameters of every pair those of the step's start.  Threads past the
// npair^2 blocks rotate columns (2l, 2l+1) of row i of V.
__global__ void jacobi_pairstep_kernel(int np, const double2* __restrict__ Ain, double2* __restrict__ Aout,
                                       const double2* __restrict__ Vin, double2* __restrict__ Vout, int batch) {
  const int npair = np / 2;
  const int64_t nblk = (int64_t)npair * npair, nv = (int64_t)np * npair, per = nblk + nv;
  const int64_t total = per * batch;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / per);
    const int64_t r = e - (int64_t)b * per;
    const double2* M = Ain + (int64_t)b * np * np;
    if (r < nblk) {
      const int k = (int)(r / npair), l = (int)(r % npair);
      const double4_t Rk = pair_rotation(M, np, 2 * k, 2 * k + 1);
      const double4_t Rl = pair_rotation(M, np, 2 * l, 2 * l + 1);
      const double2* row0 = M + (int64_t)(2 * k) * np + 2 * l;
      double2 X[2][2] = {{row0[0], row0[1]}, {row0[np], row0[np + 1]}};
      if (Rk[1] != 0.0) {
        const double2 ep = make_double2(Rk[2], Rk[3]);  // e^{i phi_k}
        for (int v = 0; v < 2; ++v) {
          const double2 x = X[0][v], ey = cmul(ep, X[1][v]);
          X[0][v] = csub(cscale(x, Rk[0]), cscale(ey, Rk[1]));
          X[1][v] = cadd(cscale(x, Rk[1]), cscale(ey, Rk[0]));
        }
      }
      if (Rl[1] != 0.0) {
        const double2 em = make_double2(Rl[2], -Rl[3]);  // e^{-i phi_l}
        for (int u = 0; u < 2; ++u) {
          const double2 x = X[u][0], ey = cmul(em, X[u][1]);
          X[u][0] = csub(cscale(x, Rl[0]), cscale(ey, Rl[1]));
          X[u][1] = cadd(cscale(x, Rl[1]), cscale(ey, Rl[0]));
        }
      }
      double2* O = Aout + (int64_t)b * np * np;
      const int r0 = pair_next(2 * k, np), r1 = pair_next(2 * k + 1, np);
      const int c0 = pair_next(2 * l, np), c1 = pair_next(2 * l + 1, np);
      O[(int64_t)r0 * np + c0] = X[0][0];
      O[(int64_t)r0 * np + c1] = X[0][1];
      O[(int64_t)r1 * np + c0] = X[1][0];
      O[(int64_t)r1 * np + c1] = X[1][1];
    } else {
      const int64_t rv = r - nblk;
      const int i = (int)(rv / npair), l = (int)(rv % npair);
      const double2* Vi = Vin + (int64_t)b * np * np + (int64_t)i * np;
      double2* Vo = Vout + (int64_t)b * np * np + (int64_t)i * np;
      const double4_t R = pair_rotation(M, np, 2 * l, 2 * l + 1);
      double2 x = Vi[2 * l], y = Vi[2 * l + 1];
      if (R[1] != 0.0) {
        const double2 ey = cmul(make_double2(R[2], -R[3]), y);
        const double2 nx = csub(cscale(x, R[0]), cscale(ey, R[1]));
        y = cadd(cscale(x, R[1]), cscale(ey, R[0]));
        x = nx;
      }
      Vo[pair_next(2 * l, np)] = x;
      Vo[pair_next(2 * l + 1, np)] = y;
    }
  }
}

// off-diagonal and total squared Frobenius norms per batch entry (atomics)
__global__ void offdiag_kernel(int n, const double2* A, double* out, int batch) {
  const int b = blockIdx.y;
  const double2* M = A + (int64_t)b * n * n;
  double off = 0.0, tot = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < (int64_t)n * n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const double2 v = M[e];
    const double a = v.x * v.x + v.y * v.y;
    tot += a;
    if (e / n != e % n) off += a;
  }
  for (int o = 32; o > 0; o >>= 1) {
    off += __shfl_xor(off, o);
    tot += __shfl_xor(tot, o);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(out + 2 * b, off);
    atomicAdd(out + 2 * b + 1, tot);
  }
}


__global__ void zero_kernel(double* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0.0;
}

__global__ void diag_kernel(int n, const double2* A, double* w, int batch) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * batch) return;
  const int b = (int)(e / n), i = (int)(e % n);
  w[e] = A[(int64_t)b * n * n + (int64_t)i * n + i].x;
}

// phi[b][m][j] = V[b][j][order[b][m]] (V: np x np, rows the n components)
__global__ void gather_kernel(int n, int np, int nk, const double2* V, const int* order, double2* phi, int batch) {
  const int64_t total = (int64_t)nk * n * batch;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / ((int64_t)nk * n));
    const int64_t r = e - (int64_t)b * nk * n;
    const int m = (int)(r / n), j = (int)(r % n);
    phi[e] = V[(int64_t)b * np * np + (int64_t)j * np + order[b * nk + m]];
  }
}

// ---------------------------------------------------------------- SOCS image
struct SocsSetup {
  double2 ampvc, ampab, phase0;
  double k, dx, dy, kx0, ky0, z0;
  int lpmx, lpmy;  // thin spectrum window
  int lpx, lpy;    // pupil half widths
};

// the four amplitude vectors of socs.py:83-98 (thin amplitude from the indicator spectrum)
__global__ void socs_amps_kernel(SocsSetup a, const double2* fmask, const double2* a0, const double2* ax,
                                 const double2* ay, const int* pupil, int ncut, double2* amps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncut) return;
  const int lp = pupil[2 * i], mp = pupil[2 * i + 1];
  double2 fm = cmul(csub(a.ampvc, a.ampab), fmask[(int64_t)(lp + a.lpmx) * (2 * a.lpmy + 1) + (mp + a.lpmy)]);
  if (lp == 0 && mp == 0) fm = cadd(fm, a.ampab);
  const double kxp = 2.0 * PI * lp / a.dx, kyp = 2.0 * PI * mp / a.dy;
  const double ph = -(a.kx0 * kxp + kxp * kxp / 2.0) / a.k * a.z0 - (a.ky0 * kyp + kyp * kyp / 2.0) / a.k * a.z0;
  double sp, cp;
  sincos(ph, &sp, &cp);
  const double2 famp = cdiv(cmul(fm, make_double2(cp, sp)), a.phase0);
  amps[i] = cadd(famp, a0[i]);                                                // Axs0
  amps[ncut + i] = cscale(ax[i], a.dx / 2.0 / PI);                            // Axsx
  amps[2 * ncut + i] = cscale(ay[i], a.dy / 2.0 / PI);                        // Axsy
  amps[3 * ncut + i] = cadd(cscale(ax[i], lp / 2.0), cscale(ay[i], mp / 2.0));  // Axsxy
}

// field spectra: field f = (amplitude vector va[f], kernel set vk[f], kernel m[f])
__global__ void socs_spectra_kernel(const SocsSetup a, int nf, const int* fdesc, const double2* amps,
                                    const double2* phi, int nk, const int* pupil, int ncut, double2* F) {
  const int64_t total = (int64_t)nf * ncut;
  const int nl = 2 * a.lpx + 1, nm = 2 * a.lpy + 1;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int f = (int)(e / ncut), i = (int)(e % ncut);
    const int va = fdesc[3 * f], vk = fdesc[3 * f + 1], m = fdesc[3 * f + 2];
    const double2 v = cmul(amps[(int64_t)va * ncut + i], phi[((int64_t)vk * nk + m) * ncut + i]);
    F[(int64_t)f * nl * nm + (int64_t)(pupil[2 * i] + a.lpx) * nm + (pupil[2 * i + 1] + a.lpy)] = v;
  }
}

// G[f][l][y] = sum_m F[f][l][m] e^{2 pi i m y / nout}
__global__ void socs_ypass_kernel(int nf, int nl, int nm, int lpy, int nout, const double2* F, const double2* tw,
                                  double2* G) {
  const int64_t total = (int64_t)nf * nl * nout;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int y = (int)(e % nout);
    const int64_t fl = e / nout;
    const double2* f = F + fl * nm;
    double2 acc = make_double2(0.0, 0.0);
    for (int mm = 0; mm < nm; ++mm) {
      const int m = mm - lpy;
      acc = cadd(acc, cmul(f[mm], tw[(((int64_t)m * y) % nout + nout) % nout]));
    }
    G[e] = acc;
  }
}

// fields[f][x][y] = sum_l e^{2 pi i l x / nout} G[f][l][y]
__global__ void socs_xpass_kernel(int nf, int nl, int lpx, int nout, const double2* G, const double2* tw,
                                  double2* fld) {
  const int64_t total = (int64_t)nf * nout * nout;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int f = (int)(e / ((int64_t)nout * nout));
    const int64_t r = e - (int64_t)f * nout * nout;
    const int x = (int)(r / nout), y = (int)(r % nout);
    double2 acc = make_double2(0.0, 0.0);
    for (int l = 0; l < nl; ++l)
      acc = cadd(acc, cmul(tw[(((int64_t)(l - lpx) * x) % nout + nout) % nout], G[((int64_t)f * nl + l) * nout + y]));
    fld[e] = acc;
  }
}

__device__ inline double re_dot(double2 a, double2 b) { return a.x * b.x + a.y * b.y; }

// small image (socs.py:100-142): fields ordered [f0_0..f0_{n0-1}] then, per m < nxy,
// (f0x, f0y, fx, fy, fxy)
__global__ void socs_combine_kernel(int nout, int n0, int nxy, const double2* fld, const double* alpha, int nk,
                                    double* small) {
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= nout * nout) return;
  const int64_t np = (int64_t)nout * nout;
  double acc = 0.0;
  for (int m = 0; m < n0; ++m) {
    const double2 f0 = fld[m * np + pix];
    acc += alpha[m] * (f0.x * f0.x + f0.y * f0.y);
    if (m < nxy) {
      const double2* g = fld + (n0 + 5 * (int64_t)m) * np + pix;
      const double2 f0x = g[0], f0y = g[np], fx = g[2 * np], fy = g[3 * np], fxy = g[4 * np];
      acc += 2.0 * alpha[nk + m] * re_dot(f0x, fx) + 2.0 * alpha[2 * nk + m] * re_dot(f0y, fy) +
             2.0 * alpha[m] * re_dot(f0, fxy);
    }
  }
  small[pix] = acc;
}

// S[u][v] = (1/nout^2) sum_{x,y} small[x][y] e^{-2 pi i (u x + v y)/nout}, u, v in [-nout/2, nout/2)
// Centred spectrum of the small image, S[u][v] = (1/nout^2) sum_{x,y}
// small[x][y] e^{-2 pi i (u x + v y)/nout}, u, v in [-nout/2, nout/2) (the
// fft2 / nout^2 of socs.py:143), in two separable passes:
// R[x][v] = sum_y small[x][y] e^{-2 pi i v y/nout}
__global__ void socs_spec_y_kernel(int nout, const double* small, const double2* twm, double2* R) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nout * nout) return;
  const int x = e / nout, v = e % nout - nout / 2;
  double2 acc = make_double2(0.0, 0.0);
  for (int y = 0; y < nout; ++y) acc = cadd(acc, cscale(twm[(((int64_t)v * y) % nout + nout) % nout], small[x * nout + y]));
  R[e] = acc;
}

// S[u][v] = (1/nout^2) sum_x e^{-2 pi i u x/nout} R[x][v]
__global__ void socs_spec_x_kernel(int nout, const double2* R, const double2* twm, double2* S) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nout * nout) return;
  const int ui = e / nout, vi = e % nout;
  const int u = ui - nout / 2;
  double2 acc = make_double2(0.0, 0.0);
  for (int x = 0; x < nout; ++x) acc = cadd(acc, cmul(twm[(((int64_t)u * x) % nout + nout) % nout], R[x * nout + vi]));
  S[e] = cscale(acc, 1.0 / ((double)nout * nout));
}

// T[u][y] = sum_v S[u][v] e^{2 pi i v y / ny}
__global__ void socs_up_y_kernel(int nout, int ny, const double2* S, const double2* twy, double2* T) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nout * ny) return;
  const int ui = (int)(e / ny), y = (int)(e % ny);
  double2 acc = make_double2(0.0, 0.0);
  for (int vi = 0; vi < nout; ++vi) {
    const int v = vi - nout / 2;
    acc = cadd(acc, cmul(S[ui * nout + vi], twy[(((int64_t)v * y) % ny + ny) % ny]));
  }
  T[e] = acc;
}

// img[x][y] = Re sum_u T[u][y] e^{2 pi i u x / nx} * (nx ny) / (nx ny)
__global__ void socs_up_x_kernel(int nout, int nx, int ny, const double2* T, const double2* twx, double* img) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nx * ny) return;
  const int x = (int)(e / ny), y = (int)(e % ny);
  double acc = 0.0;
  for (int ui = 0; ui < nout; ++ui) {
    const int u = ui - nout / 2;
    const double2 t = twx[(((int64_t)u * x) % nx + nx) % nx];
    const double2 v = T[(int64_t)ui * ny + y];
    acc += t.x * v.x - t.y * v.y;
  }
  img[e] = acc;
}

// ---- FP32 variants (EUVL_PREC_FP32_IMAGING): field transforms in FP32, the
// kernel sum of the small image with Kahan compensation in FP32 ----
__global__ void socs_tw32_kernel(int n, const double2* tw, float2* out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) out[e] = make_float2((float)tw[e].x, (float)tw[e].y);
}

__global__ void socs_ypass32_kernel(int nf, int nl, int nm, int lpy, int nout, const double2* F, const float2* tw,
                                    float2* G) {
  const int64_t total = (int64_t)nf * nl * nout;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int y = (int)(e % nout);
    const double2* f = F + (e / nout) * nm;
    float re = 0.0f, im = 0.0f;
    for (int mm = 0; mm < nm; ++mm) {
      const float2 t = tw[((((int64_t)(mm - lpy)) * y) % nout + nout) % nout];
      const float fr = (float)f[mm].x, fi = (float)f[mm].y;
      re = fmaf(fr, t.x, fmaf(-fi, t.y, re));
      im = fmaf(fr, t.y, fmaf(fi, t.x, im));
    }
    G[e] = make_float2(re, im);
  }
}

__global__ void socs_xpass32_kernel(int nf, int nl, int lpx, int nout, const float2* G, const float2* tw,
                                    float2* fld) {
  const int64_t total = (int64_t)nf * nout * nout;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int f = (int)(e / ((int64_t)nout * nout));
    const int64_t r = e - (int64_t)f * nout * nout;
    const int x = (int)(r / nout), y = (int)(r % nout);
    float re = 0.0f, im = 0.0f;
    for (int l = 0; l < nl; ++l) {
      const float2 t = tw[(((int64_t)(l - lpx) * x) % nout + nout) % nout];
      const float2 g = G[((int64_t)f * nl + l) * nout + y];
      re = fmaf(t.x, g.x, fmaf(-t.y, g.y, re));
      im = fmaf(t.x, g.y, fmaf(t.y, g.x, im));
    }
    fld[e] = make_float2(re, im);
  }
}

__device__ inline void kahan_add(float& s, float& c, float v) {
  const float y = v - c, t = s + y;
  c = (t - s) - y;
  s = t;
}

__device__ inline float re_dot32(float2 a, float2 b) { return a.x * b.x + a.y * b.y; }

__global__ void socs_combine32_kernel(int nout, int n0, int nxy, const float2* fld, const double* alpha, int nk,
                                      double* small) {
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= nout * nout) return;
  const int64_t np = (int64_t)nout * nout;
  float acc = 0.0f, cmp = 0.0f;
  for (int m = 0; m < n0; ++m) {
    const float2 f0 = fld[m * np + pix];
    kahan_add(acc, cmp, (float)alpha[m] * re_dot32(f0, f0));
    if (m < nxy) {
      const float2* g = fld + (n0 + 5 * (int64_t)m) * np + pix;
      const float2 f0x = g[0], f0y = g[np], fx = g[2 * np], fy = g[3 * np], fxy = g[4 * np];
      kahan_add(acc, cmp, 2.0f * (float)alpha[nk + m] * re_dot32(f0x, fx));
      kahan_add(acc, cmp, 2.0f * (float)alpha[2 * nk + m] * re_dot32(f0y, fy));
      kahan_add(acc, cmp, 2.0f * (float)alpha[m] * re_dot32(f0, fxy));
    }
  }
  small[pix] = (double)acc - (double)cmp;
}

// ---------------------------------------------------------------- leading eigenpairs
// When only the leading nk << n kernels are asked for (the SOCS image keeps a
// few tens of a TCC of several hundred pupil points), block subspace iteration
// finds them with a few small products per step instead of full Jacobi sweeps
// over the n x n matrix: Q (n x SUB_NB, orthonormal) <- orth(T Q), and a
// Rayleigh-Ritz step (the Jacobi method on Q^H T Q, SUB_NB x SUB_NB) gives the
// Ritz pairs, accepted when every leading residual |T f - theta f| is at
// rounding level.  The error of the leading nk Ritz vectors falls by
// |lambda_(SUB_NB+1)| / |lambda_nk| per step.
constexpr int SUB_NB = 64;      // block width (one gram_orth_kernel workgroup holds SUB_NB x SUB_NB)
constexpr int SUB_MAX_NK = 24;  // at most this many leading pairs, and n >= 2 SUB_NB

__device__ inline uint64_t mix64(uint64_t x) {  // splitmix64 finaliser
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

// a fixed start block: entries uniform in (-1/2, 1/2) + i (-1/2, 1/2), from a
// hash of their position (the same block on every call)
__global__ void start_block_kernel(int64_t total, double2* Y) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const double u = (double)(mix64(2 * (uint64_t)e) >> 11) * 0x1.0p-53;
    const double v = (double)(mix64(2 * (uint64_t)e + 1) >> 11) * 0x1.0p-53;
    Y[e] = make_double2(u - 0.5, v - 0.5);
  }
}

// XH[b][j][i] = conj(X[b][i][j]), X rows x cols
__global__ void conj_transpose_kernel(int rows, int cols, const double2* X, double2* XH, int batch) {
  const int64_t per = (int64_t)rows * cols, total = per * batch;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / per);
    const int64_t r = e - (int64_t)b * per;
    const int j = (int)(r / rows), i = (int)(r % rows);  // output index (j, i): writes in order
    const double2 v = X[(int64_t)b * per + (int64_t)i * cols + j];
    XH[e] = make_double2(v.x, -v.y);
  }
}

// out[b][m][j] = F[b][j][m] (F n x nk)
__global__ void transpose_kernel(int n, int nk, const double2* F, double2* out, int batch) {
  const int64_t per = (int64_t)n * nk, total = per * batch;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / per);
    const int64_t r = e - (int64_t)b * per;
    const int m = (int)(r / n), j = (int)(r % n);
    out[e] = F[(int64_t)b * per + (int64_t)j * nk + m];
  }
}

// H <- (H + H^H) / 2 in place (the Jacobi method reads one triangle's pivots)
__global__ void hermitize_kernel(int n, double2* H, int batch) {
  const int64_t per = (int64_t)n * n, total = per * batch;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / per);
    const int64_t r = e - (int64_t)b * per;
    const int i = (int)(r / n), j = (int)(r % n);
    double2* M = H + (int64_t)b * per;
    if (i < j) {
      const double2 a = M[(int64_t)i * n + j], c = M[(int64_t)j * n + i];
      const double2 m = make_double2(0.5 * (a.x + c.x), 0.5 * (a.y - c.y));
      M[(int64_t)i * n + j] = m;
      M[(int64_t)j * n + i] = make_double2(m.x, -m.y);
    } else if (i == j) {
      M[(int64_t)i * n + i].y = 0.0;
    }
  }
}

// U[b][j][m] = V[b][j][pos[b][m]] (V np x np with its columns in pair order)
__global__ void ritz_select_kernel(int nb, int nk, const double2* V, const int* pos, double2* U, int batch) {
  const int64_t per = (int64_t)nb * nk, total = per * batch;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / per);
    const int64_t r = e - (int64_t)b * per;
    const int j = (int)(r / nk), m = (int)(r % nk);
    U[e] = V[(int64_t)b * nb * nb + (int64_t)j * nb + pos[b * nk + m]];
  }
}

// res[b][m] = sum_j |W[b][j][m] - theta[b][m] F[b][j][m]|^2, one workgroup per (m, b)
__global__ __launch_bounds__(256) void ritz_residual_kernel(int n, int nk, const double2* W, const double2* F,
                                                            const double* theta, double* res) {
  const int m = blockIdx.x, b = blockIdx.y;
  const double th = theta[b * nk + m];
  const double2* Wb = W + (int64_t)b * n * nk;
  const double2* Fb = F + (int64_t)b * n * nk;
  double acc = 0.0;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    const double2 w = Wb[(int64_t)j * nk + m], f = Fb[(int64_t)j * nk + m];
    const double rx = w.x - th * f.x, ry = w.y - th * f.y;
    acc += rx * rx + ry * ry;
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  __shared__ double part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) res[b * nk + m] = part[0] + part[1] + part[2] + part[3];
}

// Orthonormalising factor of a block: for the Gram matrix G = Y^H Y
// (SUB_NB x SUB_NB, Hermitian positive definite) the unpivoted elimination of
// [G | I] leaves L^-1 on the right (G = L D L^H, L unit lower), and
// M = L^-H D^-1/2 gives (Y M)^H (Y M) = M^H G M = I.  One workgroup per batch
// entry; thread (r0, c) keeps rows r0 + q RS of both halves in registers; the
// pivot row and column cross threads through LDS buffers of parity p & 1 (one
// barrier per step).  A pivot below 1e-12 of the largest diagonal entry (a
// block that lost rank) sets fail[b].
template <int NT>
__global__ __launch_bounds__(NT) void gram_orth_kernel(const double2* __restrict__ G, double2* __restrict__ M,
                                                       int* __restrict__ fail) {
  constexpr int NB = SUB_NB;
  static_assert(NT % NB == 0 && NB == 64, "one row of the block per wave");
  constexpr int RS = NT / NB, RPT = NB / RS;
  __shared__ double2 rowg[2][NB], rowi[2][NB], colg[2][NB];
  __shared__ double dsh[NB];
  __shared__ double scale;
  const int b = blockIdx.x, t = threadIdx.x;
  const int c = t & (NB - 1), r0 = t / NB;
  const double2* Gb = G + (int64_t)b * NB * NB;
  double2 x[RPT], y[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int i = r0 + RS * q;
    x[q] = Gb[(int64_t)i * NB + c];
    y[q] = make_double2(i == c ? 1.0 : 0.0, 0.0);
    if (i == c) dsh[i] = x[q].x;
  }
  __syncthreads();
  if (t == 0) {
    double m = 0.0;
    for (int i = 0; i < NB; ++i) m = fmax(m, dsh[i]);
    scale = m;
  }
  __syncthreads();
  const double tiny = 1e-12 * scale;
  bool bad = false;
#pragma unroll 1
  for (int p = 0; p < NB; ++p) {
    const int pb = p & 1;
    if ((p % RS) == r0) {  // this wave holds row p
      const int qp = p / RS;
      double2 v = x[0], w = y[0];
#pragma unroll
      for (int q = 1; q < RPT; ++q)
        if (q == qp) {
          v = x[q];
          w = y[q];
        }
      rowg[pb][c] = v;
      rowi[pb][c] = w;
      if (c == p) dsh[p] = v.x;
    }
    if (c == p) {
#pragma unroll
      for (int q = 0; q < RPT; ++q) colg[pb][r0 + RS * q] = x[q];
    }
    __syncthreads();
    const double d = dsh[p];
    bad = bad || !(d > tiny);
    const double inv = (d > tiny) ? 1.0 / d : 0.0;
    const double2 rg = rowg[pb][c], ri = rowi[pb][c];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int i = r0 + RS * q;
      const double2 g = colg[pb][i];
      const double lx = (i > p) ? g.x * inv : 0.0, ly = (i > p) ? g.y * inv : 0.0;
      x[q].x -= lx * rg.x - ly * rg.y;
      x[q].y -= lx * rg.y + ly * rg.x;
      y[q].x -= lx * ri.x - ly * ri.y;
      y[q].y -= lx * ri.y + ly * ri.x;
    }
  }
  // M[c'][i] = conj(L^-1[i][c']) / sqrt(d_i)
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int i = r0 + RS * q;
    const double di = dsh[i];
    const double sc = (di > tiny) ? 1.0 / sqrt(di) : 0.0;
    M[(int64_t)b * NB * NB + (int64_t)c * NB + i] = make_double2(y[q].x * sc, -y[q].y * sc);
  }
  if (t == 0 && bad) fail[b] = 1;
}

__global__ void zero_ints_kernel(int* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

std::vector<std::complex<double>> twiddles(int n, double sign) {
  std::vector<std::complex<double>> t(n);
  const std::complex<double> zi(0.0, 1.0);
  for (int i = 0; i < n; ++i) t[i] = std::exp(sign * 2.0 * PI * zi * double(i) / double(n));
  return t;
}

}  // namespace

namespace {

// Jacobi sweeps on A (np x np in pair order, V accumulating the rotations)
// until the squared off-diagonal / total Frobenius ratio is at the rounding
// floor of an n x n matrix, (2 n eps)^2, or stops falling below 1e3 times it;
// Jacobi converges quadratically, so this takes few sweeps.  A and V (and the
// spare buffers A2, V2) are swapped as the out-of-place steps go.
int jacobi_sweeps(euvl_handle h, int n, int np, double2*& A, double2*& V, double2*& A2, double2*& V2, int B,
                  int* sweeps_out) {
  using namespace euvl;
  const int npair = np / 2, nsteps = np - 1;
  const double eps = 2.220446049250313e-16;
  const double floor2 = (2.0 * n * eps) * (2.0 * n * eps);
  const int64_t step_threads = ((int64_t)npair * npair + (int64_t)np * npair) * B;
  int sweep = 0;
  double worst = INFINITY, prev = INFINITY;
  for (sweep = 0; sweep <= 40; ++sweep) {
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, h->stream, h->d_red, 2 * B);
    EUVL_LAUNCH_CHECK();
    hipLaunchKernelGGL(offdiag_kernel, dim3(256, B), dim3(256), 0, h->stream, np, A, h->d_red, B);
    EUVL_LAUNCH_CHECK();
    EUVL_HIP_CHECK(hipMemcpyAsync(h->h_pinned, h->d_red, sizeof(double) * 2 * B, hipMemcpyDeviceToHost, h->stream));
    EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));
    worst = 0.0;
    for (int b = 0; b < B; ++b) {
      const double off = h->h_pinned[2 * b], tot = h->h_pinned[2 * b + 1];
      const double r = (tot > 0.0) ? off / tot : 0.0;
      worst = std::max(worst, std::isnan(r) ? INFINITY : r);
    }
    if (worst <= floor2 || (worst <= 1e3 * floor2 && worst > 0.25 * prev) || sweep == 40) break;
    prev = worst;
    for (int st = 0; st < nsteps; ++st) {
      hipLaunchKernelGGL(jacobi_pairstep_kernel, grid1(step_threads), dim3(256), 0, h->stream, np, A, A2, V, V2, B);
      std::swap(A, A2);
      std::swap(V, V2);
    }
    EUVL_LAUNCH_CHECK();
  }
  if (sweeps_out) *sweeps_out = sweep;
  if (!(worst <= 1e3 * floor2)) {
    set_error("euvl_tcc_kernels: Jacobi off-diagonal ratio %.3e after %d sweeps (floor %.3e)", worst, sweep, floor2);
    return EUVL_ERR_NOT_CONVERGED;
  }
  return EUVL_OK;
}

// The eigenvalues of a converged Jacobi run (the diagonal, one per position of
// the pair order; a sweep ends in the initial order) read back to the host,
// and for each batch entry the positions of the n eigenvalues in output order:
// |alpha| descending, ties in ascending alpha (as after the reference's
// ascending eigh and its stable bubble sort), then ascending index.  The
// padding position (odd n) is left out.
int jacobi_order(euvl_handle h, int n, int np, const double2* A, double* w, int B, std::vector<double>& hw,
                 std::vector<std::vector<int>>& pos) {
  using namespace euvl;
  hipLaunchKernelGGL(diag_kernel, grid1((int64_t)np * B), dim3(256), 0, h->stream, np, A, w, B);
  EUVL_LAUNCH_CHECK();
  hw.assign((size_t)np * B, 0.0);
  EUVL_HIP_CHECK(hipMemcpyAsync(hw.data(), w, sizeof(double) * np * B, hipMemcpyDeviceToHost, h->stream));
  EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));
  pos.assign(B, {});
  for (int b = 0; b < B; ++b) {
    std::vector<int>& idx = pos[b];
    // positions sorted by the index they hold, so that ties keep ascending index order
    for (int q = 0; q < n; ++q) {
      for (int ps = 0; ps < np; ++ps)
        if (pair_index0(ps, np) == q) {
          idx.push_back(ps);
          break;
        }
    }
    const double* hb = hw.data() + (size_t)b * np;
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return hb[x] < hb[y]; });
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return std::fabs(hb[x]) > std::fabs(hb[y]); });
  }
  return EUVL_OK;
}

constexpr int SUB_DECLINE = 1;  // leading_by_subspace: not accepted, use the full decomposition

size_t subspace_bytes(int n, int B, int nk) {
  const size_t nb = SUB_NB;
  return sizeof(double2) * B * (3 * (size_t)n * nb + 7 * nb * nb + 2 * (size_t)n * nk + nb * nk) +
         sizeof(double) * B * (nb + nk) + sizeof(int) * B * (nk + 1) + 16 * 256;
}

// The leading nk eigenpairs of the B Hermitian n x n matrices T by block
// subspace iteration with Rayleigh-Ritz steps (see SUB_NB).  The first
// Rayleigh-Ritz step comes after 8 iterations; from its residuals and the
// ratio |theta_SUB_NB| / |theta_nk| of its Ritz values the iterations still
// needed for the tolerance are predicted (plus 2), and the next step checks.
// Accepted when every leading Ritz pair has |T f - theta f| <= 1e-13 |theta_1|;
// SUB_DECLINE when a block loses rank, the ratio is above 0.8, or four steps
// have not reached the tolerance.  Writes d_alpha and d_phi as the full method.
int leading_by_subspace(euvl_handle h, int n, int B, const double2* T, int nk, double* d_alpha, double* d_phi,
                        int* iters) {
  using namespace euvl;
  const int nb = SUB_NB;
  const size_t blk = sizeof(double2) * (size_t)n * nb * B, sq = sizeof(double2) * (size_t)nb * nb * B;
  const size_t fk = sizeof(double2) * (size_t)n * nk * B;
  double2* Y = (double2*)ws_take(h, blk);
  double2* Q = (double2*)ws_take(h, blk);
  double2* YH = (double2*)ws_take(h, blk);
  double2* G = (double2*)ws_take(h, sq);
  double2* Mf = (double2*)ws_take(h, sq);
  double2* H = (double2*)ws_take(h, sq);
  double2* A = (double2*)ws_take(h, sq);
  double2* V = (double2*)ws_take(h, sq);
  double2* A2 = (double2*)ws_take(h, sq);
  double2* V2 = (double2*)ws_take(h, sq);
  double2* F = (double2*)ws_take(h, fk);
  double2* W = (double2*)ws_take(h, fk);
  double2* U = (double2*)ws_take(h, sizeof(double2) * (size_t)nb * nk * B);
  double* w = (double*)ws_take(h, sizeof(double) * nb * B);
  double* res = (double*)ws_take(h, sizeof(double) * nk * B);
  int* pos_d = (int*)ws_take(h, sizeof(int) * nk * B);
  int* fail = (int*)ws_take(h, sizeof(int) * B);
  if (!Y || !Q || !YH || !G || !Mf || !H || !A || !V || !A2 || !V2 || !F || !W || !U || !w || !res || !pos_d || !fail) {
    set_error("euvl_tcc_kernels: workspace carve failed");
    return EUVL_ERR_NOMEM;
  }
  const double2 one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  const int64_t snb = (int64_t)n * nb, snn = (int64_t)n * n, sbb = (int64_t)nb * nb;
  // Qout = Yin M with M from the Gram matrix of Yin (Qout != Yin)
  auto orth = [&](const double2* Yin, double2* Qout) -> int {
    hipLaunchKernelGGL(conj_transpose_kernel, grid1(snb * B), dim3(256), 0, h->stream, n, nb, Yin, YH, B);
    EUVL_LAUNCH_CHECK();
    EUVL_TRY(zgemm(h, nb, nb, n, one, YH, n, snb, Yin, nb, snb, zero, G, nb, sbb, zero, B));
    hipLaunchKernelGGL(gram_orth_kernel<1024>, dim3(B), dim3(1024), 0, h->stream, (const double2*)G, Mf, fail);
    EUVL_LAUNCH_CHECK();
    EUVL_TRY(zgemm(h, n, nb, nb, one, Yin, nb, snb, Mf, nb, sbb, zero, Qout, nb, snb, zero, B));
    return EUVL_OK;
  };
  hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(64), 0, h->stream, fail, B);
  EUVL_LAUNCH_CHECK();
  hipLaunchKernelGGL(start_block_kernel, grid1(snb * B), dim3(256), 0, h->stream, snb * B, Y);
  EUVL_LAUNCH_CHECK();
  EUVL_TRY(orth(Y, Q));
  const double tol = 1e-13;
  int it = 0, next = 8;
  for (int rr = 0;;) {
    EUVL_TRY(zgemm(h, n, nb, n, one, T, n, snn, Q, nb, snb, zero, Y, nb, snb, zero, B));
    EUVL_TRY(orth(Y, Q));
    if (++it < next) continue;
    // Rayleigh-Ritz on a block orthonormalised twice (to rounding)
    EUVL_TRY(orth(Q, Y));
    std::swap(Q, Y);
    EUVL_TRY(zgemm(h, n, nb, n, one, T, n, snn, Q, nb, snb, zero, Y, nb, snb, zero, B));
    hipLaunchKernelGGL(conj_transpose_kernel, grid1(snb * B), dim3(256), 0, h->stream, n, nb, (const double2*)Q, YH,
                       B);
    EUVL_LAUNCH_CHECK();
    EUVL_TRY(zgemm(h, nb, nb, n, one, YH, n, snb, Y, nb, snb, zero, H, nb, sbb, zero, B));
    hipLaunchKernelGGL(hermitize_kernel, grid1(sbb * B), dim3(256), 0, h->stream, nb, H, B);
    EUVL_LAUNCH_CHECK();
    hipLaunchKernelGGL(pair_order_init_kernel, grid1(sbb * B), dim3(256), 0, h->stream, nb, nb, (const double2*)H, A,
                       V, B);
    EUVL_LAUNCH_CHECK();
    EUVL_TRY(jacobi_sweeps(h, nb, nb, A, V, A2, V2, B, nullptr));
    std::vector<double> hw;
    std::vector<std::vector<int>> pos;
    EUVL_TRY(jacobi_order(h, nb, nb, A, w, B, hw, pos));
    std::vector<int> sel((size_t)nk * B);
    std::vector<double> theta((size_t)nk * B);
    double rate = 0.0;
    for (int b = 0; b < B; ++b) {
      for (int m = 0; m < nk; ++m) {
        sel[(size_t)b * nk + m] = pos[b][m];
        theta[(size_t)b * nk + m] = hw[(size_t)b * nb + pos[b][m]];
      }
      const double lead = std::fabs(hw[(size_t)b * nb + pos[b][nk - 1]]);
      const double last = std::fabs(hw[(size_t)b * nb + pos[b][nb - 1]]);
      rate = std::max(rate, lead > 0.0 ? last / lead : INFINITY);
    }
    EUVL_HIP_CHECK(hipMemcpyAsync(pos_d, sel.data(), sizeof(int) * nk * B, hipMemcpyHostToDevice, h->stream));
    EUVL_HIP_CHECK(hipMemcpyAsync(d_alpha, theta.data(), sizeof(double) * nk * B, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(ritz_select_kernel, grid1((int64_t)nb * nk * B), dim3(256), 0, h->stream, nb, nk,
                       (const double2*)V, (const int*)pos_d, U, B);
    EUVL_LAUNCH_CHECK();
    EUVL_TRY(zgemm(h, n, nk, nb, one, Q, nb, snb, U, nk, (int64_t)nb * nk, zero, F, nk, (int64_t)n * nk, zero, B));
    EUVL_TRY(zgemm(h, n, nk, n, one, T, n, snn, F, nk, (int64_t)n * nk, zero, W, nk, (int64_t)n * nk, zero, B));
    hipLaunchKernelGGL(ritz_residual_kernel, dim3(nk, B), dim3(256), 0, h->stream, n, nk, (const double2*)W,
                       (const double2*)F, (const double*)d_alpha, res);
    EUVL_LAUNCH_CHECK();
    EUVL_HIP_CHECK(hipMemcpyAsync(h->h_pinned, res, sizeof(double) * nk * B, hipMemcpyDeviceToHost, h->stream));
    int* hfail = (int*)(h->h_pinned + nk * B);
    EUVL_HIP_CHECK(hipMemcpyAsync(hfail, fail, sizeof(int) * B, hipMemcpyDeviceToHost, h->stream));
    EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));  // sel, theta are read by the copies above until here
    double worst = 0.0;
    bool lost = false;
    for (int b = 0; b < B; ++b) {
      lost = lost || hfail[b] != 0;
      const double t0 = std::fabs(theta[(size_t)b * nk]);
      for (int m = 0; m < nk; ++m) {
        const double r = std::sqrt(h->h_pinned[(size_t)b * nk + m]) / t0;
        worst = std::max(worst, std::isnan(r) ? INFINITY : r);
      }
    }
    if (lost) return SUB_DECLINE;
    if (worst <= tol) {
      hipLaunchKernelGGL(transpose_kernel, grid1((int64_t)n * nk * B), dim3(256), 0, h->stream, n, nk,
                         (const double2*)F, (double2*)d_phi, B);
      EUVL_LAUNCH_CHECK();
      if (iters) *iters = it;
      return EUVL_OK;
    }
    if (++rr >= 4 || !(rate < 0.8)) return SUB_DECLINE;
    next = it + std::max(2, (int)std::ceil(std::log(tol / worst) / std::log(rate)) + 2);
    if (next > 160) return SUB_DECLINE;
  }
}

}  // namespace

extern "C" int euvl_tcc_kernels(euvl_handle h, const euvl_em_params* p, int ncut, const int* h_pupil, int n_src,
                                const double* h_src, double kx0, double ky0, double z0, double z, int nk,
                                double* d_alpha, double* d_phi, int* h_sweeps) {
  using namespace euvl;
  if (!h || !p || ncut <= 1 || !h_pupil || n_src <= 0 || !h_src || nk <= 0 || nk > ncut || !d_alpha || !d_phi) {
    set_error("euvl_tcc_kernels: bad argument");
    return EUVL_ERR_ARG;
  }
  const int n = ncut, B = 3, S = n_src;
  // parallel Jacobi over a round-robin tournament of np = n rounded up to even
  // players, stored in pair order (see pair_next)
  const int np = n + (n & 1);
  Upload up;
  const size_t o_pup = up.add(h_pupil, sizeof(int) * 2 * n);
  const size_t o_src = up.add(h_src, sizeof(double) * 2 * S);
  const size_t mat = sizeof(double2) * (size_t)n * n * B;
  const size_t matp = sizeof(double2) * (size_t)np * np * B;
  const size_t pb = sizeof(double2) * (size_t)n * 3 * S;
  const bool sub = nk <= SUB_MAX_NK && n >= 2 * SUB_NB;
  EUVL_TRY(ws_reserve(h, up.host.size() + mat + 4 * matp + 4 * pb + sizeof(double) * np * B +
                             sizeof(int) * nk * B + 8192 + (sub ? subspace_bytes(n, B, nk) : 0)));
  char* blk = (char*)ws_take(h, up.host.size());
  double2* T0 = (double2*)ws_take(h, mat);
  double2* P = (double2*)ws_take(h, pb);
  double2* Q = (double2*)ws_take(h, 3 * pb);
  if (!blk || !T0 || !P || !Q) {
    set_error("euvl_tcc_kernels: workspace carve failed");
    return EUVL_ERR_NOMEM;
  }
  EUVL_HIP_CHECK(hipMemcpyAsync(blk, up.host.data(), up.host.size(), hipMemcpyHostToDevice, h->stream));
  TccSetup a{p->k, p->dx, p->dy, p->na, kx0, ky0, z0, z, p->mx, p->my};
  hipLaunchKernelGGL(tcc_phase_kernel, grid1((int64_t)n * S), dim3(256), 0, h->stream, a, (const int*)(blk + o_pup),
                     n, (const double*)(blk + o_src), S, P, Q);
  EUVL_LAUNCH_CHECK();
  // TCC_w = P Q_w  (n x 3S x n), the three weights as a batch sharing P
  const double2 one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  EUVL_TRY(zgemm(h, n, n, 3 * S, one, P, 3 * S, 0, Q, n, (int64_t)3 * S * n, zero, T0, n, (int64_t)n * n, zero, B));
  if (sub) {
    int iters = 0;
    const int rc = leading_by_subspace(h, n, B, T0, nk, d_alpha, d_phi, &iters);
    if (rc == EUVL_OK) {
      if (h_sweeps) *h_sweeps = -iters;
      EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));  // the upload block is read until here
      return EUVL_OK;
    }
    if (rc != SUB_DECLINE) return rc;
  }
  double2* A = (double2*)ws_take(h, matp);
  double2* V = (double2*)ws_take(h, matp);
  double2* A2 = (double2*)ws_take(h, matp);
  double2* V2 = (double2*)ws_take(h, matp);
  double* w = (double*)ws_take(h, sizeof(double) * np * B);
  int* order_d = (int*)ws_take(h, sizeof(int) * nk * B);
  if (!A || !V || !A2 || !V2 || !w || !order_d) {
    set_error("euvl_tcc_kernels: workspace carve failed");
    return EUVL_ERR_NOMEM;
  }
  hipLaunchKernelGGL(pair_order_init_kernel, grid1((int64_t)np * np * B), dim3(256), 0, h->stream, n, np, T0, A, V,
                     B);
  EUVL_LAUNCH_CHECK();
  EUVL_TRY(jacobi_sweeps(h, n, np, A, V, A2, V2, B, h_sweeps));
  // eigenvalues in output order (|alpha| descending, the reference's tie
  // order), the leading nk of each set
  std::vector<double> hw;
  std::vector<std::vector<int>> pos;
  EUVL_TRY(jacobi_order(h, n, np, A, w, B, hw, pos));
  std::vector<int> order((size_t)nk * B);
  std::vector<double> halpha((size_t)nk * B);
  for (int b = 0; b < B; ++b)
    for (int m = 0; m < nk; ++m) {
      order[(size_t)b * nk + m] = pos[b][m];
      halpha[(size_t)b * nk + m] = hw[(size_t)b * np + pos[b][m]];
    }
  EUVL_HIP_CHECK(hipMemcpyAsync(order_d, order.data(), sizeof(int) * nk * B, hipMemcpyHostToDevice, h->stream));
  EUVL_HIP_CHECK(hipMemcpyAsync(d_alpha, halpha.data(), sizeof(double) * nk * B, hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(gather_kernel, grid1((int64_t)nk * n * B), dim3(256), 0, h->stream, n, np, nk, V, order_d,
                     (double2*)d_phi, B);
  EUVL_LAUNCH_CHECK();
  EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));
  return EUVL_OK;
}

extern "C" int euvl_socs_image(euvl_handle h, const euvl_em_params* p, const double* d_fmask, int lpmax_x,
                               int lpmax_y, const double* ampvc, const double* ampab, const double* phase0,
                               const double* d_a0, const double* d_ax, const double* d_ay, int ncut,
                               const int* h_pupil, const double* d_alpha, const double* d_phi, int nk, int nsocs,
                               int nsocsxy, int nout, double kx0, double ky0, double z0, int nx, int ny,
                               double* d_img) {
  using namespace euvl;
  if (!h || !p || !d_fmask || !ampvc || !ampab || !phase0 || !d_a0 || !d_ax || !d_ay || ncut <= 0 || !h_pupil ||
      !d_alpha || !d_phi || nk <= 0 || nsocs < 0 || nsocsxy < 0 || nsocs > nk || nsocsxy > nsocs || nout < 2 ||
      (nout & 1) || nx <= 0 || ny <= 0 || !d_img) {
    set_error("euvl_socs_image: bad argument");
    return EUVL_ERR_ARG;
  }
  int lpx = 0, lpy = 0;
  std::vector<char> seen((size_t)nout * nout, 0);
  for (int i = 0; i < ncut; ++i) {
    lpx = std::max(lpx, std::abs(h_pupil[2 * i]));
    lpy = std::max(lpy, std::abs(h_pupil[2 * i + 1]));
    const int px = ((h_pupil[2 * i] % nout) + nout) % nout, py = ((h_pupil[2 * i + 1] % nout) + nout) % nout;
    if (seen[(size_t)px * nout + py]++) {
      set_error("euvl_socs_image: pupil points collide on the %d x %d grid", nout, nout);
      return EUVL_ERR_ARG;
    }
  }
  if (lpx > lpmax_x || lpy > lpmax_y) {
    set_error("euvl_socs_image: pupil outside the thin-mask window");
    return EUVL_ERR_ARG;
  }
  const int nl = 2 * lpx + 1, nm = 2 * lpy + 1;
  const int nf = nsocs + 5 * nsocsxy;
  std::vector<int> fdesc;
  for (int m = 0; m < nsocs; ++m) fdesc.insert(fdesc.end(), {0, 0, m});
  for (int m = 0; m < nsocsxy; ++m) {
    fdesc.insert(fdesc.end(), {0, 1, m});  // f0x = A0 phix
    fdesc.insert(fdesc.end(), {0, 2, m});  // f0y = A0 phiy
    fdesc.insert(fdesc.end(), {1, 1, m});  // fx  = Axsx phix
    fdesc.insert(fdesc.end(), {2, 2, m});  // fy  = Axsy phiy
    fdesc.insert(fdesc.end(), {3, 0, m});  // fxy = Axsxy phi0
  }
  auto tw_out = twiddles(nout, 1.0), tw_outm = twiddles(nout, -1.0), twx = twiddles(nx, 1.0), twy = twiddles(ny, 1.0);
  Upload up;
  const size_t o_pup = up.add(h_pupil, sizeof(int) * 2 * ncut);
  const size_t o_fd = up.add(fdesc.data(), sizeof(int) * fdesc.size());
  const size_t o_two = up.add(tw_out.data(), sizeof(double2) * nout);
  const size_t o_twm = up.add(tw_outm.data(), sizeof(double2) * nout);
  const size_t o_twx = up.add(twx.data(), sizeof(double2) * nx);
  const size_t o_twy = up.add(twy.data(), sizeof(double2) * ny);
  const size_t b_amps = sizeof(double2) * 4 * ncut;
  const size_t b_F = sizeof(double2) * (size_t)std::max(nf, 1) * nl * nm;
  const size_t b_G = sizeof(double2) * (size_t)std::max(nf, 1) * nl * nout;
  const size_t b_fld = sizeof(double2) * (size_t)std::max(nf, 1) * nout * nout;
  const size_t b_small = sizeof(double) * nout * nout;
  const size_t b_S = sizeof(double2) * nout * nout;
  const size_t b_T = sizeof(double2) * (size_t)nout * ny;
  EUVL_TRY(ws_reserve(h, up.host.size() + b_amps + b_F + b_G + b_fld + b_small + 2 * b_S + b_T + 16384));
  char* blk = (char*)ws_take(h, up.host.size());
  double2* amps = (double2*)ws_take(h, b_amps);
  double2* F = (double2*)ws_take(h, b_F);
  double2* G = (double2*)ws_take(h, b_G);
  double2* fld = (double2*)ws_take(h, b_fld);
  double* small = (double*)ws_take(h, b_small);
  double2* Rp = (double2*)ws_take(h, b_S);
  double2* Sp = (double2*)ws_take(h, b_S);
  double2* T = (double2*)ws_take(h, b_T);
  EUVL_TRY(upload_async(h, blk, up.host.data(), up.host.size()));
  SocsSetup a;
  a.ampvc = make_double2(ampvc[0], ampvc[1]);
  a.ampab = make_double2(ampab[0], ampab[1]);
  a.phase0 = make_double2(phase0[0], phase0[1]);
  a.k = p->k;
  a.dx = p->dx;
  a.dy = p->dy;
  a.kx0 = kx0;
  a.ky0 = ky0;
  a.z0 = z0;
  a.lpmx = lpmax_x;
  a.lpmy = lpmax_y;
  a.lpx = lpx;
  a.lpy = lpy;
  const int* pup = (const int*)(blk + o_pup);
  hipLaunchKernelGGL(socs_amps_kernel, dim3((ncut + 255) / 256), dim3(256), 0, h->stream, a, (const double2*)d_fmask,
                     (const double2*)d_a0, (const double2*)d_ax, (const double2*)d_ay, pup, ncut, amps);
  EUVL_LAUNCH_CHECK();
  const bool f32 = (h->prec_mask & EUVL_PREC_FP32_IMAGING) != 0;
  if (nf > 0) {
    EUVL_TRY(fill(h, (int64_t)nf * nl * nm, F, make_double2(0.0, 0.0)));
    hipLaunchKernelGGL(socs_spectra_kernel, grid1((int64_t)nf * ncut), dim3(256), 0, h->stream, a, nf,
                       (const int*)(blk + o_fd), amps, (const double2*)d_phi, nk, pup, ncut, F);
    EUVL_LAUNCH_CHECK();
    if (f32) {
      // FP32 copies of the twiddles go after the FP32 fields (half of the FP64 field buffer)
      float2* tw32 = (float2*)T;
      hipLaunchKernelGGL(socs_tw32_kernel, dim3((nout + 255) / 256), dim3(256), 0, h->stream, nout,
                         (const double2*)(blk + o_two), tw32);
      EUVL_LAUNCH_CHECK();
      hipLaunchKernelGGL(socs_ypass32_kernel, grid1((int64_t)nf * nl * nout), dim3(256), 0, h->stream, nf, nl, nm,
                         lpy, nout, F, tw32, (float2*)G);
      EUVL_LAUNCH_CHECK();
      hipLaunchKernelGGL(socs_xpass32_kernel, grid1((int64_t)nf * nout * nout), dim3(256), 0, h->stream, nf, nl,
                         lpx, nout, (const float2*)G, tw32, (float2*)fld);
      EUVL_LAUNCH_CHECK();
    } else {
      hipLaunchKernelGGL(socs_ypass_kernel, grid1((int64_t)nf * nl * nout), dim3(256), 0, h->stream, nf, nl, nm,
                         lpy, nout, F, (const double2*)(blk + o_two), G);
      EUVL_LAUNCH_CHECK();
      hipLaunchKernelGGL(socs_xpass_kernel, grid1((int64_t)nf * nout * nout), dim3(256), 0, h->stream, nf, nl, lpx,
                         nout, G, (const double2*)(blk + o_two), fld);
      EUVL_LAUNCH_CHECK();
    }
  }
  if (f32 && nf > 0)
    hipLaunchKernelGGL(socs_combine32_kernel, dim3((nout * nout + 255) / 256), dim3(256), 0, h->stream, nout, nsocs,
                       nsocsxy, (const float2*)fld, d_alpha, nk, small);
  else
    hipLaunchKernelGGL(socs_combine_kernel, dim3((nout * nout + 255) / 256), dim3(256), 0, h->stream, nout, nsocs,
                       nsocsxy, fld, d_alpha, nk, small);
  EUVL_LAUNCH_CHECK();
  hipLaunchKernelGGL(socs_spec_y_kernel, dim3((nout * nout + 255) / 256), dim3(256), 0, h->stream, nout, small,
                     (const double2*)(blk + o_twm), Rp);
  EUVL_LAUNCH_CHECK();
  hipLaunchKernelGGL(socs_spec_x_kernel, dim3((nout * nout + 255) / 256), dim3(256), 0, h->stream, nout,
                     (const double2*)Rp, (const double2*)(blk + o_twm), Sp);
  EUVL_LAUNCH_CHECK();
  hipLaunchKernelGGL(socs_up_y_kernel, grid1((int64_t)nout * ny), dim3(256), 0, h->stream, nout, ny, Sp,
                     (const double2*)(blk + o_twy), T);
  EUVL_LAUNCH_CHECK();
  hipLaunchKernelGGL(socs_up_x_kernel, grid1((int64_t)nx * ny), dim3(256), 0, h->stream, nout, nx, ny, T,
                     (const double2*)(blk + o_twx), d_img);
  EUVL_LAUNCH_CHECK();
  return EUVL_OK;
}
