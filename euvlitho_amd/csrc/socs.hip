// STCC-SOCS imaging: transmission cross coefficients from the source mesh,
// their Hermitian eigendecomposition (kernels), and the SOCS image.
//   eigenfunction.py:5-179  -> euvl_tcc_kernels
//   socs.py:12-154          -> euvl_socs_image
// TCC_w = sum_c P_c diag(w) P_c^H / S is one MFMA GEMM per weight (the three
// polarisation components stacked along K); the eigendecomposition is a
// parallel two-sided Jacobi method (round-robin pairs, n/2 rotations per
// step applied in one out-of-place launch over 2 x 2 blocks).
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

std::vector<std::complex<double>> twiddles(int n, double sign) {
  std::vector<std::complex<double>> t(n);
  const std::complex<double> zi(0.0, 1.0);
  for (int i = 0; i < n; ++i) t[i] = std::exp(sign * 2.0 * PI * zi * double(i) / double(n));
  return t;
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
  const int np = n + (n & 1), npair = np / 2, nsteps = np - 1;
  Upload up;
  const size_t o_pup = up.add(h_pupil, sizeof(int) * 2 * n);
  const size_t o_src = up.add(h_src, sizeof(double) * 2 * S);
  const size_t mat = sizeof(double2) * (size_t)n * n * B;
  const size_t matp = sizeof(double2) * (size_t)np * np * B;
  const size_t pb = sizeof(double2) * (size_t)n * 3 * S;
  EUVL_TRY(ws_reserve(h, up.host.size() + mat + 4 * matp + 4 * pb + sizeof(double) * np * B +
                             sizeof(int) * nk * B + 8192));
  char* blk = (char*)ws_take(h, up.host.size());
  double2* T0 = (double2*)ws_take(h, mat);
  double2* A = (double2*)ws_take(h, matp);
  double2* V = (double2*)ws_take(h, matp);
  double2* A2 = (double2*)ws_take(h, matp);
  double2* V2 = (double2*)ws_take(h, matp);
  double2* P = (double2*)ws_take(h, pb);
  double2* Q = (double2*)ws_take(h, 3 * pb);
  double* w = (double*)ws_take(h, sizeof(double) * np * B);
  int* order_d = (int*)ws_take(h, sizeof(int) * nk * B);
  if (!T0 || !A || !V || !A2 || !V2 || !P || !Q || !w || !order_d) {
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
  hipLaunchKernelGGL(pair_order_init_kernel, grid1((int64_t)np * np * B), dim3(256), 0, h->stream, n, np, T0, A, V,
                     B);
  EUVL_LAUNCH_CHECK();
  // Stop once the squared off-diagonal / total Frobenius ratio is at the
  // rounding floor of an n x n matrix, (2 n eps)^2, or stops falling below
  // 1e3 times it; Jacobi converges quadratically, so this takes few sweeps.
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
  if (h_sweeps) *h_sweeps = sweep;
  if (!(worst <= 1e3 * floor2)) {
    set_error("euvl_tcc_kernels: Jacobi off-diagonal ratio %.3e after %d sweeps (floor %.3e)", worst, sweep, floor2);
    return EUVL_ERR_NOT_CONVERGED;
  }
  // eigenvalues (the diagonal, one per position of the pair order; a sweep ends
  // in the initial order), stable sort by |alpha| descending (the reference's
  // bubble sort order), top nk; the padding position (odd n) is left out
  hipLaunchKernelGGL(diag_kernel, grid1((int64_t)np * B), dim3(256), 0, h->stream, np, A, w, B);
  EUVL_LAUNCH_CHECK();
  std::vector<double> hw((size_t)np * B);
  EUVL_HIP_CHECK(hipMemcpyAsync(hw.data(), w, sizeof(double) * np * B, hipMemcpyDeviceToHost, h->stream));
  EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));
  std::vector<int> order((size_t)nk * B);
  std::vector<double> halpha((size_t)nk * B);
  for (int b = 0; b < B; ++b) {
    // positions sorted by the index they hold, so that ties keep ascending index order
    std::vector<int> idx;
    for (int q = 0; q < n; ++q) {
      for (int pos = 0; pos < np; ++pos)
        if (pair_index0(pos, np) == q) {
          idx.push_back(pos);
          break;
        }
    }
    // ties keep ascending eigenvalue order, as after the reference's ascending eigh
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int x, int y) { return hw[(size_t)b * np + x] < hw[(size_t)b * np + y]; });
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) {
      return std::fabs(hw[(size_t)b * np + x]) > std::fabs(hw[(size_t)b * np + y]);
    });
    for (int m = 0; m < nk; ++m) {
      order[(size_t)b * nk + m] = idx[m];
      halpha[(size_t)b * nk + m] = hw[(size_t)b * np + idx[m]];
    }
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
  EUVL_HIP_CHECK(hipMemcpyAsync(blk, up.host.data(), up.host.size(), hipMemcpyHostToDevice, h->stream));
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
  EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));
  return EUVL_OK;
}
