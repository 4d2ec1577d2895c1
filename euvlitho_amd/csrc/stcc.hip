// The TCCEE image of elitho/stcc.py (cnn/intensity/stccsocs/elitho/stcc.py:12-431):
// Hopkins' sum I(x) = Re sum_{i,j} TCC_ij E_i(s0) . conj(E_j(s0) + 2 (...)) e^{i(k_i - k_j)x}
// with the three TCCs of one polarisation built from the uniform k-source,
// evaluated without an eigendecomposition:
//   stcc_fields_kernel  the twelve field terms of every pupil point (stcc.py:166-295)
//   stcc_p_kernel       P[i][s] = pupil phase of point i under source s (0 outside
//                       the NA) and the weighted conjugates w_q(s) conj(P[j][s])
//   ZGEMM               TCC_q = P W_q^H / SDIV, batched over q (stcc.py:12-103)
//   stcc_tccee_kernel   TCCEE over the difference window |dl| <= 2 lpmax_x,
//                       |dm| <= 2 lpmax_y: a gather over i of the (i, j) pairs
//                       whose orders differ by (dl, dm), in a fixed order
//                       (stcc.py:298-391)
//   stcc_image_y/x      ifft2(TCCEE, norm="forward").real on XDIV x XDIV as two
//                       separable passes over the window (stcc.py:428-431)
// The reference's quirks are kept: the propagation phase divides by k z0, the
// field terms use kx0 + 2 pi l / dx / 2, the Y branch always adds the M3D
// arrays, and the image grid is XDIV in both axes.
#include <cmath>
#include <vector>

#include "ew.h"
#include "optics.h"

namespace {

struct StccArgs {
  double k, kx0, ky0, z0, dx, dy, na;
  int mx, my, lpx, lpy, xdiv;
  int pol;  // 0 = X, 1 = Y
  double2 ab, vc, ph0;
};

// E terms, [12][ncut]: (s0, sx, sy, sxy) x (x, y, z)
__global__ void stcc_fields_kernel(StccArgs a, int ncut, const int* __restrict__ pupil,
                                   const double2* __restrict__ fmask, const double2* __restrict__ a0,
                                   const double2* __restrict__ ax, const double2* __restrict__ ay,
                                   double2* __restrict__ E) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncut) return;
  const int l = pupil[2 * i], m = pupil[2 * i + 1];
  const int nouty = 2 * a.lpy + 1;
  const int w = (l + a.lpx) * nouty + (m + a.lpy);
  // fampxx (stcc.py:415-418): FFT of hfpattern, propagation phase, / phasexx
  double2 f = cmul(make_double2(a.ab.x - a.vc.x, a.ab.y - a.vc.y), fmask[w]);
  if (l == 0 && m == 0) f = cadd(f, a.vc);
  const double kxp = 2.0 * PI * l / a.dx, kyp = 2.0 * PI * m / a.dy;
  const double arg = -(a.kx0 * kxp + kxp * kxp / 2 + a.ky0 * kyp + kyp * kyp / 2) / (a.k * a.z0);
  double s, c;
  sincos(arg, &s, &c);
  f = cdiv(cmul(f, make_double2(c, s)), a.ph0);
  double2 A = f, dA = make_double2(0.0, 0.0), eA = make_double2(0.0, 0.0);
  const double gx = a.dx / (2 * PI), gy = a.dy / (2 * PI);
  if (a0) {
    A = cadd(A, a0[w]);
    dA = cscale(ax[w], gx);
    eA = cscale(ay[w], gy);
  }
  const double k = a.k;
  const double kx = a.kx0 + 2 * PI * l / a.dx / 2.0, ky = a.ky0 + 2 * PI * m / a.dy / 2.0;
  const double kz = -sqrt(k * k - kx * kx - ky * ky);
  // i * s * v  and  -i * s * v
  auto iz = [](double sc, double2 v) { return make_double2(-sc * v.y, sc * v.x); };
  double2 t[12];
  if (a.pol == 0) {
    t[0] = csub(iz(k, A), iz(kx * kx / k, A));                                   // Exs0
    t[3] = csub(cadd(iz(-2.0 * kx / k, A), iz(k, dA)), iz(kx * kx / k, dA));     // Exsx
    t[6] = csub(iz(k, eA), iz(kx * kx / k, eA));                                 // Exsy
    t[1] = iz(-kx * ky / k, A);                                                  // Eys0
    t[4] = cadd(iz(-ky / k, A), iz(-kx * ky / k, dA));                           // Eysx
    t[7] = cadd(iz(-kx / k, A), iz(-kx * ky / k, eA));                           // Eysy
    t[2] = iz(-kx * kz / k, A);                                                  // Ezs0
    t[5] = cadd(iz(-kz / k, A), iz(-kx * kz / k, dA));                           // Ezsx
    t[8] = iz(-kx * kz / k, eA);                                                 // Ezsy
  } else {
    t[0] = iz(-kx * ky / k, A);                                                  // Exs0
    t[3] = cadd(iz(-ky / k, A), iz(-kx * ky / k, dA));                           // Exsx
    t[6] = cadd(iz(-kx / k, A), iz(-kx * ky / k, eA));                           // Exsy
    t[1] = csub(iz(1.0 / k, A), iz(ky * ky / k, A));                             // Eys0
    t[4] = csub(iz(1.0 / k, dA), iz(ky * ky / k, dA));                           // Eysx
    t[7] = csub(cadd(iz(-2.0 * ky / k, A), iz(1.0 / k, eA)), iz(ky * ky / k, eA));  // Eysy
    t[2] = iz(-ky * kz / k, A);                                                  // Ezs0
    t[5] = iz(-ky * kz / k, dA);                                                 // Ezsx
    t[8] = cadd(iz(-kz / k, A), iz(-ky * kz / k, eA));                           // Ezsy
  }
  for (int q = 0; q < 3; ++q)  // Esxy = Esx / gx * l / 2 + Esy / gy * m / 2
    t[9 + q] = cadd(cscale(t[3 + q], l / gx / 2.0), cscale(t[6 + q], m / gy / 2.0));
  for (int q = 0; q < 12; ++q) E[(int64_t)q * ncut + i] = t[q];
}

// P[i][s] and W_q[s][j] = w_q(s) conj(P[j][s]), q = 0 (1), 1 (sx), 2 (sy) over
// the denominator of the polarisation
__global__ void stcc_p_kernel(StccArgs a, int ncut, int ns, const int* __restrict__ pupil,
                              const double* __restrict__ src, double2* __restrict__ P, double2* __restrict__ W) {
  const int64_t total = (int64_t)ncut * ns;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / ns), s = (int)(e % ns);
    const double sx = src[2 * s], sy = src[2 * s + 1];
    const double ksx = 2 * PI / a.dx * pupil[2 * i] + sx, ksy = 2 * PI / a.dy * pupil[2 * i + 1] + sy;
    const bool in = (a.mx * a.mx * ksx * ksx + a.my * a.my * ksy * ksy) <= (a.k * a.na) * (a.k * a.na);
    double2 p = make_double2(0.0, 0.0);
    if (in) {
      const double arg = ((ksx + a.kx0) * (ksx + a.kx0) + (ksy + a.ky0) * (ksy + a.ky0)) / (2 * a.k) * a.z0;
      double sn, cs;
      sincos(arg, &sn, &cs);
      p = make_double2(cs, sn);
    }
    P[(int64_t)i * ns + s] = p;
    const double den = (a.pol == 0) ? a.k * a.k - (a.kx0 + sx) * (a.kx0 + sx) : a.k * a.k - (a.ky0 + sy) * (a.ky0 + sy);
    const double2 pc = cconj(p);
    const int64_t per = (int64_t)ns * ncut;
    W[(int64_t)s * ncut + i] = cscale(pc, 1.0 / den);
    W[per + (int64_t)s * ncut + i] = cscale(pc, sx / den);
    W[2 * per + (int64_t)s * ncut + i] = cscale(pc, sy / den);
  }
}

// TCCEE window [dl + 2 lpx][dm + 2 lpy]; idx: pupil index of window point
// (l + lpx, m + lpy) or -1
__global__ void stcc_tccee_kernel(int ncut, int lpx, int lpy, const int* __restrict__ pupil,
                                  const int* __restrict__ idx, const double2* __restrict__ T,
                                  const double2* __restrict__ E, double2* __restrict__ out) {
  const int wl = 4 * lpx + 1, wm = 4 * lpy + 1;
  const int bin = blockIdx.x;
  if (bin >= wl * wm) return;
  const int dl = bin / wm - 2 * lpx, dm = bin % wm - 2 * lpy;
  const int nouty = 2 * lpy + 1;
  const int64_t nn = (int64_t)ncut * ncut;
  double2 acc = make_double2(0.0, 0.0);
  __shared__ double2 red[256];
  for (int i = threadIdx.x; i < ncut; i += blockDim.x) {
    const int lj = pupil[2 * i] - dl, mj = pupil[2 * i + 1] - dm;
    if (lj < -lpx || lj > lpx || mj < -lpy || mj > lpy) continue;
    const int j = idx[(lj + lpx) * nouty + (mj + lpy)];
    if (j < 0) continue;
    double2 s00 = make_double2(0, 0), s0x = s00, s0y = s00, s0xy = s00;
    for (int c = 0; c < 3; ++c) {
      const double2 e0 = E[(int64_t)c * ncut + i];
      s00 = cadd(s00, cmul(e0, cconj(E[(int64_t)c * ncut + j])));
      s0x = cadd(s0x, cmul(e0, cconj(E[(int64_t)(3 + c) * ncut + j])));
      s0y = cadd(s0y, cmul(e0, cconj(E[(int64_t)(6 + c) * ncut + j])));
      s0xy = cadd(s0xy, cmul(e0, cconj(E[(int64_t)(9 + c) * ncut + j])));
    }
    const int64_t ij = (int64_t)i * ncut + j;
    const double2 t0 = T[ij], tx = T[nn + ij], ty = T[2 * nn + ij];
    acc = cadd(acc, cmul(t0, s00));
    acc = cadd(acc, cscale(cmul(tx, s0x), 2.0));
    acc = cadd(acc, cscale(cmul(ty, s0y), 2.0));
    acc = cadd(acc, cscale(cmul(t0, s0xy), 2.0));
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = cadd(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[bin] = red[0];
}

// e^{2 pi i n / X} with the phase reduced exactly in integers first
__device__ inline double2 twiddle(long long n, int X) {
  const long long r = ((n % X) + X) % X;
  double s, c;
  sincospi(2.0 * (double)r / (double)X, &s, &c);
  return make_double2(c, s);
}

// twiddle tables: tw_x[il][x] = e^{2 pi i (il - 2 lpx) x / X} (wl rows), then
// tw_y[im][y] = e^{2 pi i (im - 2 lpy) y / X} (wm rows): each factor once,
// not once per image pixel
__global__ void stcc_twiddle_kernel(int lpx, int lpy, int X, double2* __restrict__ tw) {
  const int wl = 4 * lpx + 1, wm = 4 * lpy + 1;
  const int64_t total = (int64_t)(wl + wm) * X;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e / X), c = (int)(e % X);
    tw[e] = (r < wl) ? twiddle((long long)(r - 2 * lpx) * c, X) : twiddle((long long)(r - wl - 2 * lpy) * c, X);
  }
}

// G[dl][y] = sum_dm T[dl][dm] e^{2 pi i dm y / X}
__global__ void stcc_image_y_kernel(int lpx, int lpy, int X, const double2* __restrict__ T,
                                    const double2* __restrict__ tw_y, double2* __restrict__ G) {
  const int wl = 4 * lpx + 1, wm = 4 * lpy + 1;
  const int64_t total = (int64_t)wl * X;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int il = (int)(e / X), y = (int)(e % X);
    double2 acc = make_double2(0.0, 0.0);
    for (int im = 0; im < wm; ++im) acc = cadd(acc, cmul(T[(int64_t)il * wm + im], tw_y[(int64_t)im * X + y]));
    G[e] = acc;
  }
}

// img[x][y] = Re sum_dl G[dl][y] e^{2 pi i dl x / X}
__global__ void stcc_image_x_kernel(int lpx, int X, const double2* __restrict__ G, const double2* __restrict__ tw_x,
                                    double* __restrict__ img) {
  const int wl = 4 * lpx + 1;
  const int64_t total = (int64_t)X * X;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(e / X), y = (int)(e % X);
    double acc = 0.0;
    for (int il = 0; il < wl; ++il) {
      const double2 g = G[(int64_t)il * X + y], t = tw_x[(int64_t)il * X + x];
      acc += g.x * t.x - g.y * t.y;
    }
    img[e] = acc;
  }
}

}  // namespace

extern "C" int euvl_stcc_image(euvl_handle h, char polar, const euvl_stcc_params* p, const double* d_fmask,
                               const double* ab, const double* vc, const double* phase0, const double* d_a0,
                               const double* d_ax, const double* d_ay, int ncut, const int* h_pupil, int n_src,
                               const double* h_src, double* d_img) {
  if (!h || !p || !d_fmask || !ab || !vc || !phase0 || ncut <= 0 || !h_pupil || n_src <= 0 || !h_src || !d_img ||
      (polar != 'X' && polar != 'Y') || p->lpmax_x < 0 || p->lpmax_y < 0 || p->xdiv <= 0 || !(p->k > 0.0)) {
    euvl::set_error("euvl_stcc_image: bad argument");
    return EUVL_ERR_ARG;
  }
  if ((d_a0 == nullptr) != (d_ax == nullptr) || (d_a0 == nullptr) != (d_ay == nullptr)) {
    euvl::set_error("euvl_stcc_image: a0, ax, ay must be all given or all null");
    return EUVL_ERR_ARG;
  }
  if (polar == 'Y' && !d_a0) {
    euvl::set_error("euvl_stcc_image: the Y branch reads the M3D arrays (stcc.py:239); pass a0, ax, ay");
    return EUVL_ERR_ARG;
  }
  const int lpx = p->lpmax_x, lpy = p->lpmax_y, noutx = 2 * lpx + 1, nouty = 2 * lpy + 1;
  std::vector<int> idx((size_t)noutx * nouty, -1);
  for (int i = 0; i < ncut; ++i) {
    const int l = h_pupil[2 * i], m = h_pupil[2 * i + 1];
    if (l < -lpx || l > lpx || m < -lpy || m > lpy) {
      euvl::set_error("euvl_stcc_image: pupil point %d (%d, %d) outside the window", i, l, m);
      return EUVL_ERR_ARG;
    }
    idx[(size_t)(l + lpx) * nouty + (m + lpy)] = i;
  }
  StccArgs a{p->k, p->kx0, p->ky0, p->z0, p->dx, p->dy, p->na, p->mx, p->my, lpx, lpy, p->xdiv, polar == 'Y' ? 1 : 0,
             make_double2(ab[0], ab[1]), make_double2(vc[0], vc[1]), make_double2(phase0[0], phase0[1])};
  Upload up;
  const size_t o_pup = up.add(h_pupil, sizeof(int) * 2 * ncut);
  const size_t o_src = up.add(h_src, sizeof(double) * 2 * n_src);
  const size_t o_idx = up.add(idx.data(), sizeof(int) * idx.size());
  const int wl = 4 * lpx + 1, wm = 4 * lpy + 1;
  const size_t bE = sizeof(double2) * 12 * ncut, bP = sizeof(double2) * (size_t)ncut * n_src;
  const size_t bT = sizeof(double2) * 3 * (size_t)ncut * ncut, bW = sizeof(double2) * (size_t)wl * wm;
  const size_t bG = sizeof(double2) * (size_t)wl * p->xdiv;
  const size_t bTw = sizeof(double2) * (size_t)(wl + wm) * p->xdiv;
  EUVL_TRY(euvl::ws_reserve(h, up.host.size() + bE + 4 * bP + bT + bW + bG + bTw + 9 * 256));
  char* blk = (char*)euvl::ws_take(h, up.host.size());
  double2* E = (double2*)euvl::ws_take(h, bE);
  double2* P = (double2*)euvl::ws_take(h, bP);
  double2* W = (double2*)euvl::ws_take(h, 3 * bP);
  double2* T = (double2*)euvl::ws_take(h, bT);
  double2* TW = (double2*)euvl::ws_take(h, bW);
  double2* G = (double2*)euvl::ws_take(h, bG);
  double2* TWID = (double2*)euvl::ws_take(h, bTw);
  EUVL_HIP_CHECK(hipMemcpyAsync(blk, up.host.data(), up.host.size(), hipMemcpyHostToDevice, h->stream));
  const int* d_pup = (const int*)(blk + o_pup);
  hipLaunchKernelGGL(stcc_fields_kernel, dim3((ncut + 127) / 128), dim3(128), 0, h->stream, a, ncut, d_pup,
                     (const double2*)d_fmask, (const double2*)d_a0, (const double2*)d_ax, (const double2*)d_ay, E);
  EUVL_LAUNCH_CHECK();
  hipLaunchKernelGGL(stcc_p_kernel, grid1((int64_t)ncut * n_src), dim3(256), 0, h->stream, a, ncut, n_src, d_pup,
                     (const double*)(blk + o_src), P, W);
  EUVL_LAUNCH_CHECK();
  {
    // TCC_q = P W_q / SDIV, q = 1, sx, sy (one batched product; A shared)
    euvl::PrecScope fp64(h, 0);
    EUVL_TRY(euvl::zgemm(h, ncut, ncut, n_src, make_double2(1.0 / n_src, 0.0), P, n_src, 0, W, ncut,
                         (int64_t)n_src * ncut, make_double2(0.0, 0.0), T, ncut, (int64_t)ncut * ncut,
                         make_double2(0.0, 0.0), 3));
  }
  hipLaunchKernelGGL(stcc_tccee_kernel, dim3(wl * wm), dim3(256), 0, h->stream, ncut, lpx, lpy, d_pup,
                     (const int*)(blk + o_idx), T, E, TW);
  EUVL_LAUNCH_CHECK();
  hipLaunchKernelGGL(stcc_twiddle_kernel, grid1((int64_t)(wl + wm) * p->xdiv), dim3(256), 0, h->stream, lpx, lpy,
                     p->xdiv, TWID);
  EUVL_LAUNCH_CHECK();
  hipLaunchKernelGGL(stcc_image_y_kernel, grid1((int64_t)wl * p->xdiv), dim3(256), 0, h->stream, lpx, lpy, p->xdiv,
                     TW, TWID + (size_t)wl * p->xdiv, G);
  EUVL_LAUNCH_CHECK();
  hipLaunchKernelGGL(stcc_image_x_kernel, grid1((int64_t)p->xdiv * p->xdiv), dim3(256), 0, h->stream, lpx, p->xdiv,
                     G, TWID, d_img);
  EUVL_LAUNCH_CHECK();
  EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));
  return EUVL_OK;
}
