// Image formation from EM amplitudes: the rigorous Abbe accumulation of
// cnn/intensity/abbe/intensity.cpp:276-408 and the M3D parameter fit of
// cnn/intensity/m3d/m3d.cpp:329-467.
//
// Abbe: the pupil field of every source point is scattered into a small dense
// spectrum grid (orders |l| <= lpx, |m| <= lpy), and the unnormalised inverse
// DFT onto the nx x ny image is evaluated separably (y first, then x) with
// |Ex|^2 + |Ey|^2 + |Ez|^2 summed over source points inside the x pass.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "ew.h"
#include "imaging.h"
#include "optics.h"

namespace {


using euvl::AbbeSetup;
using euvl::M3dSetup;

// Pupil fields: F[s][c][l+lpx][m+lpy] (c = x, y, z) for every source point s and pupil point i.
__global__ void abbe_field_kernel(AbbeSetup a, const double2* Ax, const int* modemap, const double* off,
                                  const int* src, int n_src, const int* pupil, int ncut, double2* F) {
  const int64_t total = (int64_t)n_src * ncut;
  const int nl = 2 * a.lpx + 1, nm = 2 * a.lpy + 1;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int s = (int)(e / ncut), i = (int)(e % ncut);
    const int o = src[3 * s], l0 = src[3 * s + 1], m0 = src[3 * s + 2];
    const int lin = pupil[2 * i], min_ = pupil[2 * i + 1];
    const double sx0 = off[4 * o], sy0 = off[4 * o + 1], fx = off[4 * o + 2], fy = off[4 * o + 3];
    const double kxn = 2.0 * PI / a.dx * fx + 2.0 * PI / a.dx * l0 + 2.0 * PI * lin / a.dx;
    const double kyn = 2.0 * PI / a.dy * fy + 2.0 * PI / a.dy * m0 + 2.0 * PI * min_ / a.dy;
    const double p2 = a.mx * a.mx * kxn * kxn + a.my * a.my * kyn * kyn;
    if (!(p2 <= (a.na * a.k) * (a.na * a.k))) continue;  // outside the pupil: stays zero
    // amplitude of mode (l0 + lin, m0 + min) for the incidence (l0, m0); -1000 when missing (intensity.cpp:255)
    double2 amp = make_double2(-1000.0, 0.0);
    const int ls = l0 + a.lsmx, ms = m0 + a.lsmy;
    const int l = l0 + lin, m = m0 + min_;
    if (ls >= 0 && ls < a.nsx && ms >= 0 && ms < a.nsy && abs(l) <= a.lmax && abs(m) <= a.mmax) {
      const int n = modemap[(l + a.lmax) * (2 * a.mmax + 1) + (m + a.mmax)];
      if (n >= 0) amp = Ax[(((int64_t)o * a.nsx + ls) * a.nsy + ms) * a.nrange + n];
    }
    // X: Ax = amp / sqrt(k^2 - kx^2) with kx = sx0 + 2 pi l0 / dx; Y: Ay = amp / sqrt(k^2 - ky^2)
    const double kin = a.pol ? sy0 + 2.0 * PI / a.dy * m0 : sx0 + 2.0 * PI / a.dx * l0;
    const double2 A = cscale(amp, 1.0 / sqrt(a.k * a.k - kin * kin));
    double R[3];
    rotate_col(a.k, a.mx, a.my, kxn, kyn, a.kx0, a.ky0, a.pol, R);
    const double ph = ((kxn + a.kx0) * (kxn + a.kx0) + (kyn + a.ky0) * (kyn + a.ky0)) / 2.0 / a.k * a.z0 +
                      p2 / 2.0 / a.k * a.z;
    double sp, cp;
    sincos(ph, &sp, &cp);
    // E_c = i k R(c, pol) A, times the phase
    const double2 ikA = make_double2(-a.k * A.y, a.k * A.x);
    const double2 t = cmul(ikA, make_double2(cp, sp));
    const int64_t base = (int64_t)s * 3 * nl * nm + (int64_t)(lin + a.lpx) * nm + (min_ + a.lpy);
    for (int c = 0; c < 3; ++c) F[base + (int64_t)c * nl * nm] = cscale(t, R[c]);
  }
}

// G[s][c][l][y] = sum_m F[s][c][l][m] exp(+2 pi i (m mod ny) y / ny)
__global__ void abbe_ypass_kernel(int n_sc, int nl, int nm, int lpy, int ny, const double2* F, const double2* twy,
                                  double2* G) {
  const int64_t total = (int64_t)n_sc * nl * ny;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int y = (int)(e % ny);
    const int64_t sl = e / ny;  // (s*3+c)*nl + l
    const double2* f = F + sl * nm;
    double2 acc = make_double2(0.0, 0.0);
    // (m y) mod ny for m = mm - lpy, stepped by y (0 <= y < ny) instead of a
    // 64-bit remainder per term: the same twiddles in the same order
    int idx = (int)((((int64_t)(-lpy) * y) % ny + ny) % ny);
    for (int mm = 0; mm < nm; ++mm) {
      acc = cadd(acc, cmul(f[mm], twy[idx]));
      idx += y;
      if (idx >= ny) idx -= ny;
    }
    G[e] = acc;
  }
}

// img[x][y] += w * sum_{s,c} | sum_l exp(2 pi i (l x mod nx)/nx) G[s][c][l][y] |^2,
// in two steps: the source/component range is cut into splits whose partial
// images P[s][x][y] are added in split order (abbe_xreduce_kernel, so the sum
// is deterministic).  The callers size P for xpass_splits splits; the MFMA
// passes below use at most that many.  abbe_xpass_kernel, on vector FMAs,
// serves the FP64 pass when the pupil is too wide for the MFMA pass's
// twiddle tile in LDS: a workgroup (one wave) takes 64 y values and 16 x
// values of one split, so each G value read serves 16 x.
constexpr int XP_X = 16, XP_Y = 64;
__global__ __launch_bounds__(XP_Y) void abbe_xpass_kernel(int n_sc, int nl, int lpx, int nx, int ny,
                                                          const double2* __restrict__ G,
                                                          const double2* __restrict__ twx, int scs,
                                                          double* __restrict__ P) {
  const int y = blockIdx.x * XP_Y + threadIdx.x;
  const int x0 = blockIdx.y * XP_X;
  const int s = blockIdx.z;
  extern __shared__ double2 sh_tw[];  // [XP_X][nl] twiddles of this block's x values
  for (int e = threadIdx.x; e < XP_X * nl; e += XP_Y) {
    const int xb = e / nl, ll = e % nl - lpx;
    const int64_t x = x0 + xb;
    sh_tw[e] = twx[(int)(((ll * x) % nx + nx) % nx)];
  }
  __syncthreads();
  if (y >= ny) return;
  double acc[XP_X];
#pragma unroll
  for (int xb = 0; xb < XP_X; ++xb) acc[xb] = 0.0;
  const int sc1 = min(n_sc, (s + 1) * scs);
  for (int sc = s * scs; sc < sc1; ++sc) {
    const double2* g = G + (int64_t)sc * nl * ny + y;
    double2 f[XP_X];
#pragma unroll
    for (int xb = 0; xb < XP_X; ++xb) f[xb] = make_double2(0.0, 0.0);
    for (int l = 0; l < nl; ++l) {
      const double2 gv = g[(int64_t)l * ny];
#pragma unroll
      for (int xb = 0; xb < XP_X; ++xb) f[xb] = cadd(f[xb], cmul(sh_tw[xb * nl + l], gv));
    }
#pragma unroll
    for (int xb = 0; xb < XP_X; ++xb) acc[xb] += f[xb].x * f[xb].x + f[xb].y * f[xb].y;
  }
#pragma unroll
  for (int xb = 0; xb < XP_X; ++xb)
    if (x0 + xb < nx) P[((int64_t)s * nx + x0 + xb) * ny + y] = acc[xb];
}

__global__ void abbe_xreduce_kernel(int S, int64_t per, const double* __restrict__ P, double w,
                                    double* __restrict__ img) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < per; e += (int64_t)gridDim.x * blockDim.x) {
    double acc = P[e];
    for (int s = 1; s < S; ++s) acc += P[(int64_t)s * per + e];
    img[e] += w * acc;
  }
}

// ---- FP32 imaging (EUVL_PREC_FP32_IMAGING) ----
typedef float float4_t __attribute__((ext_vector_type(4)));

__global__ void to_float2_kernel(int64_t n, const double2* in, float2* out) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    out[e] = make_float2((float)in[e].x, (float)in[e].y);
}

// G32[s][c][l][y] = sum_m F[s][c][l][m] exp(+2 pi i (m mod ny) y / ny), in FP32
// (fields rounded to FP32 as they are read; the pupil fields themselves are
// formed in FP64, whose propagation phases need it)
__global__ void abbe_ypass32_kernel(int n_sc, int nl, int nm, int lpy, int ny, const double2* F, const float2* twy,
                                    float2* G) {
  const int64_t total = (int64_t)n_sc * nl * ny;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int y = (int)(e % ny);
    const int64_t sl = e / ny;
    const double2* f = F + sl * nm;
    float re = 0.0f, im = 0.0f;
    int idx = (int)((((int64_t)(-lpy) * y) % ny + ny) % ny);  // (m y) mod ny, stepped by y
    for (int mm = 0; mm < nm; ++mm) {
      const float2 t = twy[idx];
      idx += y;
      if (idx >= ny) idx -= ny;
      const float fr = (float)f[mm].x, fi = (float)f[mm].y;
      re = fmaf(fr, t.x, fmaf(-fi, t.y, re));
      im = fmaf(fr, t.y, fmaf(fi, t.x, im));
    }
    G[e] = make_float2(re, im);
  }
}

// The x pass on FP32 MFMAs: per source/component sc the field of a 16 x 16
// pixel tile is the complex product tw_x (16 x nl) G[sc] (nl x 16) on
// v_mfma_f32_16x16x4_f32 (4 real products per k-step, FP32 accumulation over
// l); |E|^2 is added over the sc of the split with Kahan compensation in FP32,
// and the split's sum goes out in FP64 (P[s][x][y], then summed in split order
// by abbe_xreduce_kernel).  A workgroup of 4 waves takes 32 x by 32 y; lane l
// feeds A = tw(x0 + l%16, k0 + l/16), B = G(k0 + l/16, y0 + l%16), and holds
// D(x0 + 4 (l/16) + r, y0 + l%16).
constexpr int XM_T = 32;
__global__ __launch_bounds__(256) void abbe_xpass_mfma_kernel(int n_sc, int nl, int lpx, int nx, int ny,
                                                              const float2* __restrict__ G,
                                                              const float2* __restrict__ twx, int scs,
                                                              double* __restrict__ P) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const int xw = blockIdx.y * XM_T + (w >> 1) * 16, yw = blockIdx.x * XM_T + (w & 1) * 16;
  const int s = blockIdx.z;
  const int nlp = (nl + 3) & ~3;
  extern __shared__ float2 sh_twf[];  // [XM_T][nlp] twiddles of this block's x values (zero past nl)
  for (int e = threadIdx.x; e < XM_T * nlp; e += blockDim.x) {
    const int xb = e / nlp, kk = e % nlp;
    const int64_t x = blockIdx.y * XM_T + xb;
    float2 t = make_float2(0.0f, 0.0f);
    if (kk < nl) t = twx[(int)((((int64_t)(kk - lpx) * x) % nx + nx) % nx)];
    sh_twf[e] = t;
  }
  __syncthreads();
  const int y = yw + li;
  const bool yin = y < ny;
  float4_t acc = {0.0f, 0.0f, 0.0f, 0.0f}, cmp = {0.0f, 0.0f, 0.0f, 0.0f};
  const int sc1 = min(n_sc, (s + 1) * scs);
  const float2* tw = sh_twf + (int64_t)((w >> 1) * 16 + li) * nlp;
  for (int sc = s * scs; sc < sc1; ++sc) {
    const float2* g = G + (int64_t)sc * nl * ny;
    float4_t dre = {0.0f, 0.0f, 0.0f, 0.0f}, dim = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < nlp; k0 += 4) {
      const int k = k0 + lk;
      const float2 a = tw[k];
      const float2 b = (k < nl && yin) ? g[(int64_t)k * ny + y] : make_float2(0.0f, 0.0f);
      dre = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.x, dre, 0, 0, 0);
      dre = __builtin_amdgcn_mfma_f32_16x16x4f32(-a.y, b.y, dre, 0, 0, 0);
      dim = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.y, dim, 0, 0, 0);
      dim = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b.x, dim, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float v = dre[r] * dre[r] + dim[r] * dim[r];
      const float yk = v - cmp[r];
      const float t = acc[r] + yk;
      cmp[r] = (t - acc[r]) - yk;
      acc[r] = t;
    }
  }
  if (!yin) return;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int x = xw + 4 * lk + r;
    if (x < nx) P[((int64_t)s * nx + x) * ny + y] = (double)acc[r] - (double)cmp[r];
  }
}

// The x pass in FP64 on v_mfma_f64_16x16x4_f64: abbe_xpass_mfma_kernel with
// FP64 operands and accumulation (no compensation).  Lane l feeds
// A = tw(x0 + l%16, k0 + l/16), B = G(k0 + l/16, y0 + l%16) as there, and its
// FP64 accumulator holds D(x0 + 4 r + l/16, y0 + l%16).
__global__ __launch_bounds__(256) void abbe_xpass_mfma64_kernel(int n_sc, int nl, int lpx, int nx, int ny,
                                                                const double2* __restrict__ G,
                                                                const double2* __restrict__ twx, int scs,
                                                                double* __restrict__ P) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const int xw = blockIdx.y * XM_T + (w >> 1) * 16, yw = blockIdx.x * XM_T + (w & 1) * 16;
  const int s = blockIdx.z;
  const int nlp = (nl + 3) & ~3;
  extern __shared__ double2 sh_twd[];  // [XM_T][nlp] twiddles of this block's x values (zero past nl)
  for (int e = threadIdx.x; e < XM_T * nlp; e += blockDim.x) {
    const int xb = e / nlp, kk = e % nlp;
    const int64_t x = blockIdx.y * XM_T + xb;
    double2 t = make_double2(0.0, 0.0);
    if (kk < nl) t = twx[(int)((((int64_t)(kk - lpx) * x) % nx + nx) % nx)];
    sh_twd[e] = t;
  }
  __syncthreads();
  const int y = yw + li;
  const bool yin = y < ny;
  double4_t acc = {0.0, 0.0, 0.0, 0.0};
  const int sc1 = min(n_sc, (s + 1) * scs);
  const double2* tw = sh_twd + (int64_t)((w >> 1) * 16 + li) * nlp;
  for (int sc = s * scs; sc < sc1; ++sc) {
    const double2* g = G + (int64_t)sc * nl * ny;
    double4_t dre = {0.0, 0.0, 0.0, 0.0}, dim = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < nlp; k0 += 4) {
      const int k = k0 + lk;
      const double2 a = tw[k];
      const double2 b = (k < nl && yin) ? g[(int64_t)k * ny + y] : make_double2(0.0, 0.0);
      dre = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, dre, 0, 0, 0);
      dre = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, b.y, dre, 0, 0, 0);
      dim = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.y, dim, 0, 0, 0);
      dim = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.x, dim, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] += dre[r] * dre[r] + dim[r] * dim[r];
  }
  if (!yin) return;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int x = xw + 4 * r + lk;
    if (x < nx) P[((int64_t)s * nx + x) * ny + y] = acc[r];
  }
}

// splits of the source/component range for P's size: about 4096 tiles of
// XP_X x XP_Y pixels x splits
int xpass_splits(int n_sc, int nx, int ny) {
  const int64_t blocks = (int64_t)((ny + XP_Y - 1) / XP_Y) * ((nx + XP_X - 1) / XP_X);
  const int64_t want = (4096 + blocks - 1) / blocks;
  return (int)std::max<int64_t>(1, std::min<int64_t>(want, std::min(n_sc, 64)));
}

int xpass(euvl_handle h, int n_sc, int nl, int lpx, int nx, int ny, const double2* G, const double2* twx, double w,
          double* img, double* P) {
  // about 1024 workgroups of 4 waves, and never more splits than P holds
  const int64_t blocks = (int64_t)((ny + XM_T - 1) / XM_T) * ((nx + XM_T - 1) / XM_T);
  const int S0 = (int)std::max<int64_t>(
      1, std::min<int64_t>({(1024 + blocks - 1) / blocks, (int64_t)std::min(n_sc, 64),
                            (int64_t)xpass_splits(n_sc, nx, ny)}));
  const int nlp = (nl + 3) & ~3;
  int S;
  if (sizeof(double2) * XM_T * nlp <= 48 * 1024) {
    const int scs = (n_sc + S0 - 1) / S0;
    S = (n_sc + scs - 1) / scs;
    hipLaunchKernelGGL(abbe_xpass_mfma64_kernel, dim3((ny + XM_T - 1) / XM_T, (nx + XM_T - 1) / XM_T, S),
                       dim3(256), sizeof(double2) * XM_T * nlp, h->stream, n_sc, nl, lpx, nx, ny, G, twx, scs, P);
  } else {  // a wide pupil: the vector pass (16 x values per workgroup: half the LDS per l)
    const int S1 = xpass_splits(n_sc, nx, ny);
    const int scs = (n_sc + S1 - 1) / S1;
    S = (n_sc + scs - 1) / scs;
    hipLaunchKernelGGL(abbe_xpass_kernel, dim3((ny + XP_Y - 1) / XP_Y, (nx + XP_X - 1) / XP_X, S), dim3(XP_Y),
                       sizeof(double2) * XP_X * nl, h->stream, n_sc, nl, lpx, nx, ny, G, twx, scs, P);
  }
  EUVL_LAUNCH_CHECK();
  hipLaunchKernelGGL(abbe_xreduce_kernel, grid1((int64_t)nx * ny), dim3(256), 0, h->stream, S, (int64_t)nx * ny, P,
                     w, img);
  EUVL_LAUNCH_CHECK();
  return EUVL_OK;
}

// M3D fit of one clip, one thread per pupil point (m3d.cpp:329-467).

__device__ inline bool in_na_d(const M3dSetup& a, int l, int m) {
  const double u = l * a.mx / a.dx, v = m * a.my / a.dy, c = a.na / a.lambda;
  return u * u + v * v <= c * c;
}

int xpass32(euvl_handle h, int n_sc, int nl, int lpx, int nx, int ny, const float2* G, const float2* twx, double w,
            double* img, double* P) {
  const int64_t blocks = (int64_t)((ny + XM_T - 1) / XM_T) * ((nx + XM_T - 1) / XM_T);
  const int S0 = (int)std::max<int64_t>(1, std::min<int64_t>((1024 + blocks - 1) / blocks, std::min(n_sc, 64)));
  const int scs = (n_sc + S0 - 1) / S0, S = (n_sc + scs - 1) / scs;
  const int nlp = (nl + 3) & ~3;
  hipLaunchKernelGGL(abbe_xpass_mfma_kernel, dim3((ny + XM_T - 1) / XM_T, (nx + XM_T - 1) / XM_T, S), dim3(256),
                     sizeof(float2) * XM_T * nlp, h->stream, n_sc, nl, lpx, nx, ny, G, twx, scs, P);
  EUVL_LAUNCH_CHECK();
  hipLaunchKernelGGL(abbe_xreduce_kernel, grid1((int64_t)nx * ny), dim3(256), 0, h->stream, S, (int64_t)nx * ny, P,
                     w, img);
  EUVL_LAUNCH_CHECK();
  return EUVL_OK;
}

// blockIdx.y = clip: fmask, Ax and the outputs advance by their clip strides
__global__ void m3d_fit_kernel(M3dSetup a, const double2* fmask, const double2* Ax, const int* modemap,
                               const double2* vc, const double2* ab, const int* pupil, int ncut, double2* a0,
                               double2* ax, double2* ay, int64_t f_stride, int64_t a_stride) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= ncut) return;
  fmask += blockIdx.y * f_stride;
  Ax += blockIdx.y * a_stride;
  a0 += (int64_t)blockIdx.y * ncut;
  ax += (int64_t)blockIdx.y * ncut;
  ay += (int64_t)blockIdx.y * ncut;
  const int lin = pupil[3 * n], min_ = pupil[3 * n + 1], xin = pupil[3 * n + 2];
  const int noy = 2 * a.lpmy + 1;
  const int cx = a.lsmx, cy = a.lsmy;
  // centre thin-mask amplitude (m3d.cpp:329-356), divided by its phase (m3d.cpp:385-393)
  const double kxp = 2.0 * PI * lin / a.dx, kyp = 2.0 * PI * min_ / a.dy;
  const double ph = -(a.kx0 * kxp + kxp * kxp / 2.0 + a.ky0 * kyp + kyp * kyp / 2.0) / a.k * a.z0;
  double sp, cp;
  sincos(ph, &sp, &cp);
  const double2 vc0 = vc[cx * a.nsy + cy], ab0 = ab[cx * a.nsy + cy];
  double2 fc = cmul(cmul(fmask[(lin + a.lpmx) * noy + (min_ + a.lpmy)], make_double2(cp, sp)), csub(vc0, ab0));
  if (lin == 0 && min_ == 0) fc = cadd(fc, ab0);
  const double2 ph0 = cscale(vc0, 1.0 / hypot(vc0.x, vc0.y));
  fc = cdiv(fc, ph0);
  const double lp = lin, mp = min_;
  double2 f1 = make_double2(0.0, 0.0), fx = f1, fy = f1;
  double c1 = 0, sx = 0, sy = 0, sx2 = 0, sy2 = 0, sxy = 0;
  for (int is = 0; is < a.nsx; ++is)
    for (int js = 0; js < a.nsy; ++js) {
      const int ls_ = is - a.lsmx, ms_ = js - a.lsmy;
      if (!in_na_d(a, ls_, ms_) || !in_na_d(a, lin + ls_, min_ + ms_)) continue;
      const int l = lin + ls_, m = min_ + ms_;
      double2 am = make_double2(0.0, 0.0);
      if (abs(l) <= a.lmax && abs(m) <= a.mmax) {
        const int md = modemap[(l + a.lmax) * (2 * a.mmax + 1) + (m + a.mmax)];
        if (md >= 0) am = Ax[((int64_t)is * a.nsy + js) * a.nrange + md];
      }
      const double2 v = vc[is * a.nsy + js];
      const double2 phs = cscale(v, 1.0 / hypot(v.x, v.y));
      const double2 f = csub(cdiv(am, phs), fc);
      const double ls = ls_ + lp / 2.0, ms = ms_ + mp / 2.0;
      f1 = cadd(f1, f);
      fx = cadd(fx, cscale(f, ls));
      fy = cadd(fy, cscale(f, ms));
      c1 += 1;
      sx += ls;
      sy += ms;
      sx2 += ls * ls;
      sy2 += ms * ms;
      sxy += ls * ms;
    }
  const double cd = c1 * sx2 * sy2 + 2.0 * sx * sxy * sy - c1 * sxy * sxy - sx2 * sy * sy - sy2 * sx * sx;
  if (xin && cd == 0.0) {
    // Deviation: the reference divides by cd = 0 (NaN/Inf) when its snum >= 8
    // sources lie on one row or one column (the sums of half-integer positions
    // are exact, so cd is exactly 0 then).  Here the slope along the direction
    // that varies is fitted with a0 and the other slope is 0.
    const double vx = c1 * sx2 - sx * sx, vy = c1 * sy2 - sy * sy;
    ax[n] = ay[n] = make_double2(0.0, 0.0);
    if (vx > 0.0) {
      ax[n] = cscale(csub(cscale(fx, c1), cscale(f1, sx)), 1.0 / vx);
      a0[n] = cscale(csub(f1, cscale(ax[n], sx)), 1.0 / c1);
    } else if (vy > 0.0) {
      ay[n] = cscale(csub(cscale(fy, c1), cscale(f1, sy)), 1.0 / vy);
      a0[n] = cscale(csub(f1, cscale(ay[n], sy)), 1.0 / c1);
    } else {
      a0[n] = cscale(f1, 1.0 / c1);
    }
  } else if (xin) {
    const double i = 1.0 / cd;
    a0[n] = cscale(cadd(cadd(cscale(f1, sx2 * sy2 - sxy * sxy), cscale(fx, -sx * sy2 + sy * sxy)),
                        cscale(fy, sx * sxy - sy * sx2)), i);
    ax[n] = cscale(cadd(cadd(cscale(f1, -sx * sy2 + sy * sxy), cscale(fx, c1 * sy2 - sy * sy)),
                        cscale(fy, -c1 * sxy + sx * sy)), i);
    ay[n] = cscale(cadd(cadd(cscale(f1, sx * sxy - sy * sx2), cscale(fx, -c1 * sxy + sx * sy)),
                        cscale(fy, c1 * sx2 - sx * sx)), i);
  } else {
    a0[n] = cscale(f1, 1.0 / c1);
    ax[n] = make_double2(0.0, 0.0);
    ay[n] = make_double2(0.0, 0.0);
  }
}

// Pupil fields of the thin-mask and M3D-linear models on the 0.2-degree
// source mesh (linear.py:130-184 with electro_field.high_na_electro_field and
// the phase of int_abbe, linear.py:24-28).  Ft: thin mask, Fl: linear model;
// layout [s][c][l+lpx][m+lpy] as abbe_field_kernel.
struct LinearSetup {
  double2 ampvc, ampab, phase0;
  double co;
  int lpmx, lpmy;  // window of the indicator spectrum
};

__global__ void linear_field_kernel(AbbeSetup a, LinearSetup L, const double2* fmask, const double2* a0,
                                    const double2* ax, const double2* ay, const int* pupil, int ncut,
                                    const double* src, int n_src, double2* Fl, double2* Ft) {
  const int64_t total = (int64_t)n_src * ncut;
  const int nl = 2 * a.lpx + 1, nm = 2 * a.lpy + 1;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int s = (int)(e / ncut), i = (int)(e % ncut);
    const int lp = pupil[2 * i], mp = pupil[2 * i + 1];
    const double dsx = src[2 * s], dsy = src[2 * s + 1];
    const double kxn = dsx + 2.0 * PI * lp / a.dx;
    const double kyn = dsy + 2.0 * PI * mp / a.dy;
    const double p2 = a.mx * a.mx * kxn * kxn + a.my * a.my * kyn * kyn;
    const double nak = a.na * a.k;
    if (!(p2 <= nak * nak) || p2 < (nak * L.co) * (nak * L.co)) continue;  // zero outside the (annular) pupil
    // thin-mask amplitude: fmask = ampab delta + (ampvc - ampab) m^, times phaseX phaseY / phase0
    double2 fm = cmul(csub(L.ampvc, L.ampab), fmask[(int64_t)(lp + L.lpmx) * (2 * L.lpmy + 1) + (mp + L.lpmy)]);
    if (lp == 0 && mp == 0) fm = cadd(fm, L.ampab);
    const double kxp = 2.0 * PI * lp / a.dx, kyp = 2.0 * PI * mp / a.dy;
    const double ph = -(a.kx0 * kxp + kxp * kxp / 2.0) / a.k * a.z0 - (a.ky0 * kyp + kyp * kyp / 2.0) / a.k * a.z0;
    double sp, cp;
    sincos(ph, &sp, &cp);
    const double2 famp = cdiv(cmul(fm, make_double2(cp, sp)), L.phase0);
    const double kx = a.kx0 + dsx;
    const double inv = 1.0 / sqrt(a.k * a.k - kx * kx);
    const double ls = dsx / (2.0 * PI / a.dx), ms = dsy / (2.0 * PI / a.dy);
    const double2 At = cscale(famp, inv);
    double2 Al = cadd(famp, a0[i]);
    Al = cadd(Al, cscale(ax[i], ls + lp / 2.0));
    Al = cadd(Al, cscale(ay[i], ms + mp / 2.0));
    Al = cscale(Al, inv);
    double R[3];
    rotate_col(a.k, a.mx, a.my, kxn, kyn, a.kx0, a.ky0, 0, R);
    const double phz = ((kxn + a.kx0) * (kxn + a.kx0) + (kyn + a.ky0) * (kyn + a.ky0)) / (2.0 * a.k) * a.z0 +
                       p2 / (2.0 * a.k) * a.z;
    double sz, cz;
    sincos(phz, &sz, &cz);
    const double2 pz = make_double2(cz, sz);
    const double2 tt = cmul(make_double2(-a.k * At.y, a.k * At.x), pz);  // i k At e^{i phz}
    const double2 tl = cmul(make_double2(-a.k * Al.y, a.k * Al.x), pz);
    const int64_t base = (int64_t)s * 3 * nl * nm + (int64_t)(lp + a.lpx) * nm + (mp + a.lpy);
    for (int c = 0; c < 3; ++c) {
      Ft[base + (int64_t)c * nl * nm] = cscale(tt, R[c]);
      Fl[base + (int64_t)c * nl * nm] = cscale(tl, R[c]);
    }
  }
}

__global__ void zero_d_kernel(int64_t count, double* out) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (int64_t)gridDim.x * blockDim.x)
    out[e] = 0.0;
}


}  // namespace

namespace euvl {

void image_twiddles(int nx, int ny, std::vector<double2>& twx, std::vector<double2>& twy) {
  const std::complex<double> zi(0.0, 1.0);
  twx.resize(nx);
  twy.resize(ny);
  for (int i = 0; i < nx; ++i) {
    const std::complex<double> e = std::exp(2.0 * PI * zi * double(i) / double(nx));
    twx[i] = make_double2(e.real(), e.imag());
  }
  for (int i = 0; i < ny; ++i) {
    const std::complex<double> e = std::exp(2.0 * PI * zi * double(i) / double(ny));
    twy[i] = make_double2(e.real(), e.imag());
  }
}

void pupil_window(const int* pupil, int stride, int ncut, int& lpx, int& lpy) {
  lpx = lpy = 0;
  for (int i = 0; i < ncut; ++i) {
    lpx = std::max(lpx, std::abs(pupil[stride * i]));
    lpy = std::max(lpy, std::abs(pupil[stride * i + 1]));
  }
}

AbbeSetup abbe_setup(char polar, const euvl_em_params* p, double kx0, double ky0, double z0, double z, int nrange,
                     int lmax, int mmax, int lpx, int lpy) {
  AbbeSetup a{};
  a.k = p->k;
  a.dx = p->dx;
  a.dy = p->dy;
  a.na = p->na;
  a.kx0 = kx0;
  a.ky0 = ky0;
  a.z0 = z0;
  a.z = z;
  a.pol = polar == 'Y' ? 1 : 0;
  a.mx = p->mx;
  a.my = p->my;
  a.lsmx = p->lsmax_x;
  a.lsmy = p->lsmax_y;
  a.nsx = 2 * p->lsmax_x + 1;
  a.nsy = 2 * p->lsmax_y + 1;
  a.lmax = lmax;
  a.mmax = mmax;
  a.nrange = nrange;
  a.lpx = lpx;
  a.lpy = lpy;
  return a;
}

M3dSetup m3d_setup(const euvl_em_params* p, double kx0, double ky0, double z0, int lpmax_x, int lpmax_y,
                   int nrange, int lmax, int mmax) {
  M3dSetup a;
  a.k = p->k;
  a.dx = p->dx;
  a.dy = p->dy;
  a.na = p->na;
  a.lambda = p->wavelength;
  a.kx0 = kx0;
  a.ky0 = ky0;
  a.z0 = z0;
  a.mx = p->mx;
  a.my = p->my;
  a.lsmx = p->lsmax_x;
  a.lsmy = p->lsmax_y;
  a.nsx = 2 * p->lsmax_x + 1;
  a.nsy = 2 * p->lsmax_y + 1;
  a.lpmx = lpmax_x;
  a.lpmy = lpmax_y;
  a.lmax = lmax;
  a.mmax = mmax;
  a.nrange = nrange;
  return a;
}

// source chunk of the Abbe accumulation: the two spectrum buffers stay bounded (~256 MB)
int abbe_chunk(int n_src, int nl, int nm, int ny) {
  const int64_t per_src = (int64_t)3 * nl * (nm + ny);
  int chunk = (int)std::max<int64_t>(1, (int64_t)(16 << 20) / per_src);
  return std::max(1, std::min(chunk, n_src));
}

// bytes of the x pass's split partials: at most 64 splits in either precision
size_t xpass_partial_bytes(int n_sc, int nx, int ny) {
  return sizeof(double) * (size_t)std::min(n_sc, 64) * nx * ny;
}

size_t abbe_workspace_bytes(int n_src, int lpx, int lpy, int nx, int ny) {
  const int nl = 2 * lpx + 1, nm = 2 * lpy + 1;
  const int chunk = abbe_chunk(n_src, nl, nm, ny);
  const size_t fbytes = sizeof(double2) * (size_t)chunk * 3 * nl * nm;
  const size_t gbytes = sizeof(double2) * (size_t)chunk * 3 * nl * ny;  // FP64 or FP32 (and its twiddles)
  return fbytes + gbytes + xpass_partial_bytes(chunk * 3, nx, ny) + sizeof(float2) * (nx + ny) + 4 * 256;
}

// Abbe images of n_clips clips of one problem: clip c reads Ax + c ax_stride and
// accumulates weight x |E|^2 into img + c img_stride (zeroed first when `zero`).
// Takes its spectrum buffers from the workspace.
// Abbe images of n_clips clips of one problem: clip c reads Ax + c ax_stride and
// accumulates weight x |E|^2 into img + c img_stride (zeroed first when `zero`).
// With EUVL_PREC_FP32_IMAGING the transforms run in FP32 (the x pass on FP32
// MFMAs, compensated sums over source points).  Takes its spectrum buffers
// from the workspace.
int abbe_enqueue(euvl_handle h, const AbbeSetup& a, const double2* Ax, int64_t ax_stride, int n_clips,
                 const ImagingDev& t, int n_src, int ncut, int nx, int ny, double weight, double* img,
                 int64_t img_stride, bool zero) {
  const bool f32 = (h->prec_mask & EUVL_PREC_FP32_IMAGING) != 0;
  const int nl = 2 * a.lpx + 1, nm = 2 * a.lpy + 1;
  const int chunk = abbe_chunk(n_src, nl, nm, ny);
  const size_t fbytes = sizeof(double2) * (size_t)chunk * 3 * nl * nm;
  const size_t gbytes = sizeof(double2) * (size_t)chunk * 3 * nl * ny;
  double2* F = (double2*)ws_take(h, fbytes);
  double2* G = (double2*)ws_take(h, gbytes);
  double* Pxp = (double*)ws_take(h, xpass_partial_bytes(chunk * 3, nx, ny));
  float2* tw32 = f32 ? (float2*)ws_take(h, sizeof(float2) * (nx + ny)) : nullptr;
  if (!F || !G || !Pxp || (f32 && !tw32)) {
    set_error("abbe image: workspace too small");
    return EUVL_ERR_NOMEM;
  }
  if (f32) {
    hipLaunchKernelGGL(to_float2_kernel, grid1(nx), dim3(256), 0, h->stream, (int64_t)nx, t.twx, tw32);
    EUVL_LAUNCH_CHECK();
    hipLaunchKernelGGL(to_float2_kernel, grid1(ny), dim3(256), 0, h->stream, (int64_t)ny, t.twy, tw32 + nx);
    EUVL_LAUNCH_CHECK();
  }
  for (int c = 0; c < n_clips; ++c) {
    double* im = img + c * img_stride;
    if (zero) {
      hipLaunchKernelGGL(zero_d_kernel, grid1((int64_t)nx * ny), dim3(256), 0, h->stream, (int64_t)nx * ny, im);
      EUVL_LAUNCH_CHECK();
    }
    for (int s0 = 0; s0 < n_src; s0 += chunk) {
      const int ns = std::min(chunk, n_src - s0);
      EUVL_TRY(fill(h, (int64_t)ns * 3 * nl * nm, F, make_double2(0.0, 0.0)));
      hipLaunchKernelGGL(abbe_field_kernel, grid1((int64_t)ns * ncut), dim3(256), 0, h->stream, a,
                         Ax + c * ax_stride, t.map, t.off, t.src + 3 * s0, ns, t.pup2, ncut, F);
      EUVL_LAUNCH_CHECK();
      if (f32) {
        float2* G32 = (float2*)G;
        hipLaunchKernelGGL(abbe_ypass32_kernel, grid1((int64_t)ns * 3 * nl * ny), dim3(256), 0, h->stream, ns * 3,
                           nl, nm, a.lpy, ny, F, tw32 + nx, G32);
        EUVL_LAUNCH_CHECK();
        EUVL_TRY(xpass32(h, ns * 3, nl, a.lpx, nx, ny, G32, tw32, weight, im, Pxp));
      } else {
        hipLaunchKernelGGL(abbe_ypass_kernel, grid1((int64_t)ns * 3 * nl * ny), dim3(256), 0, h->stream, ns * 3,
                           nl, nm, a.lpy, ny, F, t.twy, G);
        EUVL_LAUNCH_CHECK();
        EUVL_TRY(xpass(h, ns * 3, nl, a.lpx, nx, ny, G, t.twx, weight, im, Pxp));
      }
    }
  }
  return EUVL_OK;
}

// M3D fits of n_clips clips (fmask + c f_stride: the clip's thin-mask window; Ax + c ax_stride:
// its amplitudes at the central offset); outputs [clip][ncut]
int m3d_enqueue(euvl_handle h, const M3dSetup& a, const double2* fmask, int64_t f_stride, const double2* Ax,
                int64_t ax_stride, int n_clips, const ImagingDev& t, int ncut, double2* a0, double2* ax,
                double2* ay) {
  hipLaunchKernelGGL(m3d_fit_kernel, dim3((ncut + 127) / 128, n_clips), dim3(128), 0, h->stream, a, fmask, Ax, t.map,
                     t.vc, t.ab, t.pup3, ncut, a0, ax, ay, f_stride, ax_stride);
  EUVL_LAUNCH_CHECK();
  return EUVL_OK;
}

}  // namespace euvl

extern "C" int euvl_abbe_image(euvl_handle h, char polar, const euvl_em_params* p, const double* d_Ax, int nrange,
                               int lmax,
                               int mmax, const int* h_modemap, int n_off, const double* h_off, int n_src,
                               const int* h_src, int ncut, const int* h_pupil, double kx0, double ky0, double z0,
                               double z, int nx, int ny, double weight, double* d_img) {
  using namespace euvl;
  if (!h || !p || !d_Ax || !h_modemap || n_off <= 0 || !h_off || n_src < 0 || ncut <= 0 || !h_pupil || nx <= 0 ||
      ny <= 0 || !d_img || (polar != 'X' && polar != 'Y')) {
    set_error("euvl_abbe_image: bad argument");
    return EUVL_ERR_ARG;
  }
  if (n_src == 0) return EUVL_OK;
  for (int s = 0; s < n_src; ++s)
    if (h_src[3 * s] < 0 || h_src[3 * s] >= n_off) {
      set_error("euvl_abbe_image: source %d names offset %d of %d", s, h_src[3 * s], n_off);
      return EUVL_ERR_ARG;
    }
  int lpx = 0, lpy = 0;
  pupil_window(h_pupil, 2, ncut, lpx, lpy);
  const AbbeSetup a = abbe_setup(polar, p, kx0, ky0, z0, z, nrange, lmax, mmax, lpx, lpy);
  std::vector<double2> twx, twy;
  image_twiddles(nx, ny, twx, twy);
  Upload up;
  const size_t o_map = up.add(h_modemap, sizeof(int) * (2 * lmax + 1) * (2 * mmax + 1));
  const size_t o_off = up.add(h_off, sizeof(double) * 4 * n_off);
  const size_t o_src = up.add(h_src, sizeof(int) * 3 * n_src);
  const size_t o_pup = up.add(h_pupil, sizeof(int) * 2 * ncut);
  const size_t o_twx = up.add(twx.data(), sizeof(double2) * nx);
  const size_t o_twy = up.add(twy.data(), sizeof(double2) * ny);
  EUVL_TRY(ws_reserve(h, up.host.size() + abbe_workspace_bytes(n_src, lpx, lpy, nx, ny) + 4096));
  char* blk = (char*)ws_take(h, up.host.size());
  EUVL_TRY(upload_async(h, blk, up.host.data(), up.host.size()));
  ImagingDev t;
  t.map = (const int*)(blk + o_map);
  t.off = (const double*)(blk + o_off);
  t.src = (const int*)(blk + o_src);
  t.pup2 = (const int*)(blk + o_pup);
  t.twx = (const double2*)(blk + o_twx);
  t.twy = (const double2*)(blk + o_twy);
  return abbe_enqueue(h, a, (const double2*)d_Ax, 0, 1, t, n_src, ncut, nx, ny, weight, d_img, 0, false);
}

extern "C" int euvl_m3d_fit(euvl_handle h, const euvl_em_params* p, const double* d_fmask, int lpmax_x, int lpmax_y,
                            const double* d_Ax, int nrange, int lmax, int mmax, const int* h_modemap,
                            const double* h_vc, const double* h_ab, double kx0, double ky0, double z0, int ncut,
                            const int* h_pupil, double* d_a0, double* d_ax, double* d_ay) {
  using namespace euvl;
  if (!h || !p || !d_fmask || !d_Ax || !h_modemap || !h_vc || !h_ab || ncut <= 0 || !h_pupil || !d_a0 || !d_ax ||
      !d_ay) {
    set_error("euvl_m3d_fit: bad argument");
    return EUVL_ERR_ARG;
  }
  const M3dSetup a = m3d_setup(p, kx0, ky0, z0, lpmax_x, lpmax_y, nrange, lmax, mmax);
  for (int i = 0; i < ncut; ++i)
    if (std::abs(h_pupil[3 * i]) > lpmax_x || std::abs(h_pupil[3 * i + 1]) > lpmax_y) {
      set_error("euvl_m3d_fit: pupil point %d outside the thin-mask window", i);
      return EUVL_ERR_ARG;
    }
  Upload up;
  const size_t o_map = up.add(h_modemap, sizeof(int) * (2 * lmax + 1) * (2 * mmax + 1));
  const size_t o_vc = up.add(h_vc, sizeof(double2) * a.nsx * a.nsy);
  const size_t o_ab = up.add(h_ab, sizeof(double2) * a.nsx * a.nsy);
  const size_t o_pup = up.add(h_pupil, sizeof(int) * 3 * ncut);
  EUVL_TRY(ws_reserve(h, up.host.size() + 4096));
  char* blk = (char*)ws_take(h, up.host.size());
  EUVL_TRY(upload_async(h, blk, up.host.data(), up.host.size()));
  ImagingDev t;
  t.map = (const int*)(blk + o_map);
  t.vc = (const double2*)(blk + o_vc);
  t.ab = (const double2*)(blk + o_ab);
  t.pup3 = (const int*)(blk + o_pup);
  return m3d_enqueue(h, a, (const double2*)d_fmask, 0, (const double2*)d_Ax, 0, 1, t, ncut, (double2*)d_a0,
                     (double2*)d_ax, (double2*)d_ay);
}

extern "C" int euvl_linear_images(euvl_handle h, const euvl_em_params* p, const double* d_fmask, int lpmax_x,
                                  int lpmax_y, const double* ampvc, const double* ampab, const double* phase0,
                                  const double* d_a0, const double* d_ax, const double* d_ay, int ncut,
                                  const int* h_pupil, int n_src, const double* h_src, double kx0, double ky0,
                                  double z0, double z, double co, int nx, int ny, double* d_intlinear,
                                  double* d_intft) {
  using namespace euvl;
  if (!h || !p || !d_fmask || !ampvc || !ampab || !phase0 || !d_a0 || !d_ax || !d_ay || ncut <= 0 || !h_pupil ||
      n_src <= 0 || !h_src || nx <= 0 || ny <= 0 || !d_intlinear || !d_intft) {
    set_error("euvl_linear_images: bad argument");
    return EUVL_ERR_ARG;
  }
  int lpx = 0, lpy = 0;
  for (int i = 0; i < ncut; ++i) {
    lpx = std::max(lpx, std::abs(h_pupil[2 * i]));
    lpy = std::max(lpy, std::abs(h_pupil[2 * i + 1]));
  }
  if (lpx > lpmax_x || lpy > lpmax_y) {
    set_error("euvl_linear_images: pupil outside the thin-mask window");
    return EUVL_ERR_ARG;
  }
  const int nl = 2 * lpx + 1, nm = 2 * lpy + 1;
  AbbeSetup a{};
  a.k = p->k;
  a.dx = p->dx;
  a.dy = p->dy;
  a.na = p->na;
  a.kx0 = kx0;
  a.ky0 = ky0;
  a.z0 = z0;
  a.z = z;
  a.mx = p->mx;
  a.my = p->my;
  a.lpx = lpx;
  a.lpy = lpy;
  LinearSetup L;
  L.ampvc = make_double2(ampvc[0], ampvc[1]);
  L.ampab = make_double2(ampab[0], ampab[1]);
  L.phase0 = make_double2(phase0[0], phase0[1]);
  L.co = co;
  L.lpmx = lpmax_x;
  L.lpmy = lpmax_y;
  std::vector<std::complex<double>> twx(nx), twy(ny);
  const std::complex<double> zi(0.0, 1.0);
  for (int i = 0; i < nx; ++i) twx[i] = std::exp(2.0 * PI * zi * double(i) / double(nx));
  for (int i = 0; i < ny; ++i) twy[i] = std::exp(2.0 * PI * zi * double(i) / double(ny));
  Upload up;
  const size_t o_src = up.add(h_src, sizeof(double) * 2 * n_src);
  const size_t o_pup = up.add(h_pupil, sizeof(int) * 2 * ncut);
  const size_t o_twx = up.add(twx.data(), sizeof(double2) * nx);
  const size_t o_twy = up.add(twy.data(), sizeof(double2) * ny);
  const int64_t per_src = (int64_t)3 * nl * (2 * nm + ny);
  int chunk = (int)std::max<int64_t>(1, (int64_t)(16 << 20) / per_src);
  if (chunk > n_src) chunk = n_src;
  const size_t fbytes = sizeof(double2) * (size_t)chunk * 3 * nl * nm;
  const size_t gbytes = sizeof(double2) * (size_t)chunk * 3 * nl * ny;
  const size_t pbytes = sizeof(double) * (size_t)xpass_splits(chunk * 3, nx, ny) * nx * ny;
  EUVL_TRY(ws_reserve(h, up.host.size() + 2 * fbytes + gbytes + pbytes + 4096));
  char* blk = (char*)ws_take(h, up.host.size());
  double2* Fl = (double2*)ws_take(h, fbytes);
  double2* Ft = (double2*)ws_take(h, fbytes);
  double2* G = (double2*)ws_take(h, gbytes);
  double* Pxp = (double*)ws_take(h, pbytes);
  EUVL_TRY(upload_async(h, blk, up.host.data(), up.host.size()));
  hipLaunchKernelGGL(zero_d_kernel, grid1((int64_t)nx * ny), dim3(256), 0, h->stream, (int64_t)nx * ny, d_intlinear);
  EUVL_LAUNCH_CHECK();
  hipLaunchKernelGGL(zero_d_kernel, grid1((int64_t)nx * ny), dim3(256), 0, h->stream, (int64_t)nx * ny, d_intft);
  EUVL_LAUNCH_CHECK();
  const double w = 1.0 / n_src;
  for (int s0 = 0; s0 < n_src; s0 += chunk) {
    const int ns = std::min(chunk, n_src - s0);
    EUVL_TRY(fill(h, (int64_t)ns * 3 * nl * nm, Fl, make_double2(0.0, 0.0)));
    EUVL_TRY(fill(h, (int64_t)ns * 3 * nl * nm, Ft, make_double2(0.0, 0.0)));
    hipLaunchKernelGGL(linear_field_kernel, grid1((int64_t)ns * ncut), dim3(256), 0, h->stream, a, L,
                       (const double2*)d_fmask, (const double2*)d_a0, (const double2*)d_ax, (const double2*)d_ay,
                       (const int*)(blk + o_pup), ncut, (const double*)(blk + o_src) + 2 * s0, ns, Fl, Ft);
    EUVL_LAUNCH_CHECK();
    for (int which = 0; which < 2; ++which) {
      hipLaunchKernelGGL(abbe_ypass_kernel, grid1((int64_t)ns * 3 * nl * ny), dim3(256), 0, h->stream, ns * 3, nl, nm,
                         lpy, ny, which == 0 ? Fl : Ft, (const double2*)(blk + o_twy), G);
      EUVL_LAUNCH_CHECK();
      EUVL_TRY(xpass(h, ns * 3, nl, lpx, nx, ny, G, (const double2*)(blk + o_twx), w,
                     which == 0 ? d_intlinear : d_intft, Pxp));
    }
  }
  return EUVL_OK;
}
