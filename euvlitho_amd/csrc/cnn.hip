// The CNN M3D predictor (cnn/cnn/predict/predict.py:16-151): six networks of
// five conv blocks and two fully connected layers map a 512 x 512
// down-sampled mask to the M3D parameters re0, im0 (1901 pupil points) and
// rex, imx, rey, imy (1749).  FP32 throughout, as the reference runs.
//
//   compmask_kernel   mask.bin bits -> 512 x 512 block means (compmask.py:17-41)
//   conv_block_kernel conv 3x3 (circular padding) + bias + ReLU + 2x2 max
//                     pool + BatchNorm (eval), one launch per block, as an
//                     implicit GEMM on v_mfma_f32_16x16x4f32
//   fc_splitk_kernel  fully connected layer y = x W^T as a split-K GEMM on
//                     v_mfma_f32_16x16x4f32 (batch x features tiles of 64 x 64;
//                     the weights, 1 GB per fc1, are read once per batch)
//   fc_reduce_kernel  the sum of the K slices + bias (+ ReLU, / factor),
//                     in a fixed order (deterministic)
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"
#include "optics.h"

namespace {

typedef float float4_t __attribute__((ext_vector_type(4)));

// compmask: out[c][r][x] = (1/f^2) * sum of the f x f block of mask bits
// (rows r*f.., columns x*f..; f = ndiv / 512), row r of the bit array being
// mask.bin row j (no flip), i.e. rowmat * mask0 * colmat of compmask.py.
__global__ void compmask_kernel(int ndiv, int nout, int n_clips, const uint8_t* __restrict__ bits,
                                float* __restrict__ out) {
  const int f = ndiv / nout;
  const int64_t per = (int64_t)nout * nout;
  const int64_t total = per * n_clips;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e / per);
    const int r = (int)((e % per) / nout), x = (int)(e % nout);
    const uint8_t* clip = bits + (int64_t)c * ndiv * ndiv / 8;
    int cnt = 0;
    for (int dy = 0; dy < f; ++dy)
      for (int dx = 0; dx < f; ++dx) {
        const int64_t idx = (int64_t)(r * f + dy) * ndiv + x * f + dx;
        cnt += (clip[idx >> 3] >> (7 - (idx & 7))) & 1;
      }
    out[e] = (float)cnt / (float)(f * f);
  }
}

// One conv block.  A workgroup (4 waves) computes a 16 x 16 pre-pool tile of
// 16*NTU output channels for one sample (tiles past an edge compute pixels
// that are not stored; H and W are even, so a pooling window is all inside or
// all outside).  GEMM view: M = pixels, N = output
// channels, K = input channel x tap, staged CK input channels at a time (the
// tile with its circular halo and the weights) through LDS.  Rows of each
// 16-row MFMA tile are ordered so that rows 4q..4q+3 are the 2 x 2 pixels of
// pooled pixel q: the accumulator element r of lane l is row 4(l/16) + r
// (gfx950 f32 16x16x4 layout), so each lane holds one whole pooling window
// and the max pool is a max over its four accumulator elements.
template <int NTU, int CK>
__global__ __launch_bounds__(256) void conv_block_kernel(int cin, int cout, int H, int W, const float* __restrict__ X,
                                                        const float* __restrict__ Wt, const float* __restrict__ bias,
                                                        const float* __restrict__ scale,
                                                        const float* __restrict__ shift, float* __restrict__ Y) {
  constexpr int KC = CK * 9;          // K per chunk
  constexpr int STEPS = (KC + 3) / 4;  // MFMA steps of 4 (the last padded with zeros when KC % 4 != 0)
  __shared__ float xs[CK][18][19];
  __shared__ float ws[16 * NTU][KC + 1];
  const int tiles_x = (W + 15) / 16;
  const int ty0 = (blockIdx.x / tiles_x) * 16, tx0 = (blockIdx.x % tiles_x) * 16;
  const int co0 = blockIdx.y * 16 * NTU, b = blockIdx.z;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lg = lane >> 4, li = lane & 15;
  // pixel of A-operand row li in M-tile t: pooled pixel q = 4t + li/4 of the
  // wave (2 rows of 8), position li % 4 inside its 2 x 2 window
  int py[4], px[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int q = 4 * t + li / 4, pos = li % 4;
    py[t] = 4 * wave + 2 * (q / 8) + pos / 2;
    px[t] = 2 * (q % 8) + pos % 2;
  }
  float4_t acc[4][NTU];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int u = 0; u < NTU; ++u) acc[t][u] = (float4_t){0.f, 0.f, 0.f, 0.f};
  const float* Xb = X + (int64_t)b * cin * H * W;
  for (int c0 = 0; c0 < cin; c0 += CK) {
    __syncthreads();
    for (int e = tid; e < CK * 18 * 18; e += 256) {
      const int ci = e / 324, r = (e % 324) / 18, c = e % 18;
      const int gy = (ty0 + r - 1 + H) % H, gx = (tx0 + c - 1 + W) % W;
      xs[ci][r][c] = (c0 + ci < cin) ? Xb[((int64_t)(c0 + ci) * H + gy) * W + gx] : 0.f;
    }
    for (int e = tid; e < 16 * NTU * KC; e += 256) {
      const int co = e / KC, k = e % KC;
      const int ci = c0 + k / 9;
      ws[co][k] = (ci < cin && co0 + co < cout) ? Wt[((int64_t)(co0 + co) * cin + ci) * 9 + k % 9] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < STEPS; ++s) {
      const int k = 4 * s + lg;
      const bool kin = (KC % 4 == 0) || k < KC;
      const int kc = kin ? k : 0;
      const int ci = kc / 9, tap = kc % 9, dy = tap / 3, dx = tap % 3;
      float a[4], w[NTU];
#pragma unroll
      for (int t = 0; t < 4; ++t) a[t] = kin ? xs[ci][py[t] + dy][px[t] + dx] : 0.f;
#pragma unroll
      for (int u = 0; u < NTU; ++u) w[u] = kin ? ws[16 * u + li][kc] : 0.f;
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int u = 0; u < NTU; ++u) acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[t], w[u], acc[t][u], 0, 0, 0);
    }
  }
  // epilogue: lane holds the window of pooled pixel q = 4t + lg for channel 16u + li
  const int Ho = H / 2, Wo = W / 2;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int u = 0; u < NTU; ++u) {
      const int co = co0 + 16 * u + li;
      if (co >= cout) continue;
      const float m = fmaxf(fmaxf(acc[t][u][0], acc[t][u][1]), fmaxf(acc[t][u][2], acc[t][u][3]));
      const float v = fmaxf(m + bias[co], 0.f);  // max(relu(x + b)) = relu(max(x) + b)
      const int q = 4 * t + lg;
      const int oy = (ty0 + 4 * wave + 2 * (q / 8)) / 2, ox = (tx0 + 2 * (q % 8)) / 2;
      if (oy < Ho && ox < Wo) Y[(((int64_t)b * cout + co) * Ho + oy) * Wo + ox] = v * scale[co] + shift[co];
    }
}

// part[s][b][n] = sum over k in slice s of x[b][k] W[n][k]: a 64 (batch) x
// 64 (features) tile per workgroup of 4 waves (2 x 2, each 32 x 32 in 2 x 2
// MFMA tiles); K staged through LDS in slabs of 32, k-major, with a register
// prefetch of the next slab (16-byte loads where the slab is whole and the
// rows are 16-byte aligned).
__global__ __launch_bounds__(256) void fc_splitk_kernel(int B, int N, int K, int kslice, int vec,
                                                       const float* __restrict__ x, const float* __restrict__ W,
                                                       float* __restrict__ part) {
  constexpr int BK = 32;
  __shared__ float xs[BK][68];
  __shared__ float wsm[BK][68];
  const int n0 = blockIdx.x * 64, s = blockIdx.y, b0 = blockIdx.z * 64;
  const int k_begin = s * kslice, k_end = min(K, k_begin + kslice);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lg = lane >> 4, li = lane & 15;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  // slab loads: thread -> row r (0..63), 8 consecutive k at c
  const int r = tid >> 2, c = (tid & 3) * 8;
  float rx[8], rw[8];
  auto load = [&](int k0) {
    const bool whole = vec && k0 + BK <= k_end;
    const bool xr = b0 + r < B, wr_ = n0 + r < N;
    if (whole) {
#pragma unroll
      for (int h2 = 0; h2 < 2; ++h2) {
        const float4 vx = xr ? *reinterpret_cast<const float4*>(x + (int64_t)(b0 + r) * K + k0 + c + 4 * h2)
                             : make_float4(0.f, 0.f, 0.f, 0.f);
        const float4 vw = wr_ ? *reinterpret_cast<const float4*>(W + (int64_t)(n0 + r) * K + k0 + c + 4 * h2)
                              : make_float4(0.f, 0.f, 0.f, 0.f);
        rx[4 * h2] = vx.x; rx[4 * h2 + 1] = vx.y; rx[4 * h2 + 2] = vx.z; rx[4 * h2 + 3] = vx.w;
        rw[4 * h2] = vw.x; rw[4 * h2 + 1] = vw.y; rw[4 * h2 + 2] = vw.z; rw[4 * h2 + 3] = vw.w;
      }
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int k = k0 + c + q;
        rx[q] = (xr && k < k_end) ? x[(int64_t)(b0 + r) * K + k] : 0.f;
        rw[q] = (wr_ && k < k_end) ? W[(int64_t)(n0 + r) * K + k] : 0.f;
      }
    }
  };
  float4_t acc[2][2];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int u = 0; u < 2; ++u) acc[t][u] = (float4_t){0.f, 0.f, 0.f, 0.f};
  load(k_begin);
  for (int k0 = k_begin; k0 < k_end; k0 += BK) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      xs[c + q][r] = rx[q];
      wsm[c + q][r] = rw[q];
    }
    __syncthreads();
    if (k0 + BK < k_end) load(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      float a[2], bw[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) a[t] = xs[kk + lg][wr + 16 * t + li];
#pragma unroll
      for (int u = 0; u < 2; ++u) bw[u] = wsm[kk + lg][wc + 16 * u + li];
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u) acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[t], bw[u], acc[t][u], 0, 0, 0);
    }
  }
  // accumulator element q of lane l: row 4 (l / 16) + q (batch), column l % 16 (feature)
  float* P = part + (int64_t)s * B * N;
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int b = b0 + wr + 16 * t + 4 * lg + q, n = n0 + wc + 16 * u + li;
        if (b < B && n < N) P[(int64_t)b * N + n] = acc[t][u][q];
      }
}

// y[b][n] = act(sum_s part[s][b][n] + bias[n]) / div[n], slices summed in order
__global__ void fc_reduce_kernel(int B, int N, int S, const float* __restrict__ part, const float* __restrict__ bias,
                                 int relu, const float* __restrict__ div, float* __restrict__ y) {
  const int64_t total = (int64_t)B * N;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int n = (int)(e % N);
    float v = 0.f;
    for (int s = 0; s < S; ++s) v += part[(int64_t)s * total + e];
    v += bias[n];
    if (relu) v = fmaxf(v, 0.f);
    if (div) v = v / div[n];
    y[e] = v;
  }
}

}  // namespace

extern "C" int euvl_cnn_compmask(euvl_handle h, const uint8_t* d_bits, int ndiv, int n_clips, int nout,
                                 float* d_out) {
  if (!h || !d_bits || !d_out || ndiv <= 0 || nout <= 0 || n_clips <= 0 || ndiv % nout != 0 ||
      ((int64_t)ndiv * ndiv) % 8 != 0) {
    euvl::set_error("euvl_cnn_compmask: bad argument (ndiv must be a multiple of nout)");
    return EUVL_ERR_ARG;
  }
  const int64_t total = (int64_t)nout * nout * n_clips;
  hipLaunchKernelGGL(compmask_kernel, grid1(total), dim3(256), 0, h->stream, ndiv, nout, n_clips, d_bits, d_out);
  EUVL_LAUNCH_CHECK();
  return EUVL_OK;
}

extern "C" int euvl_cnn_conv_block(euvl_handle h, int batch, int cin, int cout, int height, int width,
                                   const float* d_x, const float* d_w, const float* d_bias, const float* d_scale,
                                   const float* d_shift, float* d_y) {
  if (!h || !d_x || !d_w || !d_bias || !d_scale || !d_shift || !d_y || batch <= 0 || batch > 65535 || cin <= 0 ||
      cout <= 0 || height <= 0 || width <= 0 || height % 2 != 0 || width % 2 != 0) {
    euvl::set_error("euvl_cnn_conv_block: bad argument (height and width must be even)");
    return EUVL_ERR_ARG;
  }
  const int tiles = ((height + 15) / 16) * ((width + 15) / 16);
  // one input channel (the first block): chunks of 1 channel; otherwise 8
  if (cout <= 16 && cin == 1)
    hipLaunchKernelGGL((conv_block_kernel<1, 1>), dim3(tiles, (cout + 15) / 16, batch), dim3(256), 0, h->stream, cin,
                       cout, height, width, d_x, d_w, d_bias, d_scale, d_shift, d_y);
  else if (cout <= 16)
    hipLaunchKernelGGL((conv_block_kernel<1, 8>), dim3(tiles, (cout + 15) / 16, batch), dim3(256), 0, h->stream, cin,
                       cout, height, width, d_x, d_w, d_bias, d_scale, d_shift, d_y);
  else if (cin == 1)
    hipLaunchKernelGGL((conv_block_kernel<2, 1>), dim3(tiles, (cout + 31) / 32, batch), dim3(256), 0, h->stream, cin,
                       cout, height, width, d_x, d_w, d_bias, d_scale, d_shift, d_y);
  else
    hipLaunchKernelGGL((conv_block_kernel<2, 8>), dim3(tiles, (cout + 31) / 32, batch), dim3(256), 0, h->stream, cin,
                       cout, height, width, d_x, d_w, d_bias, d_scale, d_shift, d_y);
  EUVL_LAUNCH_CHECK();
  return EUVL_OK;
}

extern "C" int euvl_cnn_linear(euvl_handle h, int batch, int in_features, int out_features, const float* d_x,
                               const float* d_w, const float* d_bias, int relu, const float* d_div, float* d_y) {
  if (!h || !d_x || !d_w || !d_bias || !d_y || batch <= 0 || batch > 65535 * 64 || in_features <= 0 ||
      out_features <= 0) {
    euvl::set_error("euvl_cnn_linear: bad argument");
    return EUVL_ERR_ARG;
  }
  // K slices: enough workgroups to fill the device (~2048), slices of at
  // least 256 and a multiple of 32 (the slab)
  const int tiles = ((out_features + 63) / 64) * ((batch + 63) / 64);
  int S = std::max(1, std::min((2048 + tiles - 1) / tiles, (in_features + 255) / 256));
  int kslice = ((in_features + S - 1) / S + 31) / 32 * 32;
  S = (in_features + kslice - 1) / kslice;
  const size_t part_bytes = sizeof(float) * (size_t)S * batch * out_features;
  EUVL_TRY(euvl::ws_reserve(h, part_bytes + 256));
  float* part = (float*)euvl::ws_take(h, part_bytes);
  // 16-byte loads when every row of x and W starts on a 16-byte boundary
  const int vec = (in_features % 4 == 0) && ((uintptr_t)d_x % 16 == 0) && ((uintptr_t)d_w % 16 == 0);
  hipLaunchKernelGGL(fc_splitk_kernel, dim3((out_features + 63) / 64, S, (batch + 63) / 64), dim3(256), 0, h->stream,
                     batch, out_features, in_features, kslice, vec, d_x, d_w, part);
  EUVL_LAUNCH_CHECK();
  hipLaunchKernelGGL(fc_reduce_kernel, grid1((int64_t)batch * out_features), dim3(256), 0, h->stream, batch,
                     out_features, S, part, d_bias, relu, d_div, d_y);
  EUVL_LAUNCH_CHECK();
  return EUVL_OK;
}
