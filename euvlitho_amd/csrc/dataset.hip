// Training-data augmentation of the CNN's data loader
// (cnn/cnn/model/re0/re0.py:24-62, MaskampDataset.__getitem__): a random
// integer shift (lshift, mshift) of each down-sampled mask, as
// np.roll(x, (mshift, -lshift), axis=(0, 1)), and the matching phase ramp of
// its M3D amplitudes, ramp = re cos(2 pi (l lshift + m mshift) / n) - im sin,
// iamp = re sin + im cos (rshift / ishift, re0.py:24-36), with the float32
// cos/sin tables of the reference (computed in double, rounded once) and the
// float32 products and sums in the reference's order, unfused.
#include <algorithm>

#include "common.h"
#include "optics.h"

namespace {

// one sample per blockIdx.y; masks are float16 bit patterns (moved, not computed)
__global__ void roll_kernel(int n, int nout, const uint16_t* __restrict__ in, const int2* __restrict__ shifts,
                            uint16_t* __restrict__ out) {
  const int s = blockIdx.y;
  const int64_t per = (int64_t)nout * nout;
  const int ls = shifts[s].x, ms = shifts[s].y;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < per; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / nout), j = (int)(e % nout);
    const int si = ((i - ms) % nout + nout) % nout, sj = ((j + ls) % nout + nout) % nout;
    out[s * per + e] = in[s * per + (int64_t)si * nout + sj];
  }
}

__global__ void ramp_kernel(int n, int ncut, int nout, const int* __restrict__ lorder, const int* __restrict__ morder,
                            const float* __restrict__ re, const float* __restrict__ im, const int2* __restrict__ shifts,
                            const float* __restrict__ cs, const float* __restrict__ sn, float* __restrict__ out_re,
                            float* __restrict__ out_im) {
#pragma clang fp contract(off)
  const int s = blockIdx.y;
  const int ls = shifts[s].x, ms = shifts[s].y;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ncut; i += gridDim.x * blockDim.x) {
    const long long t = (long long)lorder[i] * ls + (long long)morder[i] * ms;
    const int lm = (int)(((t % nout) + nout) % nout);  // Python's non-negative %
    const float a = re[(int64_t)s * ncut + i], b = im[(int64_t)s * ncut + i];
    const float c = cs[lm], sv = sn[lm];
    const float p1 = a * c, p2 = b * sv, p3 = a * sv, p4 = b * c;
    out_re[(int64_t)s * ncut + i] = p1 - p2;
    out_im[(int64_t)s * ncut + i] = p3 + p4;
  }
}

}  // namespace

extern "C" int euvl_augment_batch(euvl_handle h, int n, int nout, const uint16_t* d_masks, const int* h_shifts,
                                  int ncut, const int* h_lorder, const int* h_morder, const float* d_re,
                                  const float* d_im, const float* h_cos, const float* h_sin, uint16_t* d_out_masks,
                                  float* d_out_re, float* d_out_im) {
  if (!h || n <= 0 || n > 65535 || nout <= 0 || ncut < 0 || !h_shifts || !h_cos || !h_sin ||
      (ncut > 0 && (!h_lorder || !h_morder || !d_re || !d_im || !d_out_re || !d_out_im)) ||
      (d_masks == nullptr) != (d_out_masks == nullptr)) {
    euvl::set_error("euvl_augment_batch: bad argument");
    return EUVL_ERR_ARG;
  }
  Upload up;
  const size_t o_sh = up.add(h_shifts, sizeof(int) * 2 * n);
  const size_t o_l = up.add(h_lorder, sizeof(int) * ncut);
  const size_t o_m = up.add(h_morder, sizeof(int) * ncut);
  const size_t o_c = up.add(h_cos, sizeof(float) * nout);
  const size_t o_s = up.add(h_sin, sizeof(float) * nout);
  EUVL_TRY(euvl::ws_reserve(h, up.host.size() + 256));
  char* blk = (char*)euvl::ws_take(h, up.host.size());
  EUVL_HIP_CHECK(hipMemcpyAsync(blk, up.host.data(), up.host.size(), hipMemcpyHostToDevice, h->stream));
  const int2* sh = (const int2*)(blk + o_sh);
  if (d_masks) {
    const int64_t per = (int64_t)nout * nout;
    hipLaunchKernelGGL(roll_kernel, dim3((unsigned)std::min<int64_t>((per + 255) / 256, 1024), n), dim3(256), 0,
                       h->stream, n, nout, d_masks, sh, d_out_masks);
    EUVL_LAUNCH_CHECK();
  }
  if (ncut > 0) {
    hipLaunchKernelGGL(ramp_kernel, dim3((ncut + 255) / 256, n), dim3(256), 0, h->stream, n, ncut, nout,
                       (const int*)(blk + o_l), (const int*)(blk + o_m), d_re, d_im, sh, (const float*)(blk + o_c),
                       (const float*)(blk + o_s), d_out_re, d_out_im);
    EUVL_LAUNCH_CHECK();
  }
  // the staging block is host memory read by an asynchronous copy
  EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));
  return EUVL_OK;
}
