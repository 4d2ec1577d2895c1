// Handle lifecycle, error reporting and the device workspace of libeuvl.
#include <algorithm>
#include <cstdarg>
#include <cstring>

#include "common.h"

namespace {
thread_local char g_err[1024] = "";
}

namespace euvl {

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int ws_reserve(euvl_handle h, size_t bytes) {
  bytes = (bytes + 255) & ~size_t(255);
  h->ws_top = 0;
  if (bytes <= h->ws_cap) return EUVL_OK;
  if (h->ws) {
    EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));
    EUVL_HIP_CHECK(hipFree(h->ws));
    h->ws = nullptr;
    h->ws_cap = 0;
  }
  if (hipMalloc(&h->ws, bytes) != hipSuccess) {
    (void)hipGetLastError();
    set_error("workspace allocation of %zu bytes failed", bytes);
    return EUVL_ERR_NOMEM;
  }
  h->ws_cap = bytes;
  return EUVL_OK;
}

void* aux_reserve(euvl_handle h, size_t bytes) {
  if (bytes <= h->aux_cap) return h->aux;
  if (h->aux) {
    if (hipStreamSynchronize(h->stream) != hipSuccess) return nullptr;
    (void)hipFree(h->aux);
    h->aux = nullptr;
    h->aux_cap = 0;
  }
  if (hipMalloc(&h->aux, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  h->aux_cap = bytes;
  return h->aux;
}

int upload_async(euvl_handle h, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return EUVL_OK;
  const int k = h->stage_next;
  h->stage_next = (k + 1) % euvl_handle_s::NSTAGE;
  if (h->stage_used[k]) EUVL_HIP_CHECK(hipEventSynchronize(h->stage_ev[k]));  // its last copy has read it
  if (!h->stage_ev[k]) EUVL_HIP_CHECK(hipEventCreateWithFlags(&h->stage_ev[k], hipEventDisableTiming));
  if (bytes > h->stage_cap[k]) {
    if (h->stage_buf[k]) EUVL_HIP_CHECK(hipHostFree(h->stage_buf[k]));
    h->stage_buf[k] = nullptr;
    h->stage_cap[k] = 0;
    const size_t cap = std::max<size_t>((bytes + 4095) & ~size_t(4095), 64 * 1024);
    if (hipHostMalloc((void**)&h->stage_buf[k], cap, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      h->stage_buf[k] = nullptr;
      set_error("upload: pinned staging buffer of %zu bytes failed", cap);
      return EUVL_ERR_NOMEM;
    }
    h->stage_cap[k] = cap;
  }
  memcpy(h->stage_buf[k], src, bytes);
  EUVL_HIP_CHECK(hipMemcpyAsync(dst, h->stage_buf[k], bytes, hipMemcpyHostToDevice, h->stream));
  EUVL_HIP_CHECK(hipEventRecord(h->stage_ev[k], h->stream));
  h->stage_used[k] = true;
  return EUVL_OK;
}

void* ws_take(euvl_handle h, size_t bytes) {
  bytes = (bytes + 255) & ~size_t(255);
  if (h->ws_top + bytes > h->ws_cap) return nullptr;
  void* p = h->ws + h->ws_top;
  h->ws_top += bytes;
  return p;
}

}  // namespace euvl

extern "C" const char* euvl_last_error(void) { return g_err; }

extern "C" int euvl_create(euvl_handle* out) {
  if (!out) return EUVL_ERR_ARG;
  euvl_handle h = new euvl_handle_s();
  hipError_t e = hipGetDevice(&h->device);
  if (e == hipSuccess) e = hipHostMalloc((void**)&h->h_pinned, 4096, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc((void**)&h->d_red, 4096);
  if (e != hipSuccess) {
    euvl::set_error("euvl_create: %s", hipGetErrorString(e));
    delete h;
    return EUVL_ERR_HIP;
  }
  *out = h;
  return EUVL_OK;
}

extern "C" int euvl_destroy(euvl_handle h) {
  if (!h) return EUVL_OK;
  (void)hipStreamSynchronize(h->stream);
  for (hipEvent_t e : h->ev_pool) (void)hipEventDestroy(e);
  for (hipEvent_t e : h->raster_ev)
    if (e) (void)hipEventDestroy(e);
  for (int k = 0; k < euvl_handle_s::NSTAGE; ++k) {
    if (h->stage_ev[k]) (void)hipEventDestroy(h->stage_ev[k]);
    if (h->stage_buf[k]) (void)hipHostFree(h->stage_buf[k]);
  }
  if (h->ws) (void)hipFree(h->ws);
  if (h->aux) (void)hipFree(h->aux);
  if (h->d_red) (void)hipFree(h->d_red);
  if (h->h_pinned) (void)hipHostFree(h->h_pinned);
  delete h;
  return EUVL_OK;
}

extern "C" int euvl_set_stream(euvl_handle h, void* s) {
  if (!h) return EUVL_ERR_ARG;
  // work enqueued on the old stream may still read the workspace that the
  // next call reuses on the new one
  if ((hipStream_t)s != h->stream) EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));
  h->stream = (hipStream_t)s;
  return EUVL_OK;
}

extern "C" int64_t euvl_workspace_bytes(euvl_handle h) { return h ? (int64_t)h->ws_cap : 0; }

extern "C" int euvl_set_precision(euvl_handle h, int mask) {
  const int all = EUVL_PREC_FP32_MATFUN | EUVL_PREC_FP32_COUPLING | EUVL_PREC_FP32_LU | EUVL_PREC_FP32_GEMM |
                  EUVL_PREC_GEMM_4M | EUVL_PREC_FP32_IMAGING;
  if (!h || (mask & ~all)) {
    euvl::set_error("euvl_set_precision: unknown bits in mask 0x%x", mask);
    return EUVL_ERR_ARG;
  }
  h->prec_mask = mask;
  return EUVL_OK;
}

extern "C" int euvl_reserve_workspace(euvl_handle h, int64_t bytes) {
  if (!h || bytes < 0) {
    euvl::set_error("euvl_reserve_workspace: bad argument");
    return EUVL_ERR_ARG;
  }
  EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));
  return euvl::ws_reserve(h, (size_t)bytes);
}

extern "C" int euvl_release_workspace(euvl_handle h) {
  if (!h) return EUVL_ERR_ARG;
  EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));
  if (h->ws) EUVL_HIP_CHECK(hipFree(h->ws));
  if (h->aux) EUVL_HIP_CHECK(hipFree(h->aux));
  h->ws = nullptr;
  h->aux = nullptr;
  h->ws_cap = h->aux_cap = 0;
  h->ws_top = 0;
  return EUVL_OK;
}
