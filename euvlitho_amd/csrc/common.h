// Shared helpers of libeuvl: complex arithmetic on double2, error plumbing,
// the handle and its device workspace.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/euvl.h"

typedef double double4_t __attribute__((ext_vector_type(4)));

__host__ __device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__host__ __device__ inline double2 cdiv(double2 a, double2 b) {
  // Smith's algorithm (scaled) for a / b
  if (fabs(b.x) >= fabs(b.y)) {
    double r = b.y / b.x, den = b.x + b.y * r;
    return make_double2((a.x + a.y * r) / den, (a.y - a.x * r) / den);
  } else {
    double r = b.x / b.y, den = b.x * r + b.y;
    return make_double2((a.x * r + a.y) / den, (a.y * r - a.x) / den);
  }
}
__device__ inline double2 cexpi(double2 z) {  // exp(z)
  double e = exp(z.x), s, c;
  sincos(z.y, &s, &c);
  return make_double2(e * c, e * s);
}
__device__ inline double2 csqrt_principal(double2 z) {
  // principal branch: Re >= 0, branch cut on the negative real axis
  double r = hypot(z.x, z.y);
  double a = sqrt(0.5 * (r + fabs(z.x)));
  if (a == 0.0) return make_double2(0.0, 0.0);
  if (z.x >= 0.0) return make_double2(a, z.y / (2.0 * a));
  return make_double2(fabs(z.y) / (2.0 * a), copysign(a, z.y));
}

namespace euvl {

void set_error(const char* fmt, ...);

#define EUVL_HIP_CHECK(expr)                                                   \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) {                                                    \
      ::euvl::set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr,            \
                        hipGetErrorString(e_));                                \
      return EUVL_ERR_HIP;                                                     \
    }                                                                          \
  } while (0)

#define EUVL_LAUNCH_CHECK() EUVL_HIP_CHECK(hipGetLastError())

#define EUVL_TRY(expr)                                                         \
  do {                                                                         \
    int s_ = (expr);                                                           \
    if (s_ != EUVL_OK) return s_;                                              \
  } while (0)

}  // namespace euvl

struct euvl_handle_s {
  int device = 0;
  hipStream_t stream = nullptr;
  // Bump allocator over one growing device buffer; reset at the start of each
  // top-level entry point that uses it.
  char* ws = nullptr;
  size_t ws_cap = 0;
  size_t ws_top = 0;
  size_t ws_peak = 0;  // largest request seen (grow target)
  // separate growable buffer for kernel-internal scratch (LU panels)
  char* aux = nullptr;
  size_t aux_cap = 0;
  // small pinned host buffer for reductions read back by the host
  double* h_pinned = nullptr;
  double* d_red = nullptr;  // device reduction slots
  // optional per-launch timing of the GEMM kernel (euvl_gemm_stats)
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  struct GemmRecord {
    int ev;        // index of the start event (the end event follows)
    double flops;  // 8 m n k batch
    double bytes;  // A, B read once, C written (and read when beta != 0)
    bool fp32;     // products on FP32 MFMA
  };
  std::vector<GemmRecord> ev_used;
  size_t ev_next = 0;
  // GEMM compute precision: a GEMM runs on FP32 MFMA when the bit of the
  // current scope is set in prec_mask (euvl_set_precision), else on FP64 MFMA
  int prec_mask = 0;
  int prec_scope = 0;
  // events around the last mask raster launch (euvl_mask_raster_ms)
  hipEvent_t raster_ev[2] = {nullptr, nullptr};
  // pinned staging buffers of the host tables that entry points upload
  // (upload_async): a copy may read its buffer after the call has returned,
  // so a buffer is reused only once the event after its last copy is done
  static constexpr int NSTAGE = 8;
  char* stage_buf[NSTAGE] = {};
  size_t stage_cap[NSTAGE] = {};
  hipEvent_t stage_ev[NSTAGE] = {};
  bool stage_used[NSTAGE] = {};
  int stage_next = 0;
};

namespace euvl {
// Marks the GEMMs issued while it lives as belonging to one precision scope
// (EUVL_PREC_* bit); restores the enclosing scope on exit.
struct PrecScope {
  euvl_handle h;
  int saved;
  PrecScope(euvl_handle h_, int scope) : h(h_), saved(h_->prec_scope) { h->prec_scope = scope; }
  ~PrecScope() { h->prec_scope = saved; }
};
}  // namespace euvl

namespace euvl {
// Ensure the workspace holds at least `bytes`; invalidates earlier pointers.
int ws_reserve(euvl_handle h, size_t bytes);
// Carve `bytes` (256-byte aligned) from the reserved workspace.
void* ws_take(euvl_handle h, size_t bytes);
// Ensure the aux buffer holds at least `bytes` and return it.
void* aux_reserve(euvl_handle h, size_t bytes);
// Copies `bytes` of host memory at src to dst (device) on the handle's stream
// through a pinned staging buffer, without waiting for the copy: src may be
// freed when this returns.  Not for code under graph capture or that must not
// allocate (a staging buffer may be allocated or grown).
int upload_async(euvl_handle h, void* dst, const void* src, size_t bytes);
}  // namespace euvl
