// Mask producers: band layouts drawn on the host, rasterised into packed
// mask.bin bits on the device.
//
// A layout is the AND of up to EUVL_MASK_MAX_LAYERS layers, optionally
// inverted.  A layer cuts one axis of the clip into bands: every position
// along that axis selects a pattern over the other axis (a 0/1 byte per
// pixel in a pattern pool), or a constant 1 (-1) or 0 (-2).  The line/space
// generators of the reference are all of this form:
//   cnn/intensity/mask/mVBF.h:1-70 (maskgen / randmask): one layer of column
//     bands of width cd, vacuum or a line/space pattern along y;
//   elitho/mask_pattern.py:18-94 (LinePattern): the same, either direction;
//   elitho/mask_pattern.py:97-196 (RandomLineSpacePattern): the AND of row
//     bands and column bands of random widths.
// The draws are a few hundred random numbers per clip; the raster is one
// thread per output byte (8 pixels of one row), written MSB-first at bit
// index ndivx*j + i as mask.cpp:9-26 packs them.
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.h"
#include "optics.h"

namespace {

// The additive-feedback generator behind the C library's rand() (glibc
// TYPE_3: degree 31, separation 3), from its published description:
//   r_0 = seed (0 -> 1); r_i = 16807 r_{i-1} mod (2^31 - 1) for 1 <= i < 31
//   (Schrage's decomposition, as srandom does); r_i = r_{i-31} for
//   31 <= i < 34; r_i = r_{i-31} + r_{i-3} (mod 2^32) for i >= 34; the k-th
//   output is r_{k+344} >> 1.
// mask.cpp never calls srand, so its masks come from seed 1.
class CRand {
 public:
  explicit CRand(unsigned seed) {
    int32_t r[344];
    int32_t w = (int32_t)seed;
    if (w == 0) w = 1;
    r[0] = w;
    for (int i = 1; i < 31; ++i) {
      const int32_t hi = r[i - 1] / 127773, lo = r[i - 1] % 127773;
      int32_t word = 16807 * lo - 2836 * hi;
      if (word < 0) word += 2147483647;
      r[i] = word;
    }
    for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
    for (int i = 34; i < 344; ++i) r[i] = (int32_t)((uint32_t)r[i - 31] + (uint32_t)r[i - 3]);
    for (int i = 0; i < 31; ++i) ring_[i] = (uint32_t)r[313 + i];
    pos_ = 0;  // ring_[pos_] holds r_{n-31} for the next index n
  }
  int next() {
    const uint32_t v = ring_[pos_] + ring_[(pos_ + 28) % 31];  // r_{n-31} + r_{n-3}
    ring_[pos_] = v;
    pos_ = (pos_ + 1) % 31;
    return (int)(v >> 1);
  }

 private:
  uint32_t ring_[31];
  int pos_;
};

// One clip of the reference's maskgen (mVBF.h:2-39) with randmask
// (mVBF.h:41-70), as a one-layer layout: sel[i] for every column, patterns
// over y appended to pool.
void draw_linespace(CRand& rng, int nx, int ny, int cd, int gap, int* sel, std::vector<uint8_t>& pool,
                    int64_t pool0) {
  auto randmask = [&]() -> int {  // a new pattern over y; returns its index
    const int64_t base = (int64_t)pool.size();
    pool.resize(base + ny);
    uint8_t* m = pool.data() + base;
    int line = rng.next() % 2;
    int sum = 0;
    while (sum < ny - gap) {
      int a = gap * (rng.next() % 5);
      if (line == 1) a += 3 * gap;
      for (int i = sum; i < std::min(sum + a, ny); ++i) m[i] = (line == 1) ? 0 : 1;
      line = (line == 1) ? 0 : 1;
      sum += a;
    }
    for (int i = sum; i < ny; ++i) m[i] = (line == 0) ? 0 : 1;
    return (int)((base - pool0) / ny);
  };
  int space = rng.next() % 2;
  int sum = 0, last = -1;
  while (sum < nx - cd) {
    if (space == 1) {
      for (int i = sum; i < std::min(sum + cd, nx); ++i) sel[i] = -1;
      space = 0;
    } else {
      last = randmask();
      for (int i = sum; i < std::min(sum + cd, nx); ++i) sel[i] = last;
      space = 1;
    }
    sum += cd;
  }
  for (int i = sum; i < nx; ++i) sel[i] = (space == 0) ? -2 : last;
}

// Device form of a layout, per layer L: axis[L]; for a column layer (axis 0)
// its patterns as a bit table [y][cb[L] bytes] (bit pi % 8 of byte pi / 8 is
// pattern pi at that y: the 64 lanes of a wave, one row of pixels, read a few
// bytes instead of one cache line per band); for a row layer (axis 1) its
// patterns as bytes [pattern][x] and its row selections sel[soff[L] + j].
// Every pattern block starts on a 16-byte boundary.
struct DevLayout {
  const int* axis;
  const int* cb;
  const int64_t* poff;
  const int64_t* soff;
  const int* sel;
  const uint8_t* pool;
};

__device__ inline bool column_bit(const uint8_t* p, int j, int cb, int pi) {
  return (p[(int64_t)j * cb + (pi >> 3)] >> (pi & 7)) & 1;
}

// One byte (8 pixels of one row) per thread: any ndivx that is a multiple of
// 8.  Column layers read their selections from sel[soff[L] + i].
__global__ void mask_raster_kernel(int nx, int ny, int n_layers, DevLayout d, int invert, uint8_t* __restrict__ bits) {
  const int bpr = nx / 8;
  const int per = bpr * ny;
  const int m = blockIdx.y;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < per; e += gridDim.x * blockDim.x) {
    const int j = e / bpr, i0 = (e - j * bpr) * 8;
    unsigned v = 0xffu;
    for (int l = 0; l < n_layers; ++l) {
      const int L = m * n_layers + l;
      const int* s = d.sel + d.soff[L];
      const uint8_t* p = d.pool + d.poff[L];
      unsigned b = 0;
      if (d.axis[L] == 0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int pi = s[i0 + q];
          const bool bit = (pi == -1) || (pi >= 0 && column_bit(p, j, d.cb[L], pi));
          b |= (unsigned)bit << (7 - q);
        }
      } else {
        const int pj = s[j];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const bool bit = (pj == -1) || (pj >= 0 && p[(int64_t)pj * nx + i0 + q] != 0);
          b |= (unsigned)bit << (7 - q);
        }
      }
      v &= b;
    }
    if (invert) v ^= 0xffu;
    bits[(int64_t)m * per + e] = (uint8_t)v;
  }
}

// 32 pixels per thread and row, for WR rows of one word column (ndivx a
// multiple of 32): one 32-bit store per row (little-endian: byte k holds
// pixels i0 + 8k .. i0 + 8k + 7, MSB first); the lanes of a wave take
// consecutive word columns, so each row's stores are contiguous.  A column
// layer arrives as runs: for word column wc, entries runs[wo[wo_off[L] + wc]
// .. wo[wo_off[L] + wc + 1]) give (selection, 32-bit mask of the word's pixels
// that take it), loaded once for the WR rows, so a word inside one band costs
// one table byte per row.  A row layer reads its pattern's 32 bytes as two
// 16-byte loads.
constexpr int WR = 8;

__global__ void mask_raster_words_kernel(int nx, int ny, int n_layers, DevLayout d, const int* __restrict__ wo,
                                         const int64_t* __restrict__ wo_off, const int2* __restrict__ runs,
                                         int invert, uint32_t* __restrict__ words) {
  const int wpr = nx / 32;
  const int groups = (ny + WR - 1) / WR;
  const int per = wpr * groups;
  const int m = blockIdx.y;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < per; t += gridDim.x * blockDim.x) {
    const int g = t / wpr, wc = t - g * wpr, i0 = wc * 32, j0 = g * WR;
    uint32_t v[WR];
#pragma unroll
    for (int r = 0; r < WR; ++r) v[r] = 0xffffffffu;  // bit 31 - q = pixel i0 + q
    for (int l = 0; l < n_layers; ++l) {
      const int L = m * n_layers + l;
      const uint8_t* p = d.pool + d.poff[L];
      uint32_t b[WR];
#pragma unroll
      for (int r = 0; r < WR; ++r) b[r] = 0;
      if (d.axis[L] == 0) {
        const int* o = wo + wo_off[L];
        const int cb = d.cb[L];
        for (int k = o[wc]; k < o[wc + 1]; ++k) {
          const int2 e = runs[k];
          if (e.x == -1) {
#pragma unroll
            for (int r = 0; r < WR; ++r) b[r] |= (uint32_t)e.y;
          } else if (e.x >= 0) {
#pragma unroll
            for (int r = 0; r < WR; ++r) {
              const int j = min(j0 + r, ny - 1);
              if (column_bit(p, j, cb, e.x)) b[r] |= (uint32_t)e.y;
            }
          }
        }
      } else {
#pragma unroll
        for (int r = 0; r < WR; ++r) {
          const int j = min(j0 + r, ny - 1);
          const int pj = d.sel[d.soff[L] + j];
          if (pj == -1) {
            b[r] = 0xffffffffu;
          } else if (pj >= 0) {
            const uint4* r4 = reinterpret_cast<const uint4*>(p + (int64_t)pj * nx + i0);
#pragma unroll
            for (int c = 0; c < 2; ++c) {
              const uint4 q4 = r4[c];
              const uint32_t rs[4] = {q4.x, q4.y, q4.z, q4.w};
#pragma unroll
              for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int k = 0; k < 4; ++k)
                  b[r] |= (uint32_t)(((rs[u] >> (8 * k)) & 0xffu) != 0) << (31 - (16 * c + 4 * u + k));
            }
          }
        }
      }
#pragma unroll
      for (int r = 0; r < WR; ++r) v[r] &= b[r];
    }
    uint32_t* out = words + (int64_t)m * wpr * ny;
#pragma unroll
    for (int r = 0; r < WR; ++r) {
      if (j0 + r < ny) {
        const uint32_t x = invert ? ~v[r] : v[r];
        // pixel order (bit 31 = i0) -> MSB-first bytes in memory order
        out[(int64_t)(j0 + r) * wpr + wc] = __builtin_bswap32(x);
      }
    }
  }
}

// Checks a layout on the host (every pattern index inside the pool), builds
// its device form, uploads it and launches the raster.
int raster(euvl_handle h, int nx, int ny, int n_masks, int n_layers, const int* h_axis, const int* h_sel,
           const int64_t* h_pat_off, const uint8_t* h_pool, int64_t pool_bytes, int invert, uint8_t* d_bits) {
  const int stride = std::max(nx, ny);
  const int nl = n_masks * n_layers;
  if ((int64_t)nx * ny / 8 >= ((int64_t)1 << 31)) {
    euvl::set_error("euvl_mask_raster: clip of %d x %d pixels is too large", nx, ny);
    return EUVL_ERR_ARG;
  }
  std::vector<int> npat(nl, 0);
  for (int L = 0; L < nl; ++L) {
    if (h_axis[L] != 0 && h_axis[L] != 1) {
      euvl::set_error("euvl_mask_raster: layer %d: axis must be 0 or 1", L);
      return EUVL_ERR_ARG;
    }
    const int npos = h_axis[L] == 0 ? nx : ny, plen = h_axis[L] == 0 ? ny : nx;
    if (h_pat_off[L] < 0 || h_pat_off[L] > pool_bytes) {
      euvl::set_error("euvl_mask_raster: layer %d: pattern offset outside the pool", L);
      return EUVL_ERR_ARG;
    }
    const int64_t avail = (pool_bytes - h_pat_off[L]) / plen;
    const int* s = h_sel + (int64_t)L * stride;
    for (int q = 0; q < npos; ++q) {
      if (s[q] < -2 || s[q] >= avail) {
        euvl::set_error("euvl_mask_raster: layer %d position %d selects pattern %d (pool holds %lld)", L, q, s[q],
                        (long long)avail);
        return EUVL_ERR_ARG;
      }
      npat[L] = std::max(npat[L], s[q] + 1);
    }
  }
  const bool words = nx % 32 == 0 && ((uintptr_t)d_bits % 4) == 0;
  std::vector<int> cb(nl, 0), dsel;
  std::vector<int64_t> poff(nl, 0), soff(nl, 0), wo_off(nl, 0);
  std::vector<uint8_t> dpool;
  std::vector<int> wo;
  std::vector<int2> runs;
  for (int L = 0; L < nl; ++L) {
    const int* s = h_sel + (int64_t)L * stride;
    const uint8_t* src = h_pool + h_pat_off[L];
    poff[L] = (int64_t)((dpool.size() + 15) & ~size_t(15));
    if (h_axis[L] == 0) {
      cb[L] = (npat[L] + 7) / 8;
      dpool.resize(poff[L] + (size_t)ny * cb[L], 0);
      uint8_t* dst = dpool.data() + poff[L];
      for (int pi = 0; pi < npat[L]; ++pi)
        for (int y = 0; y < ny; ++y)
          if (src[(size_t)pi * ny + y]) dst[(size_t)y * cb[L] + (pi >> 3)] |= (uint8_t)(1u << (pi & 7));
    } else {
      dpool.resize(poff[L] + (size_t)npat[L] * nx);
      if (npat[L] > 0) memcpy(dpool.data() + poff[L], src, (size_t)npat[L] * nx);
    }
    // selections: every layer for the byte kernel; only row layers for the word kernel
    soff[L] = (int64_t)dsel.size();
    if (!words || h_axis[L] == 1) dsel.insert(dsel.end(), s, s + (h_axis[L] == 0 ? nx : ny));
    // runs of equal selection inside every 32-column word of a column layer
    wo_off[L] = (int64_t)wo.size();
    if (words && h_axis[L] == 0) {
      for (int wc = 0; wc < nx / 32; ++wc) {
        wo.push_back((int)runs.size());
        for (int q = 0; q < 32;) {
          int e = q;
          uint32_t mask = 0;
          while (e < 32 && s[32 * wc + e] == s[32 * wc + q]) mask |= 1u << (31 - e++);
          runs.push_back(make_int2(s[32 * wc + q], (int)mask));
          q = e;
        }
      }
      wo.push_back((int)runs.size());
    }
  }
  // the uploads below must not be empty
  if (dsel.empty()) dsel.push_back(-2);
  if (dpool.empty()) dpool.push_back(0);
  if (wo.empty()) wo.push_back(0);
  if (runs.empty()) runs.push_back(make_int2(-2, 0));
  Upload up;
  const size_t o_axis = up.add(h_axis, sizeof(int) * nl);
  const size_t o_cb = up.add(cb.data(), sizeof(int) * nl);
  const size_t o_poff = up.add(poff.data(), sizeof(int64_t) * nl);
  const size_t o_soff = up.add(soff.data(), sizeof(int64_t) * nl);
  const size_t o_wooff = up.add(wo_off.data(), sizeof(int64_t) * nl);
  const size_t o_sel = up.add(dsel.data(), sizeof(int) * dsel.size());
  const size_t o_wo = up.add(wo.data(), sizeof(int) * wo.size());
  const size_t o_runs = up.add(runs.data(), sizeof(int2) * runs.size());
  const size_t o_pool = up.add(dpool.data(), dpool.size());
  EUVL_TRY(euvl::ws_reserve(h, up.host.size() + 256));
  char* blk = (char*)euvl::ws_take(h, up.host.size());
  EUVL_HIP_CHECK(hipMemcpyAsync(blk, up.host.data(), up.host.size(), hipMemcpyHostToDevice, h->stream));
  DevLayout d{(const int*)(blk + o_axis), (const int*)(blk + o_cb), (const int64_t*)(blk + o_poff),
              (const int64_t*)(blk + o_soff), (const int*)(blk + o_sel), (const uint8_t*)(blk + o_pool)};
  const int64_t per = words ? (int64_t)(nx / 32) * ((ny + WR - 1) / WR) : (int64_t)nx * ny / 8;
  const int64_t blocks = std::min<int64_t>((per + 255) / 256, 4096);
  if (!h->raster_ev[0]) {
    EUVL_HIP_CHECK(hipEventCreate(&h->raster_ev[0]));
    EUVL_HIP_CHECK(hipEventCreate(&h->raster_ev[1]));
  }
  EUVL_HIP_CHECK(hipEventRecord(h->raster_ev[0], h->stream));
  if (words)
    hipLaunchKernelGGL(mask_raster_words_kernel, dim3((unsigned)blocks, (unsigned)n_masks), dim3(256), 0, h->stream,
                       nx, ny, n_layers, d, (const int*)(blk + o_wo), (const int64_t*)(blk + o_wooff),
                       (const int2*)(blk + o_runs), invert, (uint32_t*)d_bits);
  else
    hipLaunchKernelGGL(mask_raster_kernel, dim3((unsigned)blocks, (unsigned)n_masks), dim3(256), 0, h->stream, nx,
                       ny, n_layers, d, invert, d_bits);
  EUVL_LAUNCH_CHECK();
  EUVL_HIP_CHECK(hipEventRecord(h->raster_ev[1], h->stream));
  // the staging block is host memory read by an asynchronous copy: keep it
  // alive until the copy has run
  EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));
  return EUVL_OK;
}

int check_dims(const char* who, int nx, int ny, int n_masks) {
  if (nx <= 0 || ny <= 0 || n_masks <= 0 || nx % 8 != 0 || n_masks > 65535) {
    euvl::set_error("%s: need ndivx a positive multiple of 8, ndivy > 0 and 1 <= n_masks <= 65535", who);
    return EUVL_ERR_ARG;
  }
  return EUVL_OK;
}

int plan_linespace(unsigned seed, int nx, int ny, int cd, int gap, int n_masks, std::vector<int>& sel,
                   std::vector<int64_t>& off, std::vector<uint8_t>& pool) {
  if (cd <= 0 || gap <= 0 || nx <= cd || ny <= gap) {
    euvl::set_error("euvl_mask_linespace: need cd, gap > 0, ndivx > cd and ndivy > gap");
    return EUVL_ERR_ARG;
  }
  CRand rng(seed);
  const int stride = std::max(nx, ny);
  sel.assign((size_t)n_masks * stride, -2);
  off.resize(n_masks);
  pool.clear();
  for (int m = 0; m < n_masks; ++m) {
    off[m] = (int64_t)pool.size();
    draw_linespace(rng, nx, ny, cd, gap, sel.data() + (size_t)m * stride, pool, off[m]);
  }
  return EUVL_OK;
}

}  // namespace

extern "C" int euvl_mask_raster(euvl_handle h, int ndivx, int ndivy, int n_masks, int n_layers, const int* h_axis,
                                const int* h_sel, const int64_t* h_pat_off, const uint8_t* h_pool,
                                int64_t pool_bytes, int invert, uint8_t* d_bits) {
  if (!h || !h_axis || !h_sel || !h_pat_off || (!h_pool && pool_bytes > 0) || pool_bytes < 0 || !d_bits ||
      n_layers < 1 || n_layers > EUVL_MASK_MAX_LAYERS) {
    euvl::set_error("euvl_mask_raster: bad argument");
    return EUVL_ERR_ARG;
  }
  EUVL_TRY(check_dims("euvl_mask_raster", ndivx, ndivy, n_masks));
  return raster(h, ndivx, ndivy, n_masks, n_layers, h_axis, h_sel, h_pat_off, h_pool, pool_bytes, invert != 0,
                d_bits);
}

extern "C" int euvl_mask_linespace_plan(unsigned seed, int ndivx, int ndivy, int cd, int gap, int n_masks,
                                        int* h_sel, int64_t* h_pat_off, uint8_t* h_pool, int64_t pool_cap,
                                        int64_t* h_pool_bytes) {
  if (!h_sel || !h_pat_off || !h_pool_bytes || (!h_pool && pool_cap > 0)) {
    euvl::set_error("euvl_mask_linespace_plan: bad argument");
    return EUVL_ERR_ARG;
  }
  EUVL_TRY(check_dims("euvl_mask_linespace_plan", ndivx, ndivy, n_masks));
  std::vector<int> sel;
  std::vector<int64_t> off;
  std::vector<uint8_t> pool;
  EUVL_TRY(plan_linespace(seed, ndivx, ndivy, cd, gap, n_masks, sel, off, pool));
  *h_pool_bytes = (int64_t)pool.size();
  if ((int64_t)pool.size() > pool_cap) {
    euvl::set_error("euvl_mask_linespace_plan: pool needs %lld bytes, %lld given", (long long)pool.size(),
                    (long long)pool_cap);
    return EUVL_ERR_ARG;
  }
  memcpy(h_sel, sel.data(), sizeof(int) * sel.size());
  memcpy(h_pat_off, off.data(), sizeof(int64_t) * off.size());
  if (!pool.empty()) memcpy(h_pool, pool.data(), pool.size());
  return EUVL_OK;
}

extern "C" int euvl_mask_linespace(euvl_handle h, unsigned seed, int ndivx, int ndivy, int cd, int gap, int n_masks,
                                   uint8_t* d_bits) {
  if (!h || !d_bits) {
    euvl::set_error("euvl_mask_linespace: bad argument");
    return EUVL_ERR_ARG;
  }
  EUVL_TRY(check_dims("euvl_mask_linespace", ndivx, ndivy, n_masks));
  std::vector<int> sel;
  std::vector<int64_t> off;
  std::vector<uint8_t> pool;
  EUVL_TRY(plan_linespace(seed, ndivx, ndivy, cd, gap, n_masks, sel, off, pool));
  std::vector<int> axis(n_masks, 0);
  return raster(h, ndivx, ndivy, n_masks, 1, axis.data(), sel.data(), off.data(), pool.data(),
                (int64_t)pool.size(), 0, d_bits);
}

extern "C" int euvl_mask_raster_ms(euvl_handle h, double* ms) {
  if (!h || !ms || !h->raster_ev[0]) {
    euvl::set_error("euvl_mask_raster_ms: no raster launch recorded on this handle");
    return EUVL_ERR_ARG;
  }
  EUVL_HIP_CHECK(hipEventSynchronize(h->raster_ev[1]));
  float t = 0.0f;
  EUVL_HIP_CHECK(hipEventElapsedTime(&t, h->raster_ev[0], h->raster_ev[1]));
  *ms = t;
  return EUVL_OK;
}
