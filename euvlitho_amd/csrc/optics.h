// Helpers shared by the imaging and EM translation units: launch grids, the
// packing of small host tables into one device upload, and the polarisation
// rotation of the projection optics.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace euvl {
// small host arrays packed (256-byte aligned) into one block, copied to the
// device with a single transfer
struct Upload {
  std::vector<char> host;
  size_t add(const void* p, size_t bytes) {
    size_t off = (host.size() + 255) & ~size_t(255);
    host.resize(off + bytes);
    if (bytes) memcpy(host.data() + off, p, bytes);
    return off;
  }
};
}  // namespace euvl
using euvl::Upload;

namespace {

const double PI = 3.14159265358979323846;

// grid-stride launch of `total` threads in blocks of 256 (at most 65536 blocks)
inline dim3 grid1(int64_t total) {
  int64_t b = (total + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}



// Column `col` (0: x-polarised incidence, 1: y) of the 3x2 polarisation
// rotation R = esp es^T + emp em^T of intensity.cpp:423-465
// (electro_field.polarization_rotation, electro_field.py:5-40):
// s0 = incidence (sx0, sy0, -kz), p = pupil (px, py, 0),
// p' = (MX px, MY py, -sqrt(k^2 - MX^2 px^2 - MY^2 py^2)).
__device__ inline void rotate_col(double k, int MX, int MY, double px, double py, double sx0, double sy0, int col,
                                  double R[3]) {
  double s0[3] = {sx0, sy0, -sqrt(k * k - sx0 * sx0 - sy0 * sy0)};
  double p[3] = {px, py, 0.0};
  double pp[3] = {MX * px, MY * py, -sqrt(k * k - MX * px * MX * px - MY * py * MY * py)};
  const double ez[3] = {0.0, 0.0, 1.0};
  double es[3], esp[3], em[3], emp[3];
  auto cross = [](const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
  };
  auto norm = [](const double* a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); };
  if ((px * px + py * py) > 0.00001) {
    cross(p, s0, es);
    double n = norm(es);
    for (int i = 0; i < 3; ++i) es[i] /= n;
    cross(pp, ez, esp);
    n = norm(esp);
    for (int i = 0; i < 3; ++i) esp[i] /= n;
    double ps0[3] = {p[0] + s0[0], p[1] + s0[1], p[2] + s0[2]};
    n = norm(ps0);
    for (int i = 0; i < 3; ++i) ps0[i] /= n;
    cross(es, ps0, em);
    cross(esp, pp, emp);
    for (int i = 0; i < 3; ++i) emp[i] /= k;
  } else {
    // pupil centre: the reference's special case
    cross(ez, s0, es);
    double n = norm(es);
    for (int i = 0; i < 3; ++i) es[i] /= n;
    for (int i = 0; i < 3; ++i) esp[i] = -es[i];
    cross(es, s0, em);
    for (int i = 0; i < 3; ++i) em[i] /= k;
    const double mez[3] = {0.0, 0.0, -1.0};
    cross(esp, mez, emp);
  }
  for (int i = 0; i < 3; ++i) R[i] = esp[i] * es[col] + emp[i] * em[col];
}

}  // namespace
