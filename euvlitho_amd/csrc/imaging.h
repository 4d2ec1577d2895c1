// Internal interfaces of the image-formation stages (imaging.hip), shared
// with the batched plans (plan.hip).
#pragma once
#include <vector>

#include "common.h"

namespace euvl {

struct AbbeSetup {
  double k, dx, dy, na, kx0, ky0, z0, z;
  int pol;  // 0: X (Ax = amp, Ay = 0), 1: Y (Ay = amp, Ax = 0), intensity.cpp:295-305
  int mx, my, lsmx, lsmy, nsx, nsy, lmax, mmax, nrange;
  int lpx, lpy;  // spectrum half-widths
};

struct M3dSetup {
  double k, dx, dy, na, lambda, kx0, ky0, z0;
  int mx, my, lsmx, lsmy, nsx, nsy, lpmx, lpmy, lmax, mmax, nrange;
};

// device tables of the image stages (unused members may stay null)
struct ImagingDev {
  const int* map = nullptr;     // mode map (2 lmax + 1) x (2 mmax + 1)
  const double* off = nullptr;  // (sx0, sy0, fx, fy) per offset
  const int* src = nullptr;     // (offset, l0, m0) per source point
  const int* pup2 = nullptr;    // (l, m) per pupil point
  const int* pup3 = nullptr;    // (l, m, xinput) per pupil point
  const double2* twx = nullptr;
  const double2* twy = nullptr;
  const double2* vc = nullptr;  // blank amplitudes (nsx x nsy)
  const double2* ab = nullptr;
};

void image_twiddles(int nx, int ny, std::vector<double2>& twx, std::vector<double2>& twy);
void pupil_window(const int* pupil, int stride, int ncut, int& lpx, int& lpy);
AbbeSetup abbe_setup(char polar, const euvl_em_params* p, double kx0, double ky0, double z0, double z, int nrange,
                     int lmax, int mmax, int lpx, int lpy);
M3dSetup m3d_setup(const euvl_em_params* p, double kx0, double ky0, double z0, int lpmax_x, int lpmax_y,
                   int nrange, int lmax, int mmax);
size_t abbe_workspace_bytes(int n_src, int lpx, int lpy, int nx, int ny);
int abbe_enqueue(euvl_handle h, const AbbeSetup& a, const double2* Ax, int64_t ax_stride, int n_clips,
                 const ImagingDev& t, int n_src, int ncut, int nx, int ny, double weight, double* img,
                 int64_t img_stride, bool zero);
int m3d_enqueue(euvl_handle h, const M3dSetup& a, const double2* fmask, int64_t f_stride, const double2* Ax,
                int64_t ax_stride, int n_clips, const ImagingDev& t, int ncut, double2* a0, double2* ax,
                double2* ay);

}  // namespace euvl
