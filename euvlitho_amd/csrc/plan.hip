// Batched plans: the whole mask -> EM -> M3D -> Abbe pipeline of n_clips clips
// of one problem (one order set, one set of source divisions) in one pass of
// batched kernels, with every host-side table uploaded once when the plan is
// made.  The clip axis is the reference's ndata loop (m3d.cpp:197-558,
// intensity.cpp:219-408 once per mask): every EM product runs batched over
// clips x source divisions, so small problems fill the GPU.
//
// A run in fixed mode takes no decision on the host: the doubling count of
// the matrix functions is fixed in the plan (calibrated by adaptive runs with
// a margin) and checked on the device, the top solve always takes the unpivoted LU and its residual is
// checked on the device; failures set per-entry flags that the caller reads
// once per batch and redoes in adaptive mode.  A fixed run enqueues work only
// (no synchronisation, no allocation, no host copy), so it can be captured
// into a HIP graph (euvl_graph_begin / euvl_graph_end) and replayed.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ew.h"
#include "imaging.h"
#include "optics.h"

struct euvl_plan_s {
  euvl_handle h = nullptr;
  char polar = 'X';
  euvl_em_params p{};
  std::vector<double> layer_eps, layer_d;
  int N = 0, lmax = 0, mmax = 0, n_off = 0, n_clips = 0, E = 0;
  int ncut = 0, lpmx = 0, lpmy = 0;  // M3D fit
  int n_src = 0, nx = 0, ny = 0, ncut2 = 0, lpx = 0, lpy = 0;  // Abbe
  double kx0 = 0, ky0 = 0, z0 = 0, z = 0, weight = 1.0;
  char* dev = nullptr;  // all tables
  euvl::SolveDev sd;
  euvl::ImagingDev im;
  const double2* tw_dft = nullptr;
  int doublings = -1;
  double alpha_max = 0.0, alpha_margin = 1.0;  // calibration of the doubling count
  size_t ws_need = 0, aux_need = 0;
};

struct euvl_graph_s {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // the capture's handle and the device buffers of that handle its nodes address
  euvl_handle h = nullptr;
  char* ws = nullptr;
  size_t ws_cap = 0;
  char* aux = nullptr;
  size_t aux_cap = 0;
};

namespace {

size_t al256(size_t x) { return (x + 255) & ~size_t(255); }

// Spectrum window sizes of the run: the EM window (|l| <= 2 lmax, |m| <= 2 mmax)
// and the thin-mask window of the M3D fit (|l| <= lpmax_x, |m| <= lpmax_y).
size_t em_window(const euvl_plan_s* pl) { return (size_t)(4 * pl->lmax + 1) * (4 * pl->mmax + 1); }
size_t thin_window(const euvl_plan_s* pl) { return (size_t)(2 * pl->lpmx + 1) * (2 * pl->lpmy + 1); }

size_t dft_bytes(const euvl_plan_s* pl, int mw) {
  const size_t scratch = sizeof(double2) * (size_t)pl->n_clips * pl->p.ndivx * (2 * mw + 1);
  const int S = euvl::dft_splits(pl->p.ndivx, pl->p.ndivy, mw, pl->n_clips);
  return al256(scratch) + (S > 1 ? al256(S * scratch) : 0);
}

bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) return false;
  return st != hipStreamCaptureStatusNone;
}

}  // namespace

extern "C" int euvl_plan_create(euvl_handle h, const euvl_plan_desc* d, euvl_plan* out) {
  using namespace euvl;
  if (!h || !d || !out || !d->p || (d->polar != 'X' && d->polar != 'Y') || !d->h_lindex || !d->h_mindex ||
      d->nrange <= 0 || d->n_off <= 0 || !d->h_off || d->n_clips <= 0 || d->ncut < 0 || d->n_src < 0 ||
      (d->ncut > 0 && (!d->h_pupil3 || !d->h_vc || !d->h_ab)) ||
      (d->n_src > 0 && (!d->h_src || d->nx <= 0 || d->ny <= 0 || d->ncut <= 0))) {
    set_error("euvl_plan_create: bad argument");
    return EUVL_ERR_ARG;
  }
  const euvl_em_params* p = d->p;
  if (p->ndivx <= 0 || p->ndivy <= 0 || p->n_layers < 2 || !p->h_layer_eps || !p->h_layer_d ||
      ((int64_t)p->ndivx * p->ndivy) % 8 != 0) {
    set_error("euvl_plan_create: invalid euvl_em_params");
    return EUVL_ERR_ARG;
  }
  const int E = d->n_off * d->n_clips;
  if (E > 65535 || (int64_t)d->n_clips * (4 * d->lmax + 1) * (4 * d->mmax + 1) >= ((int64_t)1 << 31)) {
    set_error("euvl_plan_create: %d clips x %d offsets is too large a batch", d->n_clips, d->n_off);
    return EUVL_ERR_ARG;
  }
  for (int i = 0; i < d->nrange; ++i)
    if (std::abs(d->h_lindex[i]) > d->lmax || std::abs(d->h_mindex[i]) > d->mmax) {
      set_error("euvl_plan_create: order (%d,%d) outside lmax=%d mmax=%d", d->h_lindex[i], d->h_mindex[i], d->lmax,
                d->mmax);
      return EUVL_ERR_ARG;
    }
  for (int s = 0; s < d->n_src; ++s)
    if (d->h_src[3 * s] < 0 || d->h_src[3 * s] >= d->n_off) {
      set_error("euvl_plan_create: source %d names offset %d of %d", s, d->h_src[3 * s], d->n_off);
      return EUVL_ERR_ARG;
    }
  euvl_plan_s* pl = new euvl_plan_s();
  pl->h = h;
  pl->polar = d->polar;
  pl->p = *p;
  pl->layer_eps.assign(p->h_layer_eps, p->h_layer_eps + 2 * p->n_layers);
  pl->layer_d.assign(p->h_layer_d, p->h_layer_d + p->n_layers);
  pl->p.h_layer_eps = pl->layer_eps.data();
  pl->p.h_layer_d = pl->layer_d.data();
  pl->N = d->nrange;
  pl->lmax = d->lmax;
  pl->mmax = d->mmax;
  pl->n_off = d->n_off;
  pl->n_clips = d->n_clips;
  pl->E = E;
  pl->kx0 = d->kx0;
  pl->ky0 = d->ky0;
  pl->z0 = d->z0;
  pl->z = d->z;
  pl->weight = d->weight;
  pl->ncut = d->ncut;
  pl->n_src = d->n_src;
  pl->nx = d->nx;
  pl->ny = d->ny;
  pl->lpmx = d->lpmax_x;
  pl->lpmy = d->lpmax_y;
  pl->alpha_margin = d->alpha_margin > 1.0 ? d->alpha_margin : 1.0;
  if (d->ncut > 0) {
    for (int i = 0; i < d->ncut; ++i)
      if (std::abs(d->h_pupil3[3 * i]) > d->lpmax_x || std::abs(d->h_pupil3[3 * i + 1]) > d->lpmax_y) {
        set_error("euvl_plan_create: pupil point %d outside the thin-mask window", i);
        delete pl;
        return EUVL_ERR_ARG;
      }
    pupil_window(d->h_pupil3, 3, d->ncut, pl->lpx, pl->lpy);
  }
  // host tables
  std::vector<double> sx(d->n_off), sy(d->n_off);
  for (int o = 0; o < d->n_off; ++o) {
    sx[o] = d->h_off[4 * o];
    sy[o] = d->h_off[4 * o + 1];
  }
  SolveHost sh;
  build_solve_tables(&pl->p, d->h_lindex, d->h_mindex, pl->N, pl->lmax, pl->mmax, pl->n_off, sx.data(), sy.data(),
                     pl->n_clips, sh);
  std::vector<double2> tw;
  dft_twiddles(p->ndivx, p->ndivy, tw);
  std::vector<int> map((size_t)(2 * d->lmax + 1) * (2 * d->mmax + 1), -1);
  for (int i = 0; i < pl->N; ++i)
    map[(size_t)(d->h_lindex[i] + d->lmax) * (2 * d->mmax + 1) + (d->h_mindex[i] + d->mmax)] = i;
  std::vector<int> pup2;
  for (int i = 0; i < d->ncut; ++i) {
    pup2.push_back(d->h_pupil3[3 * i]);
    pup2.push_back(d->h_pupil3[3 * i + 1]);
  }
  std::vector<double2> twx, twy;
  if (d->n_src > 0) image_twiddles(d->nx, d->ny, twx, twy);
  const int nsxy = (2 * p->lsmax_x + 1) * (2 * p->lsmax_y + 1);
  Upload up;
  const size_t o_tw = up.add(tw.data(), sizeof(double2) * tw.size());
  const size_t o_map = up.add(map.data(), sizeof(int) * map.size());
  const size_t o_off = up.add(d->h_off, sizeof(double) * 4 * d->n_off);
  const size_t o_src = up.add(d->h_src, sizeof(int) * 3 * (size_t)d->n_src);
  const size_t o_p2 = up.add(pup2.data(), sizeof(int) * pup2.size());
  const size_t o_p3 = up.add(d->h_pupil3, sizeof(int) * 3 * (size_t)d->ncut);
  const size_t o_twx = up.add(twx.data(), sizeof(double2) * twx.size());
  const size_t o_twy = up.add(twy.data(), sizeof(double2) * twy.size());
  const size_t o_vc = up.add(d->h_vc, d->ncut > 0 ? sizeof(double2) * nsxy : 0);
  const size_t o_ab = up.add(d->h_ab, d->ncut > 0 ? sizeof(double2) * nsxy : 0);
  const size_t solve_bytes = solve_tables_bytes(&pl->p, pl->N, sh);
  const size_t o_solve = al256(up.host.size());
  if (hipMalloc(&pl->dev, o_solve + solve_bytes + 256) != hipSuccess) {
    (void)hipGetLastError();
    set_error("euvl_plan_create: table allocation failed");
    delete pl;
    return EUVL_ERR_NOMEM;
  }
  int rc = EUVL_OK;
  if (hipMemcpyAsync(pl->dev, up.host.data(), up.host.size(), hipMemcpyHostToDevice, h->stream) != hipSuccess) {
    set_error("euvl_plan_create: table upload failed");
    rc = EUVL_ERR_HIP;
  }
  Upload up2;
  if (rc == EUVL_OK)
    rc = upload_solve_tables(h, &pl->p, d->h_lindex, d->h_mindex, pl->N, sh, up2, pl->dev + o_solve, pl->sd);
  if (rc == EUVL_OK && hipStreamSynchronize(h->stream) != hipSuccess) {
    set_error("euvl_plan_create: upload synchronisation failed");
    rc = EUVL_ERR_HIP;
  }
  if (rc != EUVL_OK) {
    (void)hipFree(pl->dev);
    delete pl;
    return rc;
  }
  pl->tw_dft = (const double2*)(pl->dev + o_tw);
  pl->im.map = (const int*)(pl->dev + o_map);
  pl->im.off = (const double*)(pl->dev + o_off);
  pl->im.src = (const int*)(pl->dev + o_src);
  pl->im.pup2 = (const int*)(pl->dev + o_p2);
  pl->im.pup3 = (const int*)(pl->dev + o_p3);
  pl->im.twx = (const double2*)(pl->dev + o_twx);
  pl->im.twy = (const double2*)(pl->dev + o_twy);
  pl->im.vc = (const double2*)(pl->dev + o_vc);
  pl->im.ab = (const double2*)(pl->dev + o_ab);
  // workspace of a run: [EM spectra | thin spectra] then, reused in turn, the DFT
  // scratch, the solve's matrices and the Abbe buffers
  const size_t head = al256(sizeof(double2) * pl->n_clips * em_window(pl)) +
                      (pl->ncut > 0 ? al256(sizeof(double2) * pl->n_clips * thin_window(pl)) : 0);
  size_t body = std::max(dft_bytes(pl, 2 * pl->mmax), pl->ncut > 0 ? dft_bytes(pl, pl->lpmy) : 0);
  body = std::max(body, solve_workspace_bytes(pl->N, E, pl->sd.ninc) + 256 * 16);
  if (pl->n_src > 0) body = std::max(body, abbe_workspace_bytes(pl->n_src, pl->lpx, pl->lpy, pl->nx, pl->ny));
  pl->ws_need = head + body + 4096;
  pl->aux_need = al256(gj_solve_scratch_bytes(E));  // the top solve's diagonal-block inverses
  *out = pl;
  return EUVL_OK;
}

extern "C" int euvl_plan_destroy(euvl_plan pl) {
  if (!pl) return EUVL_OK;
  if (pl->h) (void)hipStreamSynchronize(pl->h->stream);
  if (pl->dev) (void)hipFree(pl->dev);
  delete pl;
  return EUVL_OK;
}

extern "C" int64_t euvl_plan_workspace_bytes(euvl_plan pl) { return pl ? (int64_t)pl->ws_need : 0; }

extern "C" int euvl_plan_decisions(euvl_plan pl, int* doublings) {
  if (!pl) return EUVL_ERR_ARG;
  if (doublings) *doublings = pl->doublings;
  return EUVL_OK;
}

extern "C" int euvl_plan_set_decisions(euvl_plan pl, int doublings) {
  if (!pl || doublings < 0 || doublings > 60) {
    euvl::set_error("euvl_plan_set_decisions: bad argument");
    return EUVL_ERR_ARG;
  }
  pl->doublings = doublings;
  pl->alpha_max = 0.0;  // later adaptive runs calibrate from these values up
  return EUVL_OK;
}

extern "C" int euvl_plan_run(euvl_plan pl, const uint8_t* d_bits, double* d_Ax, double* d_a0, double* d_ax,
                             double* d_ay, double* d_img, int* d_flags, int fixed) {
  using namespace euvl;
  if (!pl || !d_bits || !d_Ax || (pl->ncut > 0 && (!d_a0 || !d_ax || !d_ay)) || (pl->n_src > 0 && !d_img) ||
      (fixed && !d_flags)) {
    set_error("euvl_plan_run: bad argument");
    return EUVL_ERR_ARG;
  }
  if (fixed && pl->doublings < 0) {
    set_error("euvl_plan_run: fixed mode before the decisions are set (run once adaptive or set them)");
    return EUVL_ERR_ARG;
  }
  euvl_handle h = pl->h;
  if (capturing(h->stream)) {
    if (!fixed) {
      set_error("euvl_plan_run: only fixed-mode runs can be captured");
      return EUVL_ERR_ARG;
    }
    if (h->ws_cap < pl->ws_need || h->aux_cap < pl->aux_need) {
      set_error("euvl_plan_run: workspace not reserved before capture (run the plan once uncaptured first)");
      return EUVL_ERR_ARG;
    }
    if (h->prof) {
      set_error("euvl_plan_run: GEMM timing is on (euvl_gemm_stats) during capture");
      return EUVL_ERR_ARG;
    }
  }
  EUVL_TRY(ws_reserve(h, pl->ws_need));
  if (!aux_reserve(h, pl->aux_need)) {
    set_error("euvl_plan_run: LU scratch allocation failed");
    return EUVL_ERR_NOMEM;
  }
  const int nc = pl->n_clips, nx = pl->p.ndivx, ny = pl->p.ndivy;
  const int64_t wem = (int64_t)em_window(pl), wthin = (int64_t)thin_window(pl);
  double2* spec = (double2*)ws_take(h, sizeof(double2) * nc * wem);
  double2* thin = pl->ncut > 0 ? (double2*)ws_take(h, sizeof(double2) * nc * wthin) : nullptr;
  const size_t mark = h->ws_top;
  // spectra of every clip: the EM window, then the thin-mask window of the fit
  for (int pass = 0; pass < (pl->ncut > 0 ? 2 : 1); ++pass) {
    h->ws_top = mark;
    const int lw = pass == 0 ? 2 * pl->lmax : pl->lpmx, mw = pass == 0 ? 2 * pl->mmax : pl->lpmy;
    const size_t scratch = sizeof(double2) * (size_t)nc * nx * (2 * mw + 1);
    const int S = dft_splits(nx, ny, mw, nc);
    double2* sc = (double2*)ws_take(h, scratch);
    double2* part = S > 1 ? (double2*)ws_take(h, S * scratch) : nullptr;
    EUVL_TRY(mask_spectrum_enqueue(h, d_bits, nx, ny, nc, lw, mw, pass == 0 ? spec : thin, sc, part, pl->tw_dft));
  }
  h->ws_top = mark;
  if (fixed) EUVL_TRY(zero_ints(h, pl->E, d_flags));
  SolveCtl ctl;
  ctl.fixed = fixed != 0;
  ctl.doublings = pl->doublings;
  ctl.d_flags = d_flags;
  EUVL_TRY(amp_s_core(h, pl->polar, &pl->p, pl->sd, pl->N, pl->lmax, pl->mmax, spec, pl->n_off, wem, pl->E,
                      (double2*)d_Ax, ctl));
  if (!fixed) {
    // calibration over every adaptive run so far: the doublings the largest
    // norm bound alpha needs (times the plan's margin)
    const double theta = cos_sinc_theta(h);
    pl->alpha_max = std::max(pl->alpha_max, ctl.alpha);
    int s = 0;
    while (std::ldexp(pl->p.dabs, -s) * std::ldexp(pl->p.dabs, -s) * pl->alpha_margin * pl->alpha_max > theta &&
           s < 60)
      ++s;
    pl->doublings = std::max({pl->doublings, s, ctl.doublings_used});
  }
  h->ws_top = mark;
  const int nsxy = (2 * pl->p.lsmax_x + 1) * (2 * pl->p.lsmax_y + 1);
  const int64_t ax_clip = (int64_t)pl->n_off * nsxy * pl->N;  // Ax of a clip: [offset][ls][ms][n]
  if (pl->ncut > 0) {
    // the fit reads the central offset (the first division, no shift) of each clip
    const M3dSetup a = m3d_setup(&pl->p, pl->kx0, pl->ky0, pl->z0, pl->lpmx, pl->lpmy, pl->N, pl->lmax, pl->mmax);
    EUVL_TRY(m3d_enqueue(h, a, thin, wthin, (const double2*)d_Ax, ax_clip, nc, pl->im, pl->ncut, (double2*)d_a0,
                         (double2*)d_ax, (double2*)d_ay));
  }
  if (pl->n_src > 0) {
    const AbbeSetup a = abbe_setup(pl->polar, &pl->p, pl->kx0, pl->ky0, pl->z0, pl->z, pl->N, pl->lmax, pl->mmax,
                                   pl->lpx, pl->lpy);
    EUVL_TRY(abbe_enqueue(h, a, (const double2*)d_Ax, ax_clip, nc, pl->im, pl->n_src, pl->ncut, pl->nx, pl->ny,
                          pl->weight, d_img, (int64_t)pl->nx * pl->ny, true));
  }
  if (!fixed) EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));
  return EUVL_OK;
}

extern "C" int euvl_graph_begin(euvl_handle h) {
  if (!h || !h->stream) {
    euvl::set_error("euvl_graph_begin: the handle needs a stream of its own (not the null stream)");
    return EUVL_ERR_ARG;
  }
  EUVL_HIP_CHECK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  return EUVL_OK;
}

extern "C" int euvl_graph_end(euvl_handle h, euvl_graph* out) {
  if (!h || !out) {
    euvl::set_error("euvl_graph_end: bad argument");
    return EUVL_ERR_ARG;
  }
  euvl_graph_s* g = new euvl_graph_s();
  hipError_t e = hipStreamEndCapture(h->stream, &g->graph);
  if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    euvl::set_error("euvl_graph_end: %s", hipGetErrorString(e));
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return EUVL_ERR_HIP;
  }
  g->h = h;
  g->ws = h->ws;
  g->ws_cap = h->ws_cap;
  g->aux = h->aux;
  g->aux_cap = h->aux_cap;
  *out = g;
  return EUVL_OK;
}

extern "C" int euvl_graph_launch(euvl_graph g, euvl_handle h) {
  if (!g || !h || g->h != h) {
    euvl::set_error("euvl_graph_launch: bad argument (a graph replays on the handle it was captured on)");
    return EUVL_ERR_ARG;
  }
  // the nodes hold addresses inside the workspace and LU scratch of the
  // capture: a buffer that moved or shrank since would be read out of bounds
  if (h->ws != g->ws || h->ws_cap < g->ws_cap || h->aux != g->aux || h->aux_cap < g->aux_cap) {
    euvl::set_error("euvl_graph_launch: the handle's workspace was reallocated after the capture; capture again");
    return EUVL_ERR_STALE;
  }
  EUVL_HIP_CHECK(hipGraphLaunch(g->exec, h->stream));
  return EUVL_OK;
}

extern "C" int euvl_graph_nodes(euvl_graph g, int64_t* n) {
  if (!g || !n) return EUVL_ERR_ARG;
  size_t c = 0;
  EUVL_HIP_CHECK(hipGraphGetNodes(g->graph, nullptr, &c));
  *n = (int64_t)c;
  return EUVL_OK;
}

extern "C" int euvl_graph_destroy(euvl_graph g) {
  if (!g) return EUVL_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return EUVL_OK;
}

extern "C" int euvl_plan_reserve(euvl_handle h, int n_plans, const euvl_plan* plans) {
  if (!h || n_plans < 0 || (n_plans > 0 && !plans)) {
    euvl::set_error("euvl_plan_reserve: bad argument");
    return EUVL_ERR_ARG;
  }
  size_t ws = 0, aux = 0;
  for (int i = 0; i < n_plans; ++i) {
    if (!plans[i] || plans[i]->h != h) {
      euvl::set_error("euvl_plan_reserve: plan %d does not belong to this handle", i);
      return EUVL_ERR_ARG;
    }
    ws = std::max(ws, plans[i]->ws_need);
    aux = std::max(aux, plans[i]->aux_need);
  }
  EUVL_HIP_CHECK(hipStreamSynchronize(h->stream));
  EUVL_TRY(euvl::ws_reserve(h, ws));
  if (aux > 0 && !euvl::aux_reserve(h, aux)) {
    euvl::set_error("euvl_plan_reserve: LU scratch allocation of %zu bytes failed", aux);
    return EUVL_ERR_NOMEM;
  }
  return EUVL_OK;
}
