// C ABI of the EFF/OLA operator: plan object + operator entry points.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/vmbd.h"
#include "../../include/vmbd_solver.h"
#include "capi_util.h"
#include "col_device.h"
#include "fft_core.h"
#include "ola_kernels.h"

using namespace vmbd;

struct vmbd_plan {
  Geo g;
  int device;
  float2* tw = nullptr;
  float* win = nullptr;
  int* win_idx = nullptr;
  int* win_slot = nullptr;   // [Pf] slot of patch k among the distinct windows of its patch row
  int* row_class = nullptr;  // [csf0][nslot] window class of each slot (-1 unused)
  int nslot = 0;
  int n_win = 0;
  float2* wsG = nullptr;  // [frames][csf1][NB][row_pitch((csf0-1)*hop0 + ext0)] shared residual rows
  float2* wsX = nullptr;  // [P][NB][row_pitch(sw0)], or shared image rows [frames][csf1][NB][row_pitch(S0)]
  float2* wsF = nullptr;  // [P][NB][row_pitch(sf0)]
  float2* Y1 = nullptr;   // [P][ext0][NBP]
  float2* Y2 = nullptr;   // [P][sf0][NBP]
  size_t bytes = 0;
  int mc_fwd = 0, mc_adj = 0;
  int max_frames = 1;
  int rows_frames = 0;  // >0: wsX/wsF hold the row transforms of the last forward (that many frames)
  vmbd_tuning tune{};   // launch shapes (0 = automatic)
  // separable windows (vmbd_plan_set_window_factors): row and column factors,
  // the x-adjoint's column normalisers [2][W] (1/sum of the covering column
  // factors, 1/number of covering patch columns), rows whose covering
  // windows are all at the floor [H], and the output columns where that holds
  float* fac_a = nullptr;
  float* fac_b = nullptr;
  float* fac_colw = nullptr;
  int* fac_zrow = nullptr;
  int fac_bcol[2] = {-1, -1}, fac_bn[2] = {0, 0}, fac_bc[2] = {0, 0};
  // image rows shared by the patches of a patch column (separable windows):
  // the row scale of each patch row [csf0][sw0], the column weights of each
  // patch column and row type [csf1][2][sw1], the row type of each canvas
  // row (1: every covering window at the floor) and 1/(covering patch rows)
  // of the rows with real weights (0 on the others) [S0]
  int xs = 0;
  float* fac_rsc = nullptr;
  float* fac_cw = nullptr;
  float* fac_einv = nullptr;
  int* fac_czr = nullptr;
};

namespace {

bool pow2_in_range(int v) { return v >= 16 && v <= 512 && (v & (v - 1)) == 0; }

// max number of patch rows whose extent covers one coordinate
int max_cover(int csf, int hop, int ext, int lo, int hi) {
  int best = 0;
  for (int u = lo; u < hi; ++u) {
    if (u < 0) continue;
    int mhi = std::min(csf - 1, u / hop);
    int l = u - ext + 1;
    int mlo = l <= 0 ? 0 : (l + hop - 1) / hop;
    best = std::max(best, mhi - mlo + 1);
  }
  return best;
}

template <class T>
bool dev_alloc(T** p, size_t n, size_t& total) {
  if (n == 0) n = 1;
  if (hipMalloc((void**)p, n * sizeof(T)) != hipSuccess) return false;
  total += n * sizeof(T);
  return true;
}

void plan_free(vmbd_plan* p) {
  if (!p) return;
  (void)hipFree(p->tw);
  (void)hipFree(p->win);
  (void)hipFree(p->win_idx);
  (void)hipFree(p->win_slot);
  (void)hipFree(p->row_class);
  (void)hipFree(p->wsG);
  (void)hipFree(p->wsX);
  (void)hipFree(p->wsF);
  (void)hipFree(p->Y1);
  (void)hipFree(p->Y2);
  (void)hipFree(p->fac_a);
  (void)hipFree(p->fac_b);
  (void)hipFree(p->fac_colw);
  (void)hipFree(p->fac_zrow);
  (void)hipFree(p->fac_rsc);
  (void)hipFree(p->fac_cw);
  (void)hipFree(p->fac_einv);
  (void)hipFree(p->fac_czr);
  delete p;
}

RowSrc src_image(const vmbd_plan* p, const float* x) {
  const Geo& g = p->g;
  RowSrc s{};
  s.src = x;
  s.pstride = 0;
  s.fstride = (long long)g.H * g.W;
  s.ld = g.W;
  s.srcH = g.H;
  s.srcW = g.W;
  s.origin = 1;
  s.sh0 = s.sh1 = 0;
  s.rlo = 0;
  s.rhi = g.sw0;
  s.clo = 0;
  s.chi = g.sw1;
  s.win = p->win;
  s.win_idx = p->win_idx;
  return s;
}

RowSrc src_psf(const vmbd_plan* p, const float* f) {
  const Geo& g = p->g;
  RowSrc s{};
  s.src = f;
  s.pstride = (long long)g.sf0 * g.sf1;
  s.fstride = 0;
  s.ld = g.sf1;
  s.srcH = g.sf0;
  s.srcW = g.sf1;
  s.origin = 0;
  s.sh0 = s.sh1 = 0;
  s.rlo = 0;
  s.rhi = g.sf0;
  s.clo = 0;
  s.chi = g.sf1;
  s.win = nullptr;
  s.win_idx = nullptr;
  return s;
}

// residual patch G_k: tile row b holds g[p_k + b - oc]; exact keeps the full
// correlation support [0, ext), the reference-local form only [oc, oc+sw).
RowSrc src_residual(const vmbd_plan* p, const float* gres, int exact) {
  const Geo& g = p->g;
  RowSrc s{};
  s.src = gres;
  s.pstride = 0;
  s.fstride = (long long)g.sy0 * g.sy1;
  s.ld = g.sy1;
  s.srcH = g.sy0;
  s.srcW = g.sy1;
  s.origin = 1;
  s.sh0 = g.oc0;
  s.sh1 = g.oc1;
  if (exact) {
    s.rlo = 0;
    s.rhi = g.ext0;
    s.clo = 0;
    s.chi = g.ext1;
  } else {
    s.rlo = g.oc0;
    s.rhi = g.oc0 + g.sw0;
    s.clo = g.oc1;
    s.chi = g.oc1 + g.sw1;
  }
  s.win = nullptr;
  s.win_idx = nullptr;
  return s;
}

ColIn raw_in(const float2* ws, const RowSrc& s) {
  return ColIn{ws, s.rlo, s.rhi - s.rlo, 0, row_pitch(s.rhi - s.rlo), 0};
}
ColIn spec_in(const void* spec, int Th) { return ColIn{(const float2*)spec, 0, Th, 1, Th, 0}; }

// Residual rows shared by the patches of a patch column.  Patch (m, n) reads
// tile row b from residual row m*hop0 + b - oc0 over columns that depend on n
// only, so the row transform of virtual row v = m*hop0 + b is the same for
// every covering m: the row pass runs once per (frame, patch column) over
// v in [rlo, rlo + (csf0-1)*hop0 + nr), on a geometry with one patch row
// whose patches are the patch columns.
struct SharedRows {
  Geo g;      // row-pass geometry: csf0 = 1, Pf = csf1
  RowSrc s;   // rows [rlo, rlo + nv)
  int nr;     // rows of one patch
  int nv;     // virtual rows of a patch column
};

int residual_virtual_rows(const Geo& g, int nr) { return (g.csf0 - 1) * g.hop0 + nr; }

SharedRows residual_rows(const vmbd_plan* p, const Geo& g, const float* gres, int exact) {
  SharedRows r;
  r.s = src_residual(p, gres, exact);
  r.nr = r.s.rhi - r.s.rlo;
  r.nv = residual_virtual_rows(g, r.nr);
  r.s.rhi = r.s.rlo + r.nv;
  r.g = g;
  r.g.csf0 = 1;
  r.g.Pf = g.csf1;
  r.g.P = (g.P / g.Pf) * g.csf1;
  return r;
}

ColIn shared_in(const float2* ws, const SharedRows& r) {
  return ColIn{ws, r.s.rlo, r.nr, 0, row_pitch(r.nv), 1};
}

// Image rows of a call: one transform per (patch, tile row), windowed by the
// patch's window; with separable windows one per (frame, patch column,
// canvas row) over the S0 = (csf0-1)*hop0 + sw0 canvas rows, weighted by the
// patch column's column weights of the row's type, and the column pass
// scales tile row r of patch row m by rsc[m][r] (and adds the floor
// columns' own terms): the rows of one image row are the same for every
// patch row that covers it, up to that scale.
struct XRows {
  Geo g;        // row-pass geometry
  RowSrc s;
  ColIn in;     // the column pass's operand
  bool shared;  // one transform per canvas row and patch column
};

// Whether a call shares the image rows (separable windows; the tuning's
// image_rows 1 forces it, 2 forbids it).  A call that transforms the image
// rows alone (the image spectra, the forward of cached PSF spectra) shares
// them: fewer transforms for a small factor load in the column pass.  A
// call that transforms image and PSF rows together, and the f-adjoint,
// which must match it bit for bit, shares them only from XS_MIN_BLOCKS
// per-patch image-row workgroups on: in smaller launches, which leave most
// of the GPU idle, the column pass's factor load lengthens the kernels more
// than the fewer row transforms save (1024^2 frames, 10x10 patches, tile
// 256: 600 workgroups per frame; 1 frame runs faster per patch, 2 and more
// shared).
constexpr long long XS_MIN_BLOCKS = 1024;

bool shared_x(const vmbd_plan* p, const Geo& g, bool alone) {
  if (!p->xs || p->tune.image_rows == 2) return false;
  if (p->tune.image_rows == 1 || alone) return true;
  const int TPW = 256 / (g.Tw / 16);
  return (long long)g.P * (((g.sw0 + 1) / 2 + TPW - 1) / TPW) >= XS_MIN_BLOCKS;
}

// alone: the call transforms the image rows in a launch of their own
XRows x_rows(const vmbd_plan* p, const Geo& g, const float* x, bool alone) {
  XRows r;
  r.g = g;
  r.s = src_image(p, x);
  r.shared = shared_x(p, g, alone);
  if (!r.shared) {
    r.in = raw_in(p->wsX, r.s);
    return r;
  }
  const int S0 = (g.csf0 - 1) * g.hop0 + g.sw0;
  r.s.rlo = 0;
  r.s.rhi = S0;
  r.s.win = nullptr;
  r.s.win_idx = nullptr;
  r.s.cw = p->fac_cw;
  r.s.czr = p->fac_czr;
  r.g.csf0 = 1;
  r.g.Pf = g.csf1;
  r.g.P = (g.P / g.Pf) * g.csf1;
  r.in = ColIn{p->wsX, 0, g.sw0, 0, row_pitch(S0), 1};
  r.in.rsc = p->fac_rsc;
  r.in.ex = x;
  r.in.exf = (long long)g.H * g.W;
  r.in.exld = g.W;
  r.in.exH = g.H;
  r.in.einv = p->fac_einv;
  for (int e = 0; e < 2; ++e) {
    r.in.ecol[e] = p->fac_bcol[e];
    r.in.en[e] = p->fac_bn[e];
    r.in.ec[e] = p->fac_bc[e];
  }
  return r;
}

float inv_scale(const Geo& g) { return 1.0f / (float)((double)g.Th * (double)g.Tw); }

OlaDst dst_forward(const vmbd_plan* p, float* y) {
  const Geo& g = p->g;
  OlaDst d{};
  d.out = y;
  d.scale = inv_scale(g);
  d.O0 = g.sy0;
  d.O1 = g.sy1;
  d.oc0 = g.oc0;
  d.oc1 = g.oc1;
  d.ext0 = g.ext0;
  d.ext1 = g.ext1;
  d.src = p->Y1;
  d.srow = g.ext0;
  d.win = nullptr;
  d.win_idx = nullptr;
  d.mc = p->mc_fwd;
  ola_layout(g, d);
  return d;
}

OlaDst dst_adjoint_x(const vmbd_plan* p, float* xbar) {
  const Geo& g = p->g;
  OlaDst d{};
  d.out = xbar;
  d.scale = inv_scale(g);
  d.O0 = g.H;
  d.O1 = g.W;
  d.oc0 = 0;
  d.oc1 = 0;
  d.ext0 = g.sw0;
  d.ext1 = g.sw1;
  d.src = p->Y1;
  d.srow = g.sw0;
  d.win = p->win;
  d.win_idx = p->win_idx;
  d.win_slot = p->win_slot;
  d.row_class = p->row_class;
  d.nslot = p->nslot;
  d.mc = p->mc_adj;
  d.bcol[0] = d.bcol[1] = -1;
  if (p->fac_a) {  // separable windows: summed form
    d.wrow = p->fac_a;
    d.wcol = p->fac_b;
    d.colw = p->fac_colw;
    d.zrow = p->fac_zrow;
    for (int e = 0; e < 2; ++e) {
      d.bcol[e] = p->fac_bcol[e];
      d.bn[e] = p->fac_bn[e];
      d.bc[e] = p->fac_bc[e];
    }
  }
  ola_layout(g, d);
  return d;
}

// row transforms of x and f: one launch over both (one geometry, or, with
// the image rows shared per patch column, a geometry per job)
hipError_t launch_xf_rows(const vmbd_plan* p, const Geo& g, const XRows& xr, const RowSrc& sf, hipStream_t st) {
  if (!xr.shared) {
    RowSrc srcs[2] = {xr.s, sf};
    float2* outs[2] = {p->wsX, p->wsF};
    return launch_rows_r2c_multi(g, srcs, outs, 2, p->tw, st);
  }
  // the shared image rows (one geometry per patch column) beside the PSF rows
  const RowSrc srcs[2] = {xr.s, sf};
  float2* outs[2] = {p->wsX, p->wsF};
  const int P[2] = {xr.g.P, g.P}, Pf[2] = {xr.g.Pf, g.Pf};
  return launch_rows_r2c_geo(g, srcs, outs, P, Pf, 2, p->tw, st, "rows_r2c_x_psf");
}

#define VMBD_LAUNCH(expr)                                            \
  do {                                                               \
    hipError_t e__ = (expr);                                         \
    if (e__ != hipSuccess) return vmbd_fail_hip(#expr, e__);         \
  } while (0)

int check_plan(const vmbd_plan* p, int nframes) {
  if (!p) return vmbd_fail(VMBD_E_ARG, "null plan");
  if (nframes < 1 || nframes > p->max_frames)
    return vmbd_fail(VMBD_E_ARG, "nframes=%d outside [1, max_frames=%d]", nframes, p->max_frames);
  // the plan's workspaces live on the device it was created on
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != p->device)
    return vmbd_fail(VMBD_E_ARG, "plan belongs to device %d, the calling thread's current device is %d", p->device,
                     dev);
  return VMBD_OK;
}

// geometry of one call: all patches of nframes frames
Geo call_geo(const vmbd_plan* p, int nframes) {
  Geo g = p->g;
  g.P = nframes * g.Pf;
  return g;
}

}  // namespace

extern "C" {

int vmbd_version(void) { return 200; }

vmbd_plan* vmbd_plan_create(const vmbd_geom* gm, const float* win_table, int32_t n_win, const int32_t* win_idx) {
  vmbd_clear_error();
  if (!gm || !win_table || !win_idx || n_win <= 0) {
    vmbd_fail(VMBD_E_ARG, "vmbd_plan_create: null geometry/window arguments");
    return nullptr;
  }
  const vmbd_geom G = *gm;
  if (G.H <= 0 || G.W <= 0 || G.sfh <= 0 || G.sfw <= 0 || G.csf0 <= 0 || G.csf1 <= 0 || G.sw0 <= 0 ||
      G.sw1 <= 0 || G.nhop0 <= 0 || G.nhop1 <= 0) {
    vmbd_fail(VMBD_E_ARG, "vmbd_plan_create: non-positive size in geometry");
    return nullptr;
  }
  if (!pow2_in_range(G.Th) || !pow2_in_range(G.Tw)) {
    vmbd_fail(VMBD_E_UNSUP, "vmbd_plan_create: tile %dx%d must be powers of two in [16,512]", G.Th, G.Tw);
    return nullptr;
  }
  if (G.sw0 + G.sfh - 1 > G.Th || G.sw1 + G.sfw - 1 > G.Tw) {
    vmbd_fail(VMBD_E_ARG, "vmbd_plan_create: tile %dx%d smaller than sw+sf-1 = %dx%d", G.Th, G.Tw,
              G.sw0 + G.sfh - 1, G.sw1 + G.sfw - 1);
    return nullptr;
  }
  Geo g{};
  g.H = G.H;
  g.W = G.W;
  g.sf0 = G.sfh;
  g.sf1 = G.sfw;
  g.csf0 = G.csf0;
  g.csf1 = G.csf1;
  g.Pf = G.csf0 * G.csf1;
  g.P = g.Pf;
  g.sw0 = G.sw0;
  g.sw1 = G.sw1;
  g.hop0 = G.nhop0;
  g.hop1 = G.nhop1;
  g.Th = G.Th;
  g.Tw = G.Tw;
  g.NB = G.Tw / 2 + 1;
  g.NBP = (g.NB + 15) & ~15;
  g.ext0 = G.sw0 + G.sfh - 1;
  g.ext1 = G.sw1 + G.sfw - 1;
  switch (G.mode) {
    case VMBD_MODE_VALID:
      g.oc0 = G.sfh - 1;
      g.oc1 = G.sfw - 1;
      g.sy0 = G.H - G.sfh + 1;
      g.sy1 = G.W - G.sfw + 1;
      break;
    case VMBD_MODE_SAME:
      g.oc0 = G.sfh / 2;
      g.oc1 = G.sfw / 2;
      g.sy0 = G.H;
      g.sy1 = G.W;
      break;
    case VMBD_MODE_FULL:
      g.oc0 = 0;
      g.oc1 = 0;
      g.sy0 = G.H + G.sfh - 1;
      g.sy1 = G.W + G.sfw - 1;
      break;
    default:
      vmbd_fail(VMBD_E_ARG, "vmbd_plan_create: Not a valid mode!");
      return nullptr;
  }
  if (g.sy0 <= 0 || g.sy1 <= 0) {
    vmbd_fail(VMBD_E_ARG, "vmbd_plan_create: PSF larger than image");
    return nullptr;
  }
  // every patch origin must lie inside the padded image (sxp >= H, W)
  const long long sxp0 = (long long)G.nhop0 * (G.csf0 - 1) + G.sw0;
  const long long sxp1 = (long long)G.nhop1 * (G.csf1 - 1) + G.sw1;
  if (sxp0 < G.H || sxp1 < G.W) {
    vmbd_fail(VMBD_E_ARG, "vmbd_plan_create: patch grid covers %lldx%lld < image %dx%d", sxp0, sxp1, G.H, G.W);
    return nullptr;
  }
  if (G.max_frames < 0 || G.max_frames > 65535) {
    vmbd_fail(VMBD_E_ARG, "vmbd_plan_create: max_frames=%d outside [0, 65535]", G.max_frames);
    return nullptr;
  }
  if ((long long)G.csf0 * G.csf1 > 65535) {
    // the column pass puts the patch on the second grid dimension (the frame on the third)
    vmbd_fail(VMBD_E_ARG, "vmbd_plan_create: csf0*csf1=%lld patches exceed 65535", (long long)G.csf0 * G.csf1);
    return nullptr;
  }
  for (int k = 0; k < g.Pf; ++k) {
    if (win_idx[k] < 0 || win_idx[k] >= n_win) {
      vmbd_fail(VMBD_E_ARG, "vmbd_plan_create: win_idx[%d]=%d outside [0,%d)", k, win_idx[k], n_win);
      return nullptr;
    }
  }
  vmbd_plan* p = new vmbd_plan();
  p->g = g;
  p->n_win = n_win;
  p->max_frames = G.max_frames > 0 ? G.max_frames : 1;
  (void)hipGetDevice(&p->device);
  // distinct windows per patch row (the x-adjoint stages one row of each)
  std::vector<int> slot(g.Pf), rclass;
  {
    std::vector<std::vector<int>> rows(g.csf0);
    for (int m = 0; m < g.csf0; ++m)
      for (int n = 0; n < g.csf1; ++n) {
        const int c = win_idx[m * g.csf1 + n];
        auto& v = rows[m];
        int sidx = (int)(std::find(v.begin(), v.end(), c) - v.begin());
        if (sidx == (int)v.size()) v.push_back(c);
        slot[m * g.csf1 + n] = sidx;
      }
    for (auto& v : rows) p->nslot = std::max(p->nslot, (int)v.size());
    rclass.assign((size_t)g.csf0 * p->nslot, -1);
    for (int m = 0; m < g.csf0; ++m)
      for (size_t j = 0; j < rows[m].size(); ++j) rclass[(size_t)m * p->nslot + j] = rows[m][j];
  }
  // rows of the OLA gather buffer
  p->mc_fwd = max_cover(g.csf0, g.hop0, g.ext0, g.oc0, g.oc0 + g.sy0);
  p->mc_adj = max_cover(g.csf0, g.hop0, g.sw0, 0, g.H);
  {
    const size_t s_fwd = ola_shared_bytes(g, dst_forward(p, nullptr));
    const size_t s_adj = ola_shared_bytes(g, dst_adjoint_x(p, nullptr));
    if (s_fwd > 160 * 1024 || s_adj > 160 * 1024) {
      vmbd_fail(VMBD_E_UNSUP, "vmbd_plan_create: OLA gather needs %zu B of LDS > 160 KiB (csf1=%d ext1=%d)",
                std::max(s_fwd, s_adj), g.csf1, g.ext1);
      delete p;
      return nullptr;
    }
  }
  size_t tot = 0;
  const size_t P = (size_t)g.Pf * p->max_frames, NB = g.NB, NBP = g.NBP;
  bool ok = dev_alloc(&p->tw, TW_LEN, tot) && dev_alloc(&p->win, (size_t)n_win * g.sw0 * g.sw1, tot) &&
            dev_alloc(&p->win_idx, g.Pf, tot) && dev_alloc(&p->wsG, (size_t)p->max_frames * g.csf1 * NB * row_pitch(residual_virtual_rows(g, g.ext0)), tot) &&
            dev_alloc(&p->wsX, P * NB * row_pitch(g.sw0), tot) && dev_alloc(&p->wsF, P * NB * row_pitch(g.sf0), tot) &&
            dev_alloc(&p->Y1, P * NBP * g.ext0, tot) && dev_alloc(&p->Y2, P * NBP * g.sf0, tot) &&
            dev_alloc(&p->win_slot, g.Pf, tot) && dev_alloc(&p->row_class, rclass.size(), tot);
  if (!ok) {
    plan_free(p);
    vmbd_fail(VMBD_E_NOMEM, "vmbd_plan_create: hipMalloc failed");
    return nullptr;
  }
  p->bytes = tot;
  std::vector<float2> tw(TW_LEN);
  for (int m = 0; m < TW_LEN; ++m) {
    const double a = -2.0 * M_PI * (double)m / (double)TW_LEN;
    tw[m] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  hipError_t e = hipMemcpy(p->tw, tw.data(), TW_LEN * sizeof(float2), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(p->win, win_table, (size_t)n_win * g.sw0 * g.sw1 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->win_idx, win_idx, (size_t)g.Pf * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->win_slot, slot.data(), slot.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(p->row_class, rclass.data(), rclass.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    plan_free(p);
    vmbd_fail_hip("vmbd_plan_create: hipMemcpy", e);
    return nullptr;
  }
  return p;
}

void vmbd_plan_destroy(vmbd_plan* plan) { plan_free(plan); }

int vmbd_plan_set_tuning(vmbd_plan* p, const vmbd_tuning* t) {
  vmbd_clear_error();
  if (!p || !t) return vmbd_fail(VMBD_E_ARG, "vmbd_plan_set_tuning: null argument");
  const int o = t->ola_threads, c = t->col_threads;
  if (o != 0 && o != 128 && o != 192 && o != 256)
    return vmbd_fail(VMBD_E_ARG, "vmbd_plan_set_tuning: ola_threads=%d not in {0, 128, 192, 256}", o);
  if (c != 0 && c != 128 && c != 256)
    return vmbd_fail(VMBD_E_ARG, "vmbd_plan_set_tuning: col_threads=%d not in {0, 128, 256}", c);
  if ((o == 128 || o == 192) && !(ola_fits(p->g, dst_forward(p, nullptr), o) &&
                                  ola_fits(p->g, dst_adjoint_x(p, nullptr), o)))
    return vmbd_fail(VMBD_E_UNSUP, "vmbd_plan_set_tuning: an output row needs more than one round of %d threads", o);
  if (t->image_rows < 0 || t->image_rows > 2)
    return vmbd_fail(VMBD_E_ARG, "vmbd_plan_set_tuning: image_rows=%d not in {0, 1, 2}", t->image_rows);
  if (t->image_rows == 1 && !p->xs)
    return vmbd_fail(VMBD_E_UNSUP, "vmbd_plan_set_tuning: image_rows=1 needs separable windows");
  p->tune = *t;
  p->rows_frames = 0;  // the image-row layout may change
  return VMBD_OK;
}

int vmbd_plan_set_window_factors(vmbd_plan* p, const float* a, const float* b, float floor_) {
  vmbd_clear_error();
  if (!p || !a || !b) return vmbd_fail(VMBD_E_ARG, "vmbd_plan_set_window_factors: null argument");
  if (!(floor_ >= 0.f)) return vmbd_fail(VMBD_E_ARG, "vmbd_plan_set_window_factors: floor must be >= 0");
  if (p->fac_a) return vmbd_fail(VMBD_E_ARG, "vmbd_plan_set_window_factors: factors already set");
  const Geo& g = p->g;
  const int sw0 = g.sw0, sw1 = g.sw1, hop0 = g.hop0, hop1 = g.hop1;
  const int S0 = (g.csf0 - 1) * hop0 + sw0, S1 = (g.csf1 - 1) * hop1 + sw1;  // overlap-add canvas
  // the shared image rows ([frames][csf1][NB][row_pitch(S0)]) live in the
  // per-patch image-row workspace; it holds them whenever hop0 <= sw0
  if ((size_t)g.csf1 * row_pitch(S0) > (size_t)g.Pf * row_pitch(sw0))
    return vmbd_fail(VMBD_E_UNSUP, "vmbd_plan_set_window_factors: shared image rows exceed the row workspace");
  // nom and its overlap-add in FP64
  std::vector<double> nom((size_t)sw0 * sw1), wimg((size_t)S0 * S1, 0.0);
  for (int r = 0; r < sw0; ++r)
    for (int c = 0; c < sw1; ++c) nom[(size_t)r * sw1 + c] = std::max((double)a[r] * (double)b[c], (double)floor_);
  for (int m = 0; m < g.csf0; ++m)
    for (int n = 0; n < g.csf1; ++n)
      for (int r = 0; r < sw0; ++r)
        for (int c = 0; c < sw1; ++c) wimg[(size_t)(m * hop0 + r) * S1 + n * hop1 + c] += nom[(size_t)r * sw1 + c];
  // the plan's windows must be nom / wimg
  std::vector<float> tab((size_t)p->n_win * sw0 * sw1);
  std::vector<int> idx(g.Pf);
  hipError_t e = hipMemcpy(tab.data(), p->win, tab.size() * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(idx.data(), p->win_idx, idx.size() * sizeof(int), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return vmbd_fail_hip("vmbd_plan_set_window_factors: hipMemcpy", e);
  for (int k = 0; k < g.Pf; ++k) {
    const int m = k / g.csf1, n = k - m * g.csf1;
    const float* w = tab.data() + (size_t)idx[k] * sw0 * sw1;
    for (int r = 0; r < sw0; ++r)
      for (int c = 0; c < sw1; ++c) {
        const double ref = nom[(size_t)r * sw1 + c] / wimg[(size_t)(m * hop0 + r) * S1 + n * hop1 + c];
        if (!(std::fabs((double)w[(size_t)r * sw1 + c] - ref) <= 1e-5 * std::fabs(ref) + 1e-30))
          return vmbd_fail(VMBD_E_UNSUP,
                           "vmbd_plan_set_window_factors: window of patch %d differs from nom/wimg at (%d, %d)", k,
                           r, c);
      }
  }
  // a row (column) of the canvas is at the floor when every covering patch
  // row (column) has a zero factor there; a pixel on such a row or column
  // has the reference weight 1/count for each covering patch
  auto floor_line = [](const float* f, int csf, int hop, int sw, int v) {
    int cover = 0;
    for (int m = 0; m < csf; ++m) {
      const int r = v - m * hop;
      if (r < 0 || r >= sw) continue;
      ++cover;
      if (f[r] != 0.f) return false;
    }
    return cover > 0;
  };
  const int H = g.H, W = g.W;  // x-adjoint output: canvas rows [0, H), columns [0, W)
  std::vector<int> zrow(H);
  for (int u = 0; u < H; ++u) zrow[u] = u < S0 && floor_line(a, g.csf0, hop0, sw0, u);
  int nb = 0, bcol[2] = {-1, -1}, bn[2] = {0, 0}, bc[2] = {0, 0};
  for (int v = 0; v < W && v < S1; ++v) {
    if (!floor_line(b, g.csf1, hop1, sw1, v)) continue;
    int cover = 0, nn = 0;
    for (int n = 0; n < g.csf1; ++n)
      if (v - n * hop1 >= 0 && v - n * hop1 < sw1) {
        ++cover;
        nn = n;
      }
    if (nb == 2 || cover != 1)
      return vmbd_fail(VMBD_E_UNSUP, "vmbd_plan_set_window_factors: floor column %d not at the canvas edge", v);
    bcol[nb] = v;
    bn[nb] = nn;
    bc[nb] = v - nn * hop1;
    ++nb;
  }
  // wimg(u, v) = A(u) B(v) away from the floor (A, B: sums of the covering
  // row and column factors); on a floor row the weights are 1/count =
  // 1/(rows * columns covering the pixel).  Per output column: 1/B(v) and
  // 1/(number of covering patch columns)
  std::vector<float> colw((size_t)2 * W, 0.f);
  for (int v = 0; v < W && v < S1; ++v) {
    double B = 0.0;
    int cn = 0;
    for (int n = 0; n < g.csf1; ++n) {
      const int c = v - n * hop1;
      if (c < 0 || c >= sw1) continue;
      B += b[c];
      ++cn;
    }
    colw[v] = B > 0.0 ? (float)(1.0 / B) : 0.f;
    colw[(size_t)W + v] = cn ? (float)(1.0 / cn) : 0.f;
  }
  // image rows shared per patch column: tile row r of patch row m is
  // rsc[m][r] times the canvas row u = m*hop0 + r weighted by the column
  // weights of u's type; wimg = A(u) B(v) away from the floor, 1/count on it
  std::vector<float> rsc((size_t)g.csf0 * sw0), cw((size_t)g.csf1 * 2 * sw1), einv(S0);
  std::vector<int> czr(S0);
  std::vector<double> A(S0, 0.0);
  std::vector<int> M(S0, 0);
  for (int u = 0; u < S0; ++u) {
    for (int m = 0; m < g.csf0; ++m) {
      const int r = u - m * hop0;
      if (r < 0 || r >= sw0) continue;
      A[u] += a[r];
      ++M[u];
    }
    czr[u] = floor_line(a, g.csf0, hop0, sw0, u);
    einv[u] = (czr[u] || !M[u]) ? 0.f : (float)(1.0 / M[u]);
  }
  for (int m = 0; m < g.csf0; ++m)
    for (int r = 0; r < sw0; ++r) {
      const int u = m * hop0 + r;
      rsc[(size_t)m * sw0 + r] = czr[u] ? (float)(1.0 / M[u]) : (A[u] > 0.0 ? (float)(a[r] / A[u]) : 0.f);
    }
  for (int n = 0; n < g.csf1; ++n)
    for (int c = 0; c < sw1; ++c) {
      const int v = n * hop1 + c;
      double B = 0.0;
      int cn = 0;
      for (int n2 = 0; n2 < g.csf1; ++n2) {
        const int c2 = v - n2 * hop1;
        if (c2 < 0 || c2 >= sw1) continue;
        B += b[c2];
        ++cn;
      }
      cw[((size_t)n * 2 + 0) * sw1 + c] = B > 0.0 ? (float)(b[c] / B) : 0.f;
      cw[((size_t)n * 2 + 1) * sw1 + c] = cn ? (float)(1.0 / cn) : 0.f;
    }
  size_t tot = 0;
  if (!(dev_alloc(&p->fac_a, sw0, tot) && dev_alloc(&p->fac_b, sw1, tot) && dev_alloc(&p->fac_colw, colw.size(), tot) &&
        dev_alloc(&p->fac_zrow, zrow.size(), tot) && dev_alloc(&p->fac_rsc, rsc.size(), tot) &&
        dev_alloc(&p->fac_cw, cw.size(), tot) && dev_alloc(&p->fac_einv, einv.size(), tot) &&
        dev_alloc(&p->fac_czr, czr.size(), tot))) {
    (void)hipFree(p->fac_a);
    (void)hipFree(p->fac_b);
    (void)hipFree(p->fac_colw);
    (void)hipFree(p->fac_zrow);
    (void)hipFree(p->fac_rsc);
    (void)hipFree(p->fac_cw);
    (void)hipFree(p->fac_einv);
    (void)hipFree(p->fac_czr);
    p->fac_a = p->fac_b = p->fac_colw = p->fac_rsc = p->fac_cw = p->fac_einv = nullptr;
    p->fac_zrow = p->fac_czr = nullptr;
    return vmbd_fail(VMBD_E_NOMEM, "vmbd_plan_set_window_factors: hipMalloc failed");
  }
  e = hipMemcpy(p->fac_a, a, sw0 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->fac_b, b, sw1 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->fac_colw, colw.data(), colw.size() * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->fac_zrow, zrow.data(), zrow.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->fac_rsc, rsc.data(), rsc.size() * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->fac_cw, cw.data(), cw.size() * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->fac_einv, einv.data(), einv.size() * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->fac_czr, czr.data(), czr.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e != hipSuccess) return vmbd_fail_hip("vmbd_plan_set_window_factors: hipMemcpy", e);
  for (int e2 = 0; e2 < 2; ++e2) {
    p->fac_bcol[e2] = bcol[e2];
    p->fac_bn[e2] = bn[e2];
    p->fac_bc[e2] = bc[e2];
  }
  p->bytes += tot;
  // the summed form's LDS must fit (it needs less than the per-patch form)
  if (ola_shared_bytes(g, dst_adjoint_x(p, nullptr)) > 160 * 1024) {
    (void)hipFree(p->fac_a);
    (void)hipFree(p->fac_b);
    (void)hipFree(p->fac_colw);
    (void)hipFree(p->fac_zrow);
    (void)hipFree(p->fac_rsc);
    (void)hipFree(p->fac_cw);
    (void)hipFree(p->fac_einv);
    (void)hipFree(p->fac_czr);
    p->fac_a = p->fac_b = p->fac_colw = p->fac_rsc = p->fac_cw = p->fac_einv = nullptr;
    p->fac_zrow = p->fac_czr = nullptr;
    p->bytes -= tot;
    return vmbd_fail(VMBD_E_UNSUP, "vmbd_plan_set_window_factors: summed x-adjoint does not fit the LDS");
  }
  p->xs = 1;          // image rows shared per patch column from now on
  p->rows_frames = 0;  // the workspaces hold no rows of the new layout
  return VMBD_OK;
}

int vmbd_plan_image_rows_shared(const vmbd_plan* p, int32_t nframes) {
  vmbd_clear_error();
  if (!p) return vmbd_fail(VMBD_E_ARG, "vmbd_plan_image_rows_shared: null plan");
  if (nframes < 1 || nframes > p->max_frames)
    return vmbd_fail(VMBD_E_ARG, "nframes=%d outside [1, max_frames=%d]", nframes, p->max_frames);
  return shared_x(p, call_geo(p, nframes), false) ? 1 : 0;
}

int vmbd_plan_get_tuning(const vmbd_plan* p, vmbd_tuning* t) {
  if (!p || !t) return vmbd_fail(VMBD_E_ARG, "vmbd_plan_get_tuning: null argument");
  *t = p->tune;
  return VMBD_OK;
}

size_t vmbd_plan_workspace_bytes(const vmbd_plan* plan) { return plan ? plan->bytes : 0; }

int vmbd_plan_out_shape(const vmbd_plan* plan, int32_t* sy0, int32_t* sy1) {
  if (!plan || !sy0 || !sy1) return vmbd_fail(VMBD_E_ARG, "vmbd_plan_out_shape: null argument");
  *sy0 = plan->g.sy0;
  *sy1 = plan->g.sy1;
  return VMBD_OK;
}

size_t vmbd_plan_spectrum_elems(const vmbd_plan* plan) {
  if (!plan) return 0;
  return (size_t)plan->g.Pf * plan->g.NB * plan->g.Th;
}

int vmbd_ola_forward(vmbd_plan* p, int32_t nframes, const float* x, const float* f, float* y,
                     vmbd_stream_t stream) {
  vmbd_clear_error();
  if (int r = check_plan(p, nframes)) return r;
  if (!x || !f || !y) return vmbd_fail(VMBD_E_ARG, "vmbd_ola_forward: null pointer");
  hipStream_t st = (hipStream_t)stream;
  const Geo g = call_geo(p, nframes);
  p->rows_frames = 0;
  const XRows xr = x_rows(p, g, x, false);
  const RowSrc sf = src_psf(p, f);
  VMBD_LAUNCH(launch_xf_rows(p, g, xr, sf, st));
  ColOut o1{p->Y1, g.ext0}, o2{nullptr, 0};
  ColIn none{nullptr, 0, 0, 0};
  VMBD_LAUNCH(launch_cols(g, COL_FWD, xr.in, raw_in(p->wsF, sf), none, o1, o2, nullptr, p->tw, st,
                          p->tune.col_threads));
  VMBD_LAUNCH(launch_rows_c2r_ola(g, dst_forward(p, y), p->tw, st, p->tune.ola_threads));
  p->rows_frames = nframes;
  return VMBD_OK;
}

int vmbd_ola_adjoint_x(vmbd_plan* p, int32_t nframes, const float* gres, const float* f, float* xbar,
                       int32_t exact, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (int r = check_plan(p, nframes)) return r;
  if (!gres || !f || !xbar) return vmbd_fail(VMBD_E_ARG, "vmbd_ola_adjoint_x: null pointer");
  hipStream_t st = (hipStream_t)stream;
  const Geo g = call_geo(p, nframes);
  p->rows_frames = 0;
  const SharedRows sr = residual_rows(p, g, gres, exact);
  const RowSrc sf = src_psf(p, f);
  VMBD_LAUNCH(launch_rows_r2c(sr.g, sr.s, p->wsG, p->tw, st));
  VMBD_LAUNCH(launch_rows_r2c(g, sf, p->wsF, p->tw, st, "rows_r2c_psf"));
  ColOut o1{p->Y1, g.sw0}, o2{nullptr, 0};
  ColIn none{};
  VMBD_LAUNCH(launch_cols(g, COL_ADJ1, shared_in(p->wsG, sr), raw_in(p->wsF, sf), none, o1, o2, nullptr, p->tw, st, p->tune.col_threads));
  VMBD_LAUNCH(launch_rows_c2r_ola(g, dst_adjoint_x(p, xbar), p->tw, st, p->tune.ola_threads));
  return VMBD_OK;
}

int vmbd_ola_adjoint_f(vmbd_plan* p, int32_t nframes, const float* gres, const float* x, float* fbar,
                       int32_t exact, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (int r = check_plan(p, nframes)) return r;
  if (!gres || !x || !fbar) return vmbd_fail(VMBD_E_ARG, "vmbd_ola_adjoint_f: null pointer");
  hipStream_t st = (hipStream_t)stream;
  const Geo g = call_geo(p, nframes);
  p->rows_frames = 0;
  const SharedRows sr = residual_rows(p, g, gres, exact);
  // the form of vmbd_ola_adjoint_xf, so that both give f-bar bit for bit
  const XRows xr = x_rows(p, g, x, false);
  VMBD_LAUNCH(launch_rows_r2c(sr.g, sr.s, p->wsG, p->tw, st));
  VMBD_LAUNCH(launch_rows_r2c(xr.g, xr.s, p->wsX, p->tw, st, "rows_r2c_x"));
  ColOut o1{p->Y2, g.sf0}, o2{nullptr, 0};
  ColIn none{};
  VMBD_LAUNCH(launch_cols(g, COL_ADJ1, shared_in(p->wsG, sr), xr.in, none, o1, o2, nullptr, p->tw, st, p->tune.col_threads));
  VMBD_LAUNCH(launch_rows_c2r_stack(g, p->Y2, g.sf0, fbar, g.sf0, g.sf1, inv_scale(g), p->tw, st));
  return VMBD_OK;
}

int vmbd_ola_adjoint_xf(vmbd_plan* p, int32_t nframes, const float* gres, const float* x, const float* f,
                        float* xbar, float* fbar, int32_t exact, int32_t flags, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (int r = check_plan(p, nframes)) return r;
  if (!gres || !xbar || !fbar) return vmbd_fail(VMBD_E_ARG, "vmbd_ola_adjoint_xf: null pointer");
  const bool reuse = (flags & VMBD_REUSE_ROWS) != 0;
  if (reuse && p->rows_frames != nframes)
    return vmbd_fail(VMBD_E_ARG, "vmbd_ola_adjoint_xf: VMBD_REUSE_ROWS without a preceding forward of %d frames",
                     nframes);
  if (!reuse && (!x || !f)) return vmbd_fail(VMBD_E_ARG, "vmbd_ola_adjoint_xf: null x/f");
  hipStream_t st = (hipStream_t)stream;
  const Geo g = call_geo(p, nframes);
  const SharedRows sr = residual_rows(p, g, gres, exact);
  const XRows xr = x_rows(p, g, x, false);
  const RowSrc sf = src_psf(p, f);
  VMBD_LAUNCH(launch_rows_r2c(sr.g, sr.s, p->wsG, p->tw, st));
  if (!reuse) {
    p->rows_frames = 0;
    VMBD_LAUNCH(launch_xf_rows(p, g, xr, sf, st));
    p->rows_frames = nframes;  // wsX/wsF now hold the rows of this x and f
  }
  ColOut o1{p->Y1, g.sw0}, o2{p->Y2, g.sf0};
  VMBD_LAUNCH(launch_cols(g, COL_ADJ2, shared_in(p->wsG, sr), raw_in(p->wsF, sf), xr.in, o1, o2, nullptr, p->tw, st,
                          p->tune.col_threads));
  VMBD_LAUNCH(launch_rows_c2r_ola(g, dst_adjoint_x(p, xbar), p->tw, st, p->tune.ola_threads));
  VMBD_LAUNCH(launch_rows_c2r_stack(g, p->Y2, g.sf0, fbar, g.sf0, g.sf1, inv_scale(g), p->tw, st));
  return VMBD_OK;
}

namespace {
// per-patch windows ws[k] = table[idx[k]] (the plan keeps the class table)
__global__ __launch_bounds__(256) void k_windows(float* ws, const float* table, const int* idx, int P, int n) {
  const long long total = (long long)P * n;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
    const long long k = e / n;
    ws[e] = table[(long long)idx[k] * n + (e - k * n)];
  }
}
}  // namespace

int vmbd_windows(const vmbd_plan* p, float* ws, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!p || !ws) return vmbd_fail(VMBD_E_ARG, "vmbd_windows: null pointer");
  const Geo& g = p->g;
  const int n = g.sw0 * g.sw1;
  const long long total = (long long)g.Pf * n;
  const unsigned blocks = (unsigned)std::min<long long>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(k_windows, dim3(blocks), dim3(256), 0, (hipStream_t)stream, ws, p->win, p->win_idx, g.Pf, n);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return vmbd_fail_hip("vmbd_windows", e);
  return VMBD_OK;
}

int vmbd_cache_psf_spectra(vmbd_plan* p, int32_t nframes, const float* f, void* spec, vmbd_stream_t stream) {
  return vmbd_spectra(p, nframes, VMBD_CACHE_PSF, f, spec, stream);
}

int vmbd_cache_image_spectra(vmbd_plan* p, int32_t nframes, const float* x, void* spec, vmbd_stream_t stream) {
  return vmbd_spectra(p, nframes, VMBD_CACHE_IMAGE, x, spec, stream);
}

int vmbd_spectra(vmbd_plan* p, int32_t nframes, int32_t kind, const float* z, void* spec, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (int r = check_plan(p, nframes)) return r;
  if (!z || !spec) return vmbd_fail(VMBD_E_ARG, "vmbd_spectra: null pointer");
  if (kind != VMBD_CACHE_PSF && kind != VMBD_CACHE_IMAGE) return vmbd_fail(VMBD_E_ARG, "vmbd_spectra: bad kind");
  hipStream_t st = (hipStream_t)stream;
  const Geo g = call_geo(p, nframes);
  p->rows_frames = 0;
  ColIn in;
  if (kind == VMBD_CACHE_PSF) {
    const RowSrc s = src_psf(p, z);
    VMBD_LAUNCH(launch_rows_r2c(g, s, p->wsF, p->tw, st, "rows_r2c_psf"));
    in = raw_in(p->wsF, s);
  } else {
    const XRows xr = x_rows(p, g, z, true);
    VMBD_LAUNCH(launch_rows_r2c(xr.g, xr.s, p->wsX, p->tw, st, "rows_r2c_x"));
    in = xr.in;
  }
  ColOut none_o{nullptr, 0};
  ColIn none{nullptr, 0, 0, 0};
  VMBD_LAUNCH(launch_cols(g, COL_SPEC, in, none, none, none_o, none_o, (float2*)spec, p->tw, st, p->tune.col_threads));
  return VMBD_OK;
}

int vmbd_ola_forward_cached_ex(vmbd_plan* p, int32_t nframes, int32_t kind, const void* spec, const float* u,
                               const float* sub, const float* mul, float* y, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (int r = check_plan(p, nframes)) return r;
  if (!spec || !u || !y) return vmbd_fail(VMBD_E_ARG, "vmbd_ola_forward_cached: null pointer");
  if (kind != VMBD_CACHE_PSF && kind != VMBD_CACHE_IMAGE)
    return vmbd_fail(VMBD_E_ARG, "vmbd_ola_forward_cached: bad kind");
  hipStream_t st = (hipStream_t)stream;
  const Geo g = call_geo(p, nframes);
  p->rows_frames = 0;
  // the free operand is the image for a PSF cache and the PSFs for an image cache
  ColIn in;
  if (kind == VMBD_CACHE_PSF) {
    const XRows xr = x_rows(p, g, u, true);
    VMBD_LAUNCH(launch_rows_r2c(xr.g, xr.s, p->wsX, p->tw, st, "rows_r2c_x"));
    in = xr.in;
  } else {
    const RowSrc s = src_psf(p, u);
    VMBD_LAUNCH(launch_rows_r2c(g, s, p->wsF, p->tw, st, "rows_r2c_psf"));
    in = raw_in(p->wsF, s);
  }
  ColOut o1{p->Y1, g.ext0}, o2{nullptr, 0};
  ColIn none{};
  VMBD_LAUNCH(launch_cols(g, COL_FWD, in, spec_in(spec, g.Th), none, o1, o2, nullptr, p->tw, st, p->tune.col_threads));
  OlaDst d = dst_forward(p, y);
  d.sub = sub;
  d.mul = mul;
  VMBD_LAUNCH(launch_rows_c2r_ola(g, d, p->tw, st, p->tune.ola_threads));
  return VMBD_OK;
}

int vmbd_ola_forward_cached(vmbd_plan* p, int32_t nframes, int32_t kind, const void* spec, const float* u, float* y,
                            vmbd_stream_t stream) {
  return vmbd_ola_forward_cached_ex(p, nframes, kind, spec, u, nullptr, nullptr, y, stream);
}

int vmbd_ola_adjoint_cached_ex(vmbd_plan* p, int32_t nframes, int32_t kind, const void* spec, const float* gres,
                               const float* z, float beta, float* out, int32_t exact, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (int r = check_plan(p, nframes)) return r;
  if (!spec || !gres || !out) return vmbd_fail(VMBD_E_ARG, "vmbd_ola_adjoint_cached: null pointer");
  if (kind != VMBD_CACHE_PSF && kind != VMBD_CACHE_IMAGE)
    return vmbd_fail(VMBD_E_ARG, "vmbd_ola_adjoint_cached: bad kind");
  hipStream_t st = (hipStream_t)stream;
  const Geo g = call_geo(p, nframes);
  const SharedRows sr = residual_rows(p, g, gres, exact);
  VMBD_LAUNCH(launch_rows_r2c(sr.g, sr.s, p->wsG, p->tw, st));
  ColIn none{};
  ColOut o2{nullptr, 0};
  if (kind == VMBD_CACHE_PSF) {
    ColOut o1{p->Y1, g.sw0};
    VMBD_LAUNCH(launch_cols(g, COL_ADJ1, shared_in(p->wsG, sr), spec_in(spec, g.Th), none, o1, o2, nullptr, p->tw, st, p->tune.col_threads));
    OlaDst d = dst_adjoint_x(p, out);
    d.axz = z;
    d.beta = beta;
    VMBD_LAUNCH(launch_rows_c2r_ola(g, d, p->tw, st, p->tune.ola_threads));
  } else {
    ColOut o1{p->Y2, g.sf0};
    VMBD_LAUNCH(launch_cols(g, COL_ADJ1, shared_in(p->wsG, sr), spec_in(spec, g.Th), none, o1, o2, nullptr, p->tw, st, p->tune.col_threads));
    VMBD_LAUNCH(launch_rows_c2r_stack(g, p->Y2, g.sf0, out, g.sf0, g.sf1, inv_scale(g), p->tw, st, z, beta));
  }
  return VMBD_OK;
}

int vmbd_ola_adjoint_cached(vmbd_plan* p, int32_t nframes, int32_t kind, const void* spec, const float* gres,
                            float* out, int32_t exact, vmbd_stream_t stream) {
  return vmbd_ola_adjoint_cached_ex(p, nframes, kind, spec, gres, nullptr, 0.f, out, exact, stream);
}

// ---------------------------------------------------------------------------
// Closed-form deconvolutions (olaGPU.py:416-733): row pass, Fourier division
// in the column pass, inverse rows into the output.  One frame.
// ---------------------------------------------------------------------------
namespace {

// rows of one deconvolution operand in the workspace: slot j of four
float2* work_rows(const vmbd_plan* p, void* work, int j) {
  const size_t per = (size_t)p->g.Pf * p->g.NB * row_pitch(p->g.sw0);
  return (float2*)work + (size_t)j * per;
}

// a patch of an (rows x cols) image placed at tile offset (off0, off1),
// rows / columns [off, off + sw) cut to the tile (chop_pad, gputools.py:364-401)
RowSrc src_chop(const vmbd_plan* p, const float* img, int rows, int cols, int off0, int off1) {
  const Geo& g = p->g;
  RowSrc s{};
  s.src = img;
  s.ld = cols;
  s.srcH = rows;
  s.srcW = cols;
  s.origin = 1;
  s.sh0 = off0;
  s.sh1 = off1;
  s.rlo = off0;
  s.rhi = std::min(off0 + g.sw0, g.Th);
  s.clo = off1;
  s.chi = std::min(off1 + g.sw1, g.Tw);
  if (s.rhi < s.rlo) s.rhi = s.rlo;
  if (s.chi < s.clo) s.chi = s.clo;
  return s;
}

// ... multiplied by the patch's window placed at tile offset (w0, w1); the
// rows and columns outside the window are cut
RowSrc windowed(const vmbd_plan* p, RowSrc s, int w0, int w1) {
  const Geo& g = p->g;
  s.win = p->win;
  s.win_idx = p->win_idx;
  s.wsh0 = w0;
  s.wsh1 = w1;
  s.rlo = std::max(s.rlo, w0);
  s.rhi = std::min(s.rhi, w0 + g.sw0);
  s.clo = std::max(s.clo, w1);
  s.chi = std::min(s.chi, w1 + g.sw1);
  if (s.rhi < s.rlo) s.rhi = s.rlo;
  if (s.chi < s.clo) s.chi = s.clo;
  return s;
}

// overlap-add of the patches' first sw rows and columns into an H x W image
// (crop_stack to sw, optional window, ola, crop to sx)
OlaDst dst_patches(const vmbd_plan* p, float* out, bool window) {
  OlaDst d = dst_adjoint_x(p, out);
  if (!window) {
    d.win = nullptr;
    d.win_idx = nullptr;
    d.win_slot = nullptr;
    d.row_class = nullptr;
    d.nslot = 0;
    d.wrow = d.wcol = d.colw = nullptr;
    d.zrow = nullptr;
    d.bcol[0] = d.bcol[1] = -1;
  }
  return d;
}

int check_deconv(const vmbd_plan* p, const char* what) {
  if (int r = check_plan(p, 1)) return r;
  if (p->g.Pf <= 0) return vmbd_fail(VMBD_E_ARG, "%s: empty plan", what);
  return VMBD_OK;
}

}  // namespace

size_t vmbd_deconv_workspace_bytes(const vmbd_plan* p) {
  if (!p) return 0;
  return 4 * (size_t)p->g.Pf * p->g.NB * row_pitch(p->g.sw0) * sizeof(float2);
}

int vmbd_deconv_direct(vmbd_plan* p, int32_t kind, const void* spec, const float* y, float alpha, float beta,
                       float* out, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (int r = check_deconv(p, "vmbd_deconv_direct")) return r;
  if (!spec || !y || !out) return vmbd_fail(VMBD_E_ARG, "vmbd_deconv_direct: null pointer");
  if (kind != VMBD_CACHE_PSF && kind != VMBD_CACHE_IMAGE) return vmbd_fail(VMBD_E_ARG, "vmbd_deconv_direct: bad kind");
  hipStream_t st = (hipStream_t)stream;
  const Geo g = call_geo(p, 1);
  p->rows_frames = 0;
  // the observation's patch at the mode offset times the window at tile
  // offset 0 (olaGPU.py:577-581)
  const RowSrc s = windowed(p, src_chop(p, y, g.sy0, g.sy1, g.oc0, g.oc1), 0, 0);
  VMBD_LAUNCH(launch_rows_r2c(g, s, p->wsX, p->tw, st));
  const ColIn in[2] = {raw_in(p->wsX, s), spec_in(spec, g.Th)};
  const DivParams prm{alpha, beta, 0.f};
  if (kind == VMBD_CACHE_PSF) {
    VMBD_LAUNCH(launch_cols_div(g, DIV_X, in, 2, prm, ColOut{p->Y1, g.sw0}, nullptr, p->tw, st));
    VMBD_LAUNCH(launch_rows_c2r_ola(g, dst_patches(p, out, false), p->tw, st));
  } else {
    VMBD_LAUNCH(launch_cols_div(g, DIV_X, in, 2, prm, ColOut{p->Y2, g.sf0}, nullptr, p->tw, st));
    VMBD_LAUNCH(launch_rows_c2r_stack(g, p->Y2, g.sf0, out, g.sf0, g.sf1, inv_scale(g), p->tw, st));
  }
  return VMBD_OK;
}

int vmbd_deconv_gdirect(vmbd_plan* p, const float* xx, const float* xy, const float* yx, const float* yy, float alpha,
                        float beta, float* fout, void* work, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (int r = check_deconv(p, "vmbd_deconv_gdirect")) return r;
  if (!xx || !xy || !yx || !yy || !fout || !work) return vmbd_fail(VMBD_E_ARG, "vmbd_deconv_gdirect: null pointer");
  hipStream_t st = (hipStream_t)stream;
  const Geo g = call_geo(p, 1);
  p->rows_frames = 0;
  // gradients of the latent image at offset 0, of the observation at the
  // mode offset (olaGPU.py:480-500)
  const RowSrc srcs[4] = {src_chop(p, xx, g.H, g.W - 1, 0, 0), src_chop(p, yx, g.sy0, g.sy1 - 1, g.oc0, g.oc1),
                          src_chop(p, xy, g.H - 1, g.W, 0, 0), src_chop(p, yy, g.sy0 - 1, g.sy1, g.oc0, g.oc1)};
  float2* outs[4] = {work_rows(p, work, 0), work_rows(p, work, 1), work_rows(p, work, 2), work_rows(p, work, 3)};
  VMBD_LAUNCH(launch_rows_r2c_multi(g, srcs, outs, 4, p->tw, st));
  const ColIn in[4] = {raw_in(outs[0], srcs[0]), raw_in(outs[1], srcs[1]), raw_in(outs[2], srcs[2]),
                       raw_in(outs[3], srcs[3])};
  VMBD_LAUNCH(launch_cols_div(g, DIV_G, in, 4, DivParams{alpha, beta, 0.f}, ColOut{p->Y2, g.sf0}, nullptr, p->tw,
                              st));
  VMBD_LAUNCH(launch_rows_c2r_stack(g, p->Y2, g.sf0, fout, g.sf0, g.sf1, inv_scale(g), p->tw, st));
  return VMBD_OK;
}

int vmbd_deconv_sparse_setup(vmbd_plan* p, const void* spec, const float* y, const float* c, float alpha, float gamma,
                             void* ftpy, float* cfac, void* work, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (int r = check_deconv(p, "vmbd_deconv_sparse_setup")) return r;
  if (!spec || !y || !c || !ftpy || !cfac || !work)
    return vmbd_fail(VMBD_E_ARG, "vmbd_deconv_sparse_setup: null pointer");
  hipStream_t st = (hipStream_t)stream;
  const Geo g = call_geo(p, 1);
  p->rows_frames = 0;
  // observation and constant image, patches and windows both at the mode
  // offset (olaGPU.py:655-672)
  const RowSrc srcs[2] = {windowed(p, src_chop(p, y, g.sy0, g.sy1, g.oc0, g.oc1), g.oc0, g.oc1),
                          windowed(p, src_chop(p, c, g.sy0, g.sy1, g.oc0, g.oc1), g.oc0, g.oc1)};
  float2* outs[2] = {work_rows(p, work, 0), work_rows(p, work, 1)};
  VMBD_LAUNCH(launch_rows_r2c_multi(g, srcs, outs, 2, p->tw, st));
  // F^H y, kept as a spectrum for every half-quadratic step (olaGPU.py:690-692)
  const ColIn yin[2] = {raw_in(outs[0], srcs[0]), spec_in(spec, g.Th)};
  VMBD_LAUNCH(launch_cols_div(g, DIV_SPROD, yin, 2, DivParams{0.f, 0.f, 0.f}, ColOut{nullptr, 0}, (float2*)ftpy,
                              p->tw, st));
  // deconvolved constant image (olaGPU.py:674-688): the filter energy of the
  // sqrt(2)/2-scaled differences is L1/2, hence alpha/2 (exact)
  const ColIn cin[2] = {raw_in(outs[1], srcs[1]), spec_in(spec, g.Th)};
  VMBD_LAUNCH(launch_cols_div(g, DIV_X, cin, 2, DivParams{0.5f * alpha, gamma, 0.f}, ColOut{p->Y1, g.sw0}, nullptr,
                              p->tw, st));
  OlaDst d = dst_patches(p, cfac, true);
  d.clip_lo = 1;
  d.lo = 0.01f;
  VMBD_LAUNCH(launch_rows_c2r_ola(g, d, p->tw, st, p->tune.ola_threads));
  // cfac = 0.01 / max(const, 0.01)
  return vmbd_ew(cfac, cfac, nullptr, 0.01f, (int64_t)g.H * g.W, VMBD_EW_SDIV, stream);
}

int vmbd_deconv_sparse_step(vmbd_plan* p, const void* spec, const void* ftpy, const float* wx, int32_t wx_rows,
                            int32_t wx_cols, const float* wy, int32_t wy_rows, int32_t wy_cols, float alpha,
                            float beta, float gamma, const float* cfac, float* x, void* work, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (int r = check_deconv(p, "vmbd_deconv_sparse_step")) return r;
  if (!spec || !ftpy || !wx || !wy || !cfac || !x || !work)
    return vmbd_fail(VMBD_E_ARG, "vmbd_deconv_sparse_step: null pointer");
  if (wx_rows <= 0 || wx_cols <= 0 || wy_rows <= 0 || wy_cols <= 0 || !(alpha * beta > 0.f))
    return vmbd_fail(VMBD_E_ARG, "vmbd_deconv_sparse_step: bad size or alpha*beta <= 0");
  hipStream_t st = (hipStream_t)stream;
  const Geo g = call_geo(p, 1);
  p->rows_frames = 0;
  // shrunk gradients at offset 0 (olaGPU.py:700-712)
  const RowSrc srcs[2] = {src_chop(p, wx, wx_rows, wx_cols, 0, 0), src_chop(p, wy, wy_rows, wy_cols, 0, 0)};
  float2* outs[2] = {work_rows(p, work, 0), work_rows(p, work, 1)};
  VMBD_LAUNCH(launch_rows_r2c_multi(g, srcs, outs, 2, p->tw, st));
  const ColIn in[4] = {raw_in(outs[0], srcs[0]), raw_in(outs[1], srcs[1]), spec_in(ftpy, g.Th), spec_in(spec, g.Th)};
  VMBD_LAUNCH(launch_cols_div(g, DIV_S, in, 4, DivParams{alpha, beta, gamma}, ColOut{p->Y1, g.sw0}, nullptr, p->tw,
                              st));
  // windows, overlap-add, times the constant-image factor, clipped at 0
  // (olaGPU.py:714-730)
  OlaDst d = dst_patches(p, x, true);
  d.mul = cfac;
  d.clip_lo = 1;
  d.lo = 0.f;
  VMBD_LAUNCH(launch_rows_c2r_ola(g, d, p->tw, st, p->tune.ola_threads));
  return VMBD_OK;
}

// ---------------------------------------------------------------------------
// Batches of plans: the passes of every group in single launches
// ---------------------------------------------------------------------------
}  // extern "C"

struct vmbd_batch {
  int device = -1;
  std::vector<vmbd_plan*> plans;
  BatchGroup* dev = nullptr;
  struct Launch {
    int pass, tile, wgs;
    size_t shm;
    BatchLaunch L;
  };
  std::vector<Launch> launches;  // in pass order
};

namespace {

// the step's parameters of one group: vmbd_ola_forward followed by
// vmbd_ola_adjoint_xf reusing the forward's row transforms
BatchGroup batch_group(vmbd_plan* p, int nframes, const float* x, const float* f, const float* gres, float* y,
                       float* xbar, float* fbar, int exact) {
  BatchGroup B{};
  const Geo g = call_geo(p, nframes);
  B.g = g;
  B.tw = p->tw;
  const XRows xr = x_rows(p, g, x, false);
  const RowSrc sf = src_psf(p, f);
  const int TPW = WG / (g.Tw / 16);
  auto blocks = [&](const RowSrc& s) { return ((s.rhi - s.rlo + 1) / 2 + TPW - 1) / TPW; };
  if (xr.shared) {  // the image rows in a pass of their own (their geometry)
    B.rows_fwd.n = 1;
    B.rows_fwd.s[0] = sf;
    B.rows_fwd.out[0] = p->wsF;
    B.rows_fwd.blk_end[0] = blocks(sf);
    B.gx = xr.g;
    B.rows_x.n = 1;
    B.rows_x.s[0] = xr.s;
    B.rows_x.out[0] = p->wsX;
    B.rows_x.blk_end[0] = blocks(xr.s);
  } else {
    B.rows_fwd.n = 2;
    B.rows_fwd.s[0] = xr.s;
    B.rows_fwd.s[1] = sf;
    B.rows_fwd.out[0] = p->wsX;
    B.rows_fwd.out[1] = p->wsF;
    B.rows_fwd.blk_end[0] = blocks(xr.s);
    B.rows_fwd.blk_end[1] = B.rows_fwd.blk_end[0] + blocks(sf);
  }
  B.cf_a = xr.in;
  B.cf_b = raw_in(p->wsF, sf);
  B.cf_o = ColOut{p->Y1, g.ext0};
  B.ola_fwd = dst_forward(p, y);
  ola_fit_rows(B.ola_fwd, (long long)B.ola_fwd.O0 * nframes);
  const SharedRows sr = residual_rows(p, g, gres, exact);
  B.gres = sr.g;
  B.rows_res.n = 1;
  B.rows_res.s[0] = sr.s;
  B.rows_res.out[0] = p->wsG;
  B.rows_res.blk_end[0] = blocks(sr.s);
  B.ca_a = shared_in(p->wsG, sr);
  B.ca_b = raw_in(p->wsF, sf);
  B.ca_c = xr.in;
  B.ca_o1 = ColOut{p->Y1, g.sw0};
  B.ca_o2 = ColOut{p->Y2, g.sf0};
  B.ola_adj = dst_adjoint_x(p, xbar);
  ola_fit_rows(B.ola_adj, (long long)B.ola_adj.O0 * nframes);
  // the grouped OLA launch runs WG threads per workgroup for every group
  B.st_src = p->Y2;
  B.st_srow = g.sf0;
  B.st_nrows = g.sf0;
  B.st_ncols = g.sf1;
  B.st_out = fbar;
  B.st_scale = inv_scale(g);
  return B;
}

void batch_free(vmbd_batch* b) {
  if (!b) return;
  (void)hipFree(b->dev);
  delete b;
}

}  // namespace

extern "C" {

vmbd_batch* vmbd_batch_create(vmbd_plan* const* plans, int32_t n, const int32_t* nframes, const float* const* x,
                              const float* const* f, const float* const* g, float* const* y, float* const* xbar,
                              float* const* fbar, int32_t exact) {
  vmbd_clear_error();
  if (!plans || !nframes || !x || !f || !g || !y || !xbar || !fbar || n <= 0) {
    vmbd_fail(VMBD_E_ARG, "vmbd_batch_create: null argument or n <= 0");
    return nullptr;
  }
  if (n > BATCH_MAX_GROUPS) {
    vmbd_fail(VMBD_E_UNSUP, "vmbd_batch_create: %d plans, at most %d", n, BATCH_MAX_GROUPS);
    return nullptr;
  }
  std::vector<BatchGroup> host((size_t)n);
  int dev = -1;
  (void)hipGetDevice(&dev);
  for (int i = 0; i < n; ++i) {
    if (check_plan(plans[i], nframes[i])) return nullptr;
    if (!x[i] || !f[i] || !g[i] || !y[i] || !xbar[i] || !fbar[i]) {
      vmbd_fail(VMBD_E_ARG, "vmbd_batch_create: null buffer of plan %d", i);
      return nullptr;
    }
    host[i] = batch_group(plans[i], nframes[i], x[i], f[i], g[i], y[i], xbar[i], fbar[i], exact);
  }
  vmbd_batch* b = new vmbd_batch();
  b->device = dev;
  b->plans.assign(plans, plans + n);
  // per pass, one launch per distinct tile; groups in plan order
  for (int pass = 0; pass < BP_N; ++pass) {
    std::vector<int> tiles;
    for (int i = 0; i < n; ++i) {
      const int t = batch_tile(host[i], pass);
      if (std::find(tiles.begin(), tiles.end(), t) == tiles.end()) tiles.push_back(t);
    }
    for (int t : tiles) {
      vmbd_batch::Launch l{};
      l.pass = pass;
      l.tile = t;
      int blk = 0;
      for (int i = 0; i < n; ++i) {
        if (batch_tile(host[i], pass) != t || batch_blocks(host[i], pass) == 0) continue;
        l.L.gidx[l.L.ng] = i;
        l.L.start[l.L.ng] = blk;
        ++l.L.ng;
        blk += (batch_blocks(host[i], pass) + 7) & ~7;
        l.shm = std::max(l.shm, batch_shared_bytes(host[i], pass));
      }
      l.L.start[l.L.ng] = blk;
      if (l.L.ng == 0) continue;  // no group of this tile has work in the pass
      l.wgs = (pass == BP_OLA_FWD || pass == BP_OLA_ADJ) ? batch_ola_threads(host.data(), l.L, pass) : WG;
      b->launches.push_back(l);
    }
  }
  size_t tot = 0;
  if (!dev_alloc(&b->dev, (size_t)n, tot)) {
    batch_free(b);
    vmbd_fail(VMBD_E_NOMEM, "vmbd_batch_create: hipMalloc failed");
    return nullptr;
  }
  hipError_t e = hipMemcpy(b->dev, host.data(), (size_t)n * sizeof(BatchGroup), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    batch_free(b);
    vmbd_fail_hip("vmbd_batch_create: hipMemcpy", e);
    return nullptr;
  }
  return b;
}

void vmbd_batch_destroy(vmbd_batch* b) { batch_free(b); }

int vmbd_batch_launches(const vmbd_batch* b) { return b ? (int)b->launches.size() : 0; }

int vmbd_batch_step(vmbd_batch* b, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!b) return vmbd_fail(VMBD_E_ARG, "vmbd_batch_step: null batch");
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != b->device)
    return vmbd_fail(VMBD_E_ARG, "vmbd_batch_step: batch belongs to device %d, current device is %d", b->device, dev);
  hipStream_t st = (hipStream_t)stream;
  for (vmbd_plan* p : b->plans) p->rows_frames = 0;
  for (const auto& l : b->launches) VMBD_LAUNCH(launch_batch(b->dev, l.L, l.pass, l.tile, l.wgs, l.shm, st));
  return VMBD_OK;
}

}  // extern "C"
