// Shared declarations of the EFF/OLA kernels (geometry + launch helpers).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace vmbd {

// Patch-grid geometry of one operator instance, passed by value to kernels.
//   image x: H x W;  PSF stack f: P x sf0 x sf1;  patch k = m*csf1 + n at
//   origin (m*hop0, n*hop1);  FFT tile Th x Tw;  half spectrum NB = Tw/2+1 slots
//   (the real bins 0 and Tw/2 share slot 0 between the passes).
//   ext = sw + sf - 1 (support of one patch's linear convolution);
//   output y: sy0 x sy1 = crop of the OLA buffer at offset (oc0, oc1).
struct Geo {
  int H, W;
  int sf0, sf1;
  int csf0, csf1;
  int Pf;  // patches per frame (csf0*csf1)
  int P;   // patches in this launch (frames * Pf); patch p = frame*Pf + k
  int sw0, sw1;
  int hop0, hop1;
  int Th, Tw, NB;
  int NBP;  // row pitch (bins) of the column-pass outputs: NB rounded up to 16 (128-byte lines)
  int ext0, ext1;
  int oc0, oc1;
  int sy0, sy1;
};

// Source of the rows fed to the row transform (chop / pad / window fused).
// Tile element (r, c) of patch p reads src[p*pstride + (o0+r)*ld + (o1+c)]
// with (o0, o1) = (origin ? (m*hop0, n*hop1) : (0,0)) - (sh0, sh1), zero when
// (r, c) is outside [rlo,rhi) x [clo,chi) or the source index is outside
// [0,srcH) x [0,srcW); multiplied by win[win_idx[p]][r - wsh0][c - wsh1]
// when win != null (the window placed at tile offset (wsh0, wsh1); the
// caller keeps [rlo,rhi) x [clo,chi) inside it).
struct RowSrc {
  const float* src;
  long long pstride;  // elements between consecutive patches (PSF stacks)
  long long fstride;  // elements between consecutive frames (images)
  int ld, srcH, srcW;
  int origin;
  int sh0, sh1;
  int rlo, rhi, clo, chi;
  const float* win;
  const int* win_idx;
  int wsh0, wsh1;
  // instead of win (when set): per-patch-column weights of one row type,
  // cw[(k*2 + czr[i])*sw1 + c] for the source row i of tile row r (the image
  // rows shared by a patch column, separable windows)
  const float* cw;
  const int* czr;
};

// Up to four row-transform jobs fused in one launch: blocks [0, blk_end[0])
// run job 0, [blk_end[0], blk_end[1]) job 1, ...
constexpr int ROW_JOBS_MAX = 4;
struct RowJobs {
  RowSrc s[ROW_JOBS_MAX];
  float2* out[ROW_JOBS_MAX];
  int blk_end[ROW_JOBS_MAX];
  int n;
};

// Row-pass output: [P][NB][row_pitch(nr)] holding tile rows [rlo, rlo+nr);
// the pitch rounds nr up to 16 rows so every bin column starts on a 128-byte
// line and a workgroup's 32 rows fill whole lines.
__host__ __device__ inline int row_pitch(int nr) { return (nr + 15) & ~15; }

// Column-pass input: either raw row-transform output (spec == 0, pitch =
// row_pitch(nr)), or a cached spectrum laid out [P][NB][Th] (spec == 1,
// pitch = Th).  shared == 1: the row transforms are shared by the patches of
// one patch column, laid out [frame][csf1][NB][pitch] over virtual rows
// v = m*hop0 + b (b in [rlo, rlo+nr)); patch (m, n) reads group
// (frame, n) from row m*hop0.
struct ColIn {
  const float2* p;
  int rlo, nr, spec, pitch;
  int shared;
  // shared image rows (separable windows): tile row r of patch (m, n) is
  // rsc[m*nr + r] times its patch column's row; where the window is at the
  // floor in a column covered by one patch column (ecol[e] >= 0: image
  // column, en: its patch column, ec: its tile column) the row takes
  // ex[frame][m*hop0 + r][ecol] * einv[m*hop0 + r] there (a constant times
  // e^{-2 pi i k ec / Tw} in bin k), which the load adds
  const float* rsc;
  const float* ex;
  long long exf;
  int exld, exH;
  const float* einv;
  int ecol[2], en[2], ec[2];
};

// Where a column-pass workgroup is: patch row and column, frame, the hop,
// the row bins, the twiddle table (the shared image rows' corrections)
struct ColCtx {
  int pm, pn, pfr, hop0, Tw;
  const float2* tw;
};

// Column-pass output rows [0, nrows) in layout [P][nrows][NBP].
struct ColOut {
  float2* p;
  int nrows;
};

// Inverse row transform + overlap-add gather + crop into an output image.
// Output pixel (i, j) sums, over every patch k whose local extent covers the
// OLA coordinate (i+oc0, j+oc1), the spatial value at local (r, c) =
// (i+oc0-m*hop0, j+oc1-n*hop1) taken from src rows [P][srow][NBP]; the value is
// multiplied by win[win_idx[k]][r][c] when win != null.
struct OlaDst {
  float* out;
  int O0, O1;
  int oc0, oc1;
  int ext0, ext1;
  const float2* src;
  int srow;
  const float* win;
  const int* win_idx;
  // windows staged in LDS per output row: patch k = m*csf1 + n uses slot
  // win_slot[k] of its patch row, slot s of patch row m holds class
  // row_class[m*nslot + s] (-1: unused); nslot distinct windows per patch row
  const int* win_slot;
  const int* row_class;
  int nslot;
  int mc;     // max number of patch rows covering one output row
  int alias;   // 1: all transforms of a row fit one round; gather buffer aliases the FFT slices
  int stream;  // 1: several rounds, each round's rows (aliasing the FFT slices) gathered into
               //    per-thread column accumulators; needs O1 <= ola_stream_cols()
  // > 0 (single-round launches): the gather buffer holds, per covering patch
  // row, ncls rows of ola_class_len() floats in overlap-add coordinates;
  // patch n writes its row at column n*hop1 of row n mod ncls, so one
  // output column reads the same column of every row (ola_layout)
  int ncls;
  // > 0 (no window, class layout): the covering patch rows are summed in
  // the row-frequency domain before one inverse transform per patch-column
  // pair, sumr output rows per 256-thread workgroup (ola_layout)
  int sumr;
  // separable windows of the summed form (null: unit weights): row weight
  // wrow[r], column weight wcol[c] of patch-local (r, c); per output column
  // j the normalisers colw[j] = 1 / sum of the covering b(c) and
  // colw[O1 + j] = 1 / number of covering patch columns; zrow[i] != 0: every
  // covering window is at the floor on row i (unit weights); output columns
  // bcol[e] (-1: none, OLA coordinates) are covered by patch column bn[e]
  // alone, at local column bc[e], where its window is at the floor: their
  // values in the other rows are the unweighted sums, evaluated apart
  const float* wrow;
  const float* wcol;
  const float* colw;
  const int* zrow;
  int bcol[2], bn[2], bc[2];
  // 1/(Th*Tw) of the unnormalised transforms, applied to the overlap-add sum
  float scale;
  // fused solver arithmetic on the stored value v = scale * sum (each optional):
  // v = ((v - sub[o]) * mul[o]) + beta * axz[o], o = output index
  const float* sub;
  const float* mul;
  const float* axz;
  float beta;
  // then v = max(v, lo) when clip_lo != 0
  int clip_lo;
  float lo;
};

enum ColOp { COL_SPEC = 0, COL_FWD = 1, COL_ADJ1 = 2, COL_ADJ2 = 3 };

// Fourier divisions of the closed-form deconvolutions (olaGPU.py:416-733,
// gputools.cu:316-406), on the half spectra of the column pass, per
// (row bin c, column frequency k) with wx = 2 pi c / Tw, wy = 2 pi k / Th:
//   DIV_X  Z = conj(B) A / (|B|^2 + alpha L1 + beta)              (xdirect)
//          L1 = 4 sin^2(wx/2) + 4 sin^2(wy/2): the [-1 1] and [-1 1]' filters
//   DIV_G  Z = (conj(B1) A1 + conj(B2) A2) / (|B1|^2 + |B2|^2 + alpha L2 + beta)
//          L2 = (1+2cos wy)^2 4 sin^2(wx/2) + 4 sin^2(wy/2) (1+2cos wx)^2  (gdirect)
//   SPROD  S = A conj(B), written as a spectrum                   (sparse, F^H y)
//   DIV_S  Z = (conj(Gx) A1 + conj(Gy) A2 + S1/(alpha beta)) / (|S2|^2/(alpha beta) + Ls + gamma)
//          Gx = r(e^{-i wx} - 1), Gy = r(e^{-i wy} - 1), r = sqrt(2)/2, Ls = |Gx|^2 + |Gy|^2  (sparse)
// A* are row-pass outputs (transformed here), B / S* cached spectra.
enum DivOp { DIV_X = 0, DIV_G = 1, DIV_SPROD = 2, DIV_S = 3 };
struct DivParams {
  float alpha, beta, gamma;
};

// Launchers (ola_kernels.hip).  Return hipError_t of the launch.
// Jobs of different patch counts in one launch (the shared image rows of a
// patch column beside the PSF rows of every patch): job j has P[j] units of
// its own geometry's patches (Pf[j] per frame); at most two jobs.
struct RowJobGeo {
  int P[2], Pf[2];
  int nb[2];    // row blocks per patch
  int uend[2];  // cumulative work units, P[j] * nb[j]
};
hipError_t launch_rows_r2c_geo(const Geo& g, const RowSrc* s, float2* const* out, const int* P, const int* Pf, int n,
                               const float2* tw, hipStream_t st, const char* name);
// name: the launch's label in the library's kernel timing (default rows_r2c)
hipError_t launch_rows_r2c(const Geo& g, const RowSrc& s, float2* out, const float2* tw, hipStream_t st,
                           const char* name = nullptr);
// several row-transform jobs in one launch (n <= ROW_JOBS_MAX)
hipError_t launch_rows_r2c_multi(const Geo& g, const RowSrc* s, float2* const* out, int n, const float2* tw,
                                 hipStream_t st);
// wgc: threads per column-pass workgroup (128 or 256; 0 = 256): the column
// strip of a workgroup is wgc / (Th/16) columns wide
hipError_t launch_cols(const Geo& g, int op, const ColIn& a, const ColIn& b, const ColIn& c, const ColOut& o1,
                       const ColOut& o2, float2* spec_out, const float2* tw, hipStream_t st, int wgc = 0);
// wgs: threads per OLA workgroup (128, 192 or 256; 0 = ola_wg's choice);
// 128 and 192 need every output row in one round (ola_fits)
hipError_t launch_rows_c2r_ola(const Geo& g, const OlaDst& d, const float2* tw, hipStream_t st, int wgs = 0);
// a single-round launch of wgs threads holds every output row's transforms
int ola_fits(const Geo& g, const OlaDst& d, int wgs);
// Column pass of a Fourier division (deconv_kernels.hip): operands a[0..na)
// (row-pass outputs or spectra, in the order of DivOp's formula: DIV_X /
// SPROD: A, B; DIV_G: B1, A1, B2, A2; DIV_S: A1, A2, S1, S2); result rows
// [0, o.nrows) to o, or (SPROD) the spectrum to spec_out.
hipError_t launch_cols_div(const Geo& g, int op, const ColIn* a, int na, DivParams prm, const ColOut& o,
                           float2* spec_out, const float2* tw, hipStream_t st);
// out = scale * inverse row transforms (+ beta * axz when axz != null)
hipError_t launch_rows_c2r_stack(const Geo& g, const float2* src, int srow, float* out, int nrows, int ncols,
                                 float scale, const float2* tw, hipStream_t st, const float* axz = nullptr, float beta = 0.f);
// Batches of several plans (frames of different geometry) whose passes run
// as single launches over a table of groups (one group = one plan's frames):
// everything the seven passes of a step (forward + both exact adjoints) need
// for one group, built once when the batch binds its buffers.
constexpr int BATCH_MAX_GROUPS = 16;
struct BatchGroup {
  Geo g;     // call geometry (P = frames * Pf)
  Geo gres;  // geometry of the shared residual rows
  Geo gx;    // geometry of the image rows shared per patch column (rows_x.n > 0)
  RowJobs rows_fwd, rows_res, rows_x;
  ColIn cf_a, cf_b;
  ColOut cf_o;
  ColIn ca_a, ca_b, ca_c;
  ColOut ca_o1, ca_o2;
  OlaDst ola_fwd, ola_adj;
  const float2* st_src;
  int st_srow, st_nrows, st_ncols;
  float* st_out;
  float st_scale;
  const float2* tw;
};
enum BatchPass { BP_ROWS_FWD = 0, BP_ROWS_X, BP_COLS_FWD, BP_OLA_FWD, BP_ROWS_RES, BP_COLS_ADJ, BP_OLA_ADJ, BP_STACK, BP_N };
// One launch of a pass over ng groups of the same tile (Tw for the row, OLA
// and stack passes, Th for the column passes): launch group gi is
// groups[gidx[gi]] and owns blocks [start[gi], start[gi+1]); every start is a
// multiple of 8, so the XCD a block lands on is the same as in a launch of
// its group alone.
struct BatchLaunch {
  int ng;
  int gidx[BATCH_MAX_GROUPS];
  int start[BATCH_MAX_GROUPS + 1];
};
// blocks the group needs in the pass (before rounding up to 8)
int batch_blocks(const BatchGroup& G, int pass);
// the pass's tile (template parameter) of the group
int batch_tile(const BatchGroup& G, int pass);
// LDS bytes of the pass's launch for the group (dynamic, OLA passes only)
size_t batch_shared_bytes(const BatchGroup& G, int pass);
// threads per workgroup of an OLA pass over the launch's groups: 128 or 192
// when every group's row fits one round of that many threads, else 256
int batch_ola_threads(const BatchGroup* host_groups, const BatchLaunch& L, int pass);
hipError_t launch_batch(const BatchGroup* dev_groups, const BatchLaunch& L, int pass, int tile, int wgs, size_t shm,
                        hipStream_t st);

size_t ola_rowbuf_bytes(const Geo& g, const OlaDst& d);
// floats per class row of the gather buffer: the overlap-add extent of a
// patch row, (csf1-1)*hop1 + ext1, rounded up to 4
__host__ __device__ inline int ola_class_len(const Geo& g, const OlaDst& d) {
  return ((g.csf1 - 1) * g.hop1 + d.ext1 + 3) & ~3;
}
// sets d.alias, d.stream and d.ncls of a destination whose other fields are set
void ola_layout(const Geo& g, OlaDst& d);
// fewer summed-form rows per workgroup for a launch of `rows` output rows
// that would not fill the GPU several times over
void ola_fit_rows(OlaDst& d, long long rows);
size_t ola_shared_bytes(const Geo& g, const OlaDst& d);
int ola_single_round(const Geo& g, const OlaDst& d);
// multi-round launches whose output rows fit the per-thread accumulators
int ola_stream_rounds(const Geo& g, const OlaDst& d);
// threads per workgroup of an OLA launch of `rows` output rows (all frames)
int ola_wg(const Geo& g, const OlaDst& d, long long rows);
// Staged x-adjoint window rows: thread t of a transform (TPT = Tw/16 threads)
// weights columns t + TPT*mm; they are kept at floats t*MW + mm of a row of
// TPT*MW floats, MW = ceil(sw1/TPT) rounded up to a multiple of 4.
__host__ __device__ inline int ola_win_mw(const Geo& g) {
  const int TPT = g.Tw / 16;
  return (((g.sw1 + TPT - 1) / TPT) + 3) & ~3;
}
__host__ __device__ inline int ola_win_rw(const Geo& g) { return (g.Tw / 16) * ola_win_mw(g); }

}  // namespace vmbd
