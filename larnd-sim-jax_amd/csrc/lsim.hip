// C ABI of the MI355X LUT simulation chain (declarations and conventions in include/lsim.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "backward.h"
#include "chop.h"
#include "loss.h"
#include "param.h"
#include "probfee.h"
#include "sdtw.h"

using namespace lsim;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess) return fail(LSIM_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Bump allocator over a caller buffer (256-byte aligned sub-buffers).
struct Carve {
  char* base;
  size_t off = 0;
  explicit Carve(void* b) : base((char*)b) {}
  template <typename T>
  T* take(int64_t n) {
    off = (off + 255) & ~(size_t)255;
    T* p = base ? (T*)(base + off) : nullptr;
    off += (size_t)(n > 0 ? n : 1) * sizeof(T);
    return p;
  }
};

}  // namespace

struct lsim_ctx {
  lsim_geometry geo;
  lsim_config cfg;
  int halo = 1;
  int cm = 5;
  // LUT tables
  bool has_lut = false;
  int lx = 0, ly = 0, nt = 0;
  float3* TM3 = nullptr;
  float* TM = nullptr;
  float* TN = nullptr;
  float* TNP = nullptr;
  float2* DM = nullptr;
  float2* DN = nullptr;
  float2* DNS = nullptr;  // DN regrouped per (tick, sub-bin x, sub-bin y): one segment's neighbour entries contiguous
  double* CTM = nullptr;
  double* CTN = nullptr;
};

namespace {

int make_dev(const lsim_ctx* c, const lsim_params* p, int n_events, Dev& d) {
  std::memset(&d, 0, sizeof(d));
  const lsim_geometry& g = c->geo;
  const lsim_config& k = c->cfg;
  d.ntpc = g.n_tpc;
  d.nx = g.n_pixels_x;
  d.ny = g.n_pixels_y;
  d.nsb = k.nb_sampling_bins_per_pixel;
  d.nb = k.nb_tran_diff_bins;
  d.nbsym = (d.nb - 1) / 2;
  d.nn = k.number_pix_neighbors;
  d.ne = d.nn > 0 ? d.nn : 1;
  d.halo = c->halo;
  d.gx = d.nx + 2 * d.halo;
  d.gy = d.ny + 2 * d.halo;
  d.n_events = n_events;
  d.n_cells = (int64_t)n_events * d.ntpc * d.gx * d.gy;
  d.pitch = g.pixel_pitch;
  d.w = g.pixel_pitch / d.nsb;
  std::memcpy(d.tpc, g.tpc_borders, sizeof(d.tpc));
  d.L = k.signal_length;
  d.nt = c->nt;
  d.nticks = k.n_ticks;
  d.nout = k.n_ticks - 1;
  d.lx = c->lx;
  d.ly = c->ly;
  d.cm = c->cm;
  d.nm = k.n_templates;
  d.ext = k.long_diff_extent;
  d.H = k.max_adc_values;
  d.interval = k.adc_interval;
  for (int m = 0; m < LSIM_MAX_TEMPLATES; ++m) d.sig[m] = k.long_diff_template[m];
  if (p) {
    d.eField = p->eField;
    d.Ab = p->Ab;
    d.kb = p->kb;
    d.lifetime = p->lifetime;
    d.long_diff = p->long_diff;
    d.tran_diff = p->tran_diff;
    d.rho = p->lArDensity;
    d.alpha = p->alpha;
    d.beta = p->beta;
    d.R_param = p->R_param;
    d.mev = p->MeVToElectrons;
    d.v = p->vdrift_eff;
    d.dv = p->dvdrift_deField;
    d.shift[0] = p->shift_x;
    d.shift[1] = p->shift_y;
    d.shift[2] = p->shift_z;
    d.margin = p->size_margin;
    d.ts = p->t_sampling;
    d.thr = p->threshold;
    d.gain = p->gain;
    d.vped = p->v_pedestal;
    d.vcm = p->v_cm;
    d.vref = p->v_ref;
    d.adcc = p->adc_counts;
    d.recomb = p->recombination_mode;
    if (d.recomb < 1 || d.recomb > 3) return fail(LSIM_E_ARG, "recombination_mode must be 1 (BOX), 2 (BIRKS) or 3 (ELLIPSOID)");
    if (!(d.v > 0)) return fail(LSIM_E_ARG, "vdrift_eff must be > 0");
    if (!(d.ts > 0)) return fail(LSIM_E_ARG, "t_sampling must be > 0");
    // segments inside the margin may sit up to floor(margin / pitch) + 1 pixels off the plane
    const int need = 1 + (int)std::floor(p->size_margin / g.pixel_pitch);
    if (need > c->halo) return fail(LSIM_E_ARG, "size_margin too large for the cell halo");
  }
  if (n_events < 0) return fail(LSIM_E_ARG, "n_events must be >= 0");
  if (d.n_cells >= (int64_t)0x7fffffff) return fail(LSIM_E_ARG, "too many events for 32-bit cell keys");
  return LSIM_OK;
}

SegIn make_segin(const lsim_segments* s) {
  SegIn in;
  in.x = s->x;
  in.y = s->y;
  in.z = s->z;
  in.dE = s->dE;
  in.dEdx = s->dEdx;
  in.ev = s->event_id;
  in.zs = s->z_start;
  in.ze = s->z_end;
  in.dx = s->dx;
  in.stride = s->stride;
  in.n = s->n;
  return in;
}

int check_segments(const lsim_segments* s, int recomb) {
  if (!s) return fail(LSIM_E_ARG, "segments is NULL");
  if (s->n < 0 || s->n >= (int64_t)0x7fffffff) return fail(LSIM_E_ARG, "segment count out of range");
  if (s->n > 0 && (!s->x || !s->y || !s->z || !s->dE || !s->dEdx || !s->event_id))
    return fail(LSIM_E_ARG, "segment column pointer is NULL");
  if (recomb == 3 && s->n > 0 && (!s->z_start || !s->z_end || !s->dx))
    return fail(LSIM_E_ARG, "ELLIPSOID recombination needs z_start, z_end and dx");
  if (s->stride < 1) return fail(LSIM_E_ARG, "segment stride must be >= 1");
  return LSIM_OK;
}

// Workspace layout shared by the forward phases and the backward.
struct Ws {
  uint32_t *k0, *v0, *k1, *v1;  // sort ping-pong
  void* sort_scratch;
  SegRec* rec;
  PairRec* prs;  // [n] window-store records (sorted order)
  int2* rng;
  int* cell_row;
  int* flag;
  int* ridx;
  int* row_cell;
  int* n_rows;
  int* hit_cnt;
  int* hit_off;
  int* scan_tmp;
  int* n_hits32;
  int* row_work;  // [n] estimated work of each pixel row in processing order (XCD band cuts)
  int* band;      // [kXcd + 1] first processing slot of each XCD's band
  uint32_t *tk0, *tv0, *tk1, *tv1;  // [n + 1] tile sort ping-pong; the processing order lands in tv0 or tv1
  float* dtot;  // [(2 nn + 1)^2][n] boundary corrections of every (segment, row offset) pair
  size_t bytes;
};

Ws layout(void* base, int64_t n, int64_t n_cells) {
  Carve c(base);
  Ws w;
  w.k0 = c.take<uint32_t>(n);
  w.v0 = c.take<uint32_t>(n);
  w.k1 = c.take<uint32_t>(n);
  w.v1 = c.take<uint32_t>(n);
  w.sort_scratch = c.take<char>((int64_t)radix_sort_scratch(n + 1));  // segments, then rows (<= max(n, 1))
  w.rec = c.take<SegRec>(n);
  w.prs = c.take<PairRec>(n);
  w.rng = c.take<int2>(n_cells);
  w.cell_row = c.take<int>(n_cells);
  w.flag = c.take<int>(n);
  w.ridx = c.take<int>(n);
  w.row_cell = c.take<int>(n);
  w.n_rows = c.take<int>(4);
  w.hit_cnt = c.take<int>(n);
  w.hit_off = c.take<int>(n);
  w.scan_tmp = c.take<int>(scan_tiles(n) + 4);
  w.n_hits32 = c.take<int>(4);
  w.row_work = c.take<int>(n);
  w.band = c.take<int>(kXcd + 1);
  w.tk0 = c.take<uint32_t>(n + 1);
  w.tv0 = c.take<uint32_t>(n + 1);
  w.tk1 = c.take<uint32_t>(n + 1);
  w.tv1 = c.take<uint32_t>(n + 1);
  w.dtot = nullptr;
  w.bytes = c.off + 256;
  return w;
}

// The pair corrections ((2 nn + 1)^2 floats per segment), carved after the rest.
void layout_corrections(Ws& w, void* base, int64_t n, int nn) {
  Carve c(base);
  c.off = w.bytes;
  w.dtot = c.take<float>((int64_t)(2 * nn + 1) * (2 * nn + 1) * n);
  w.bytes = c.off + 256;
}

int bits_for(int64_t v) {
  int b = 0;
  while (b < 32 && ((int64_t)1 << b) <= v) ++b;
  return b;
}

// Segments are sorted in two stable stages, by window key (sort_bits_window) then by cell key
// (bits_for(n_cells)); the result lands in (k0, v0) or (k1, v1) depending on the total pass count,
// a pure function of nout and n_cells, so every phase agrees without storing anything.
int sort_bits_window(const Dev& d) { return bits_for((int64_t)d.nout + 2); }

void sorted_buffers(const Ws& w, const Dev& d, uint32_t*& keys, uint32_t*& vals) {
  const int passes = (sort_bits_window(d) + 7) / 8 + (bits_for(d.n_cells) + 7) / 8;
  keys = (passes & 1) ? w.k1 : w.k0;
  vals = (passes & 1) ? w.v1 : w.v0;
}

// Fixed cost of one pixel row (set-up, row sum, cumulative charge, self-trigger) in units of
// segment-row pairs, for the work estimate the XCD bands are cut on (lsim_set_row_cost).
int g_row_cost = 256;

// Rows each XCD's band may hold: 1.5 times an equal share of max_rows (the grid of the row kernel
// is kXcd times this, so at most a third of its workgroups find no row).
int64_t band_slots(int64_t max_rows) { return (3 * max_rows + 2 * kXcd - 1) / (2 * kXcd); }

// Side of the square cell tiles of the rows' processing order (lsim_set_row_tile).
int g_row_tile = 16;

// Bits of the tile-order sort key: every tile id and the key of the rows past P.
int tile_key_bits(const Dev& d) { return bits_for(row_tiles(d, g_row_tile) + 1); }

// The processing order of the rows (sorted by tile: k_tile_keys + radix sort); the buffer it lands
// in depends only on the key width, like sorted_buffers.
const int* row_order(const Ws& w, const Dev& d) {
  return (const int*)((((tile_key_bits(d) + 7) / 8) & 1) ? w.tv1 : w.tv0);
}

// Processing order, work estimate of every row and the XCD band cuts (k_tile_keys + sort,
// k_row_work, k_bands) before the row kernel; also zeroes the hit counts of the rows past P.
int launch_row_bands(const Dev& d, const Ws& w, int64_t max_rows, hipStream_t st) {
  const int bits = tile_key_bits(d);
  const unsigned g = (unsigned)((max_rows + 255) / 256);
  hipLaunchKernelGGL(k_tile_keys, dim3(g), dim3(256), 0, st, d, g_row_tile, w.row_cell, w.n_rows, max_rows,
                     (uint32_t)(((uint64_t)1 << bits) - 1), w.tk0, w.tv0);
  hipError_t err;
  radix_sort_pairs(w.tk0, w.tv0, w.tk1, w.tv1, max_rows, bits, w.sort_scratch, st, &err);
  HIP_TRY(err);
  hipLaunchKernelGGL(k_row_work, dim3(g), dim3(256), 0, st, d, w.rng, w.row_cell, row_order(w, d), w.n_rows, max_rows,
                     g_row_cost, w.row_work, w.hit_cnt);
  hipLaunchKernelGGL(k_bands, dim3(1), dim3(kBandThreads), 0, st, w.row_work, w.n_rows, max_rows, band_slots(max_rows),
                     w.band);
  HIP_TRY(hipGetLastError());
  return LSIM_OK;
}

template <int NB, int WAVES>
void launch_rows_nb(const Dev& d, const Tables& tb, const Ws& w, int64_t n_seg, int64_t max_rows, const RowOut& ro,
                    hipStream_t st) {
  const size_t lds = rows_lds_bytes(WAVES, d.nout, d.L, (2 * d.ne + 1) * (2 * d.ne + 1));
  const int64_t grid = kXcd * band_slots(max_rows);
  hipLaunchKernelGGL((k_rows_fwd<NB, WAVES>), dim3((unsigned)grid), dim3(WAVES * kWave), lds, st, d, tb, w.rec, w.rng,
                     w.row_cell, w.n_rows, max_rows, w.band, row_order(w, d), w.dtot, w.prs, n_seg, ro);
}

template <int WAVES>
void launch_rows(const Dev& d, const Tables& tb, const Ws& w, int64_t n_seg, int64_t max_rows, const RowOut& ro,
                 hipStream_t st) {
  if (d.nb == 5) launch_rows_nb<5, WAVES>(d, tb, w, n_seg, max_rows, ro, st);
  else if (d.nb == 3) launch_rows_nb<3, WAVES>(d, tb, w, n_seg, max_rows, ro, st);
  else launch_rows_nb<1, WAVES>(d, tb, w, n_seg, max_rows, ro, st);
}

int g_row_waves = 4;

// Optional timing of the waveform kernel: hipEvents recorded on the launch stream around every
// k_rows_fwd launch while enabled (not for use under graph capture).
struct RowTimer {
  bool on = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  size_t used = 0;
} g_timer;

void timer_begin(hipStream_t st, hipEvent_t* stop_out) {
  *stop_out = nullptr;
  if (!g_timer.on) return;
  if (g_timer.used == g_timer.ev.size()) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
    g_timer.ev.emplace_back(a, b);
  }
  auto& pr = g_timer.ev[g_timer.used++];
  (void)hipEventRecord(pr.first, st);
  *stop_out = pr.second;
}

void free_tables(lsim_ctx* c) {
  (void)hipFree(c->TM3);
  (void)hipFree(c->TM);
  (void)hipFree(c->TN);
  (void)hipFree(c->TNP);
  (void)hipFree(c->DM);
  (void)hipFree(c->DN);
  (void)hipFree(c->DNS);
  (void)hipFree(c->CTM);
  (void)hipFree(c->CTN);
  c->TM = c->TN = c->TNP = nullptr;
  c->TM3 = nullptr;
  c->DM = c->DN = c->DNS = nullptr;
  c->CTM = c->CTN = nullptr;
  c->has_lut = false;
}

}  // namespace

extern "C" {

const char* lsim_last_error(void) { return g_err.c_str(); }

int lsim_version(void) { return 1; }

size_t lsim_tape_bytes(void) { return sizeof(FeeTape); }

int lsim_set_row_waves(int waves) {
  if (waves != 1 && waves != 2 && waves != 4 && waves != 8) return fail(LSIM_E_ARG, "waves must be 1, 2, 4 or 8");
  g_row_waves = waves;
  return LSIM_OK;
}

int lsim_set_row_cost(int pairs) {
  if (pairs < 1 || pairs > (1 << 20)) return fail(LSIM_E_ARG, "row cost must be in [1, 2^20] pairs");
  g_row_cost = pairs;
  return LSIM_OK;
}

int lsim_set_row_tile(int cells) {
  if (cells < 1 || cells > 4096) return fail(LSIM_E_ARG, "row tile must be in [1, 4096] cells");
  g_row_tile = cells;
  return LSIM_OK;
}

int lsim_timing(int enable) {
  g_timer.on = enable != 0;
  g_timer.used = 0;
  return LSIM_OK;
}

int lsim_timing_read(double* total_ms, int64_t* launches) {
  if (!total_ms || !launches) return fail(LSIM_E_ARG, "NULL argument");
  double tot = 0.0;
  for (size_t i = 0; i < g_timer.used; ++i) {
    HIP_TRY(hipEventSynchronize(g_timer.ev[i].second));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, g_timer.ev[i].first, g_timer.ev[i].second));
    tot += ms;
  }
  *total_ms = tot;
  *launches = (int64_t)g_timer.used;
  return LSIM_OK;
}

int lsim_ctx_create(lsim_ctx** out, const lsim_geometry* geo, const lsim_config* cfg) {
  if (!out || !geo || !cfg) return fail(LSIM_E_ARG, "NULL argument");
  *out = nullptr;
  if (geo->n_tpc < 1 || geo->n_tpc > LSIM_MAX_TPC) return fail(LSIM_E_ARG, "n_tpc out of range");
  if (geo->n_pixels_x < 1 || geo->n_pixels_y < 1) return fail(LSIM_E_ARG, "pixel counts must be >= 1");
  if (!(geo->pixel_pitch > 0)) return fail(LSIM_E_ARG, "pixel_pitch must be > 0");
  const int nb = cfg->nb_tran_diff_bins, nsb = cfg->nb_sampling_bins_per_pixel;
  if (nb != 1 && nb != 3 && nb != 5) return fail(LSIM_E_ARG, "nb_tran_diff_bins must be 1, 3 or 5");
  if (nsb < 2 || (nsb & 1)) return fail(LSIM_E_ARG, "nb_sampling_bins_per_pixel must be even and >= 2");
  if ((nb - 1) / 2 >= nsb) return fail(LSIM_E_ARG, "diffusion bins wider than a pixel are not supported");
  if (cfg->number_pix_neighbors < 0 || cfg->number_pix_neighbors > 8) return fail(LSIM_E_ARG, "number_pix_neighbors out of range");
  if (cfg->signal_length < 1 || cfg->signal_length > 3 * kChunk - 2)
    return fail(LSIM_E_ARG, "signal_length out of range (1..187)");
  if (cfg->n_ticks < 4 || cfg->n_ticks - 1 > kMaxOut) return fail(LSIM_E_ARG, "n_ticks out of range (<= 2049)");
  if (cfg->max_adc_values < 1 || cfg->max_adc_values > kMaxH) return fail(LSIM_E_ARG, "max_adc_values out of range");
  if (cfg->n_templates < 3 || cfg->n_templates > LSIM_MAX_TEMPLATES) return fail(LSIM_E_ARG, "n_templates out of range");
  if (cfg->long_diff_extent < 0 || 2 * cfg->long_diff_extent + 1 > 128) return fail(LSIM_E_ARG, "long_diff_extent out of range");
  for (int m = 0; m < cfg->n_templates; ++m)
    if (!(cfg->long_diff_template[m] > 0) || (m && !(cfg->long_diff_template[m] > cfg->long_diff_template[m - 1])))
      return fail(LSIM_E_ARG, "long_diff_template must be positive and strictly increasing");
  lsim_ctx* c = new lsim_ctx();
  c->geo = *geo;
  c->cfg = *cfg;
  c->halo = 2;
  c->cm = nsb / 2;
  *out = c;
  return LSIM_OK;
}

int lsim_ctx_destroy(lsim_ctx* c) {
  if (!c) return LSIM_OK;
  free_tables(c);
  delete c;
  return LSIM_OK;
}

int lsim_lut_set(lsim_ctx* c, const float* raw, int32_t nx, int32_t ny, int32_t nt, void* stream) {
  if (!c || !raw) return fail(LSIM_E_ARG, "NULL argument");
  hipStream_t st = (hipStream_t)stream;
  const lsim_config& k = c->cfg;
  if (nt < k.signal_length + 1 || nt > kMaxOut) return fail(LSIM_E_ARG, "LUT time axis out of range");
  const int nsb = k.nb_sampling_bins_per_pixel, n = k.number_pix_neighbors > 0 ? k.number_pix_neighbors : 1;
  const int maxc = nsb * n + nsb / 2 - 1;  // largest |offset| index a neighbour (or a diffusion bin one pixel over) reads
  if (nx <= maxc || ny <= maxc || nx < c->cm || ny < c->cm)
    return fail(LSIM_E_ARG, "LUT too small for number_pix_neighbors (needs > " + std::to_string(maxc) + " bins per axis)");
  if (c->has_lut && (nx != c->lx || ny != c->ly || nt != c->nt)) free_tables(c);
  c->lx = nx;
  c->ly = ny;
  c->nt = nt;
  const int64_t ms = (int64_t)k.n_templates * c->cm * c->cm, ns = (int64_t)nx * ny;
  const int64_t nds = (int64_t)nsb * nsb * (2 * n + 1) * (2 * n + 1);  // DNS entries per tick
  const int L = k.signal_length;
  Dev d;
  int rc = make_dev(c, nullptr, 1, d);
  if (rc) return rc;
  if (!c->has_lut) {
    // all tables or none: a failed allocation frees the ones made before it (free_tables resets
    // every pointer, and hipFree of a null pointer is a no-op)
    const std::pair<void**, size_t> req[] = {
        {(void**)&c->TM3, ms * kWinPad * sizeof(float3)}, {(void**)&c->TM, ms * L * sizeof(float)},
        {(void**)&c->TN, ns * L * sizeof(float)},         {(void**)&c->TNP, ns * kWinPad * sizeof(float)},
        {(void**)&c->DM, ms * nt * sizeof(float2)},       {(void**)&c->DN, ns * nt * sizeof(float2)},
        {(void**)&c->DNS, nds * nt * sizeof(float2)},
        {(void**)&c->CTM, ms * (L + 1) * sizeof(double)}, {(void**)&c->CTN, ns * (L + 1) * sizeof(double)}};
    for (const auto& q : req) {
      const hipError_t e = hipMalloc(q.first, q.second);
      if (e != hipSuccess) {
        *q.first = nullptr;
        free_tables(c);
        return fail(LSIM_E_HIP, std::string("hipMalloc of the LUT tables: ") + hipGetErrorString(e));
      }
    }
  }
  hipLaunchKernelGGL(k_lut_tables, dim3((unsigned)ms), dim3(256), 0, st, d, raw, 1, c->TM, (float*)nullptr, c->DM, c->CTM);
  hipLaunchKernelGGL(k_lut_tables, dim3((unsigned)ns), dim3(256), 0, st, d, raw, 0, c->TN, c->TNP, c->DN, c->CTN);
  hipLaunchKernelGGL(k_lut_tm3, dim3(1024), dim3(256), 0, st, d, c->TM, c->TM3);
  hipLaunchKernelGGL(k_lut_dns, dim3(4096), dim3(256), 0, st, d, c->DN, c->DNS);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    free_tables(c);  // no half-built tables stay behind (and none leak on the next call)
    return fail(LSIM_E_HIP, std::string("LUT table kernels: ") + hipGetErrorString(e));
  }
  c->has_lut = true;
  return LSIM_OK;
}

int lsim_forward_query(lsim_ctx* c, int64_t n_seg, int32_t n_events, size_t* ws_bytes, int64_t* max_rows_bound) {
  if (!c || !ws_bytes) return fail(LSIM_E_ARG, "NULL argument");
  Dev d;
  int rc = make_dev(c, nullptr, n_events, d);
  if (rc) return rc;
  Ws w = layout(nullptr, n_seg, d.n_cells);
  layout_corrections(w, nullptr, n_seg, d.ne);
  *ws_bytes = w.bytes;
  if (max_rows_bound) {
    const int64_t plane_pixels = (int64_t)n_events * d.ntpc * d.nx * d.ny;
    *max_rows_bound = n_seg < plane_pixels ? n_seg : plane_pixels;
  }
  return LSIM_OK;
}

int lsim_forward_prepare(lsim_ctx* c, const lsim_params* p, const lsim_segments* seg, int32_t n_events, void* ws,
                         size_t ws_bytes, int64_t* n_rows_dev, int32_t* status_dev, void* stream) {
  if (!c || !p || !ws) return fail(LSIM_E_ARG, "NULL argument");
  if (!c->has_lut) return fail(LSIM_E_STATE, "lsim_lut_set has not been called");
  hipStream_t st = (hipStream_t)stream;
  Dev d;
  int rc = make_dev(c, p, n_events, d);
  if (rc) return rc;
  rc = check_segments(seg, d.recomb);
  if (rc) return rc;
  const int64_t n = seg->n;
  Ws w = layout(ws, n, d.n_cells);
  layout_corrections(w, ws, n, d.ne);
  if (w.bytes > ws_bytes) return fail(LSIM_E_SPACE, "workspace too small: need " + std::to_string(w.bytes) + " bytes");
  SegIn in = make_segin(seg);
  int* status = status_dev;

  fill<int2>(w.rng, d.n_cells, make_int2(0, 0), st);
  fill<int>(w.cell_row, d.n_cells, -1, st);
  fill<int>(w.n_rows, 1, 0, st);
  if (n > 0) {
    const unsigned g = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_keys, dim3(g), dim3(256), 0, st, d, in, w.k0, w.v0, status);
    hipError_t err;
    const int cur = radix_sort_pairs(w.k0, w.v0, w.k1, w.v1, n, sort_bits_window(d), w.sort_scratch, st, &err);
    HIP_TRY(err);
    uint32_t* ka = cur ? w.k1 : w.k0;
    uint32_t* va = cur ? w.v1 : w.v0;
    uint32_t* kb = cur ? w.k0 : w.k1;
    uint32_t* vb = cur ? w.v0 : w.v1;
    hipLaunchKernelGGL(k_cell_keys, dim3(g), dim3(256), 0, st, d, in, va, n, ka);
    radix_sort_pairs(ka, va, kb, vb, n, bits_for(d.n_cells), w.sort_scratch, st, &err);
    HIP_TRY(err);
    uint32_t *keys, *vals;
    sorted_buffers(w, d, keys, vals);
    if (d.nb == 5) {
      hipLaunchKernelGGL(k_records<5>, dim3(g), dim3(kRecThreads), 0, st, d, in, keys, vals, n, c->DNS, c->DM, w.rec, w.prs, w.dtot);
    } else if (d.nb == 3) {
      hipLaunchKernelGGL(k_records<3>, dim3(g), dim3(kRecThreads), 0, st, d, in, keys, vals, n, c->DNS, c->DM, w.rec, w.prs, w.dtot);
    } else {
      hipLaunchKernelGGL(k_records<1>, dim3(g), dim3(kRecThreads), 0, st, d, in, keys, vals, n, c->DNS, c->DM, w.rec, w.prs, w.dtot);
    }
    hipLaunchKernelGGL(k_cells, dim3(g), dim3(256), 0, st, d, keys, n, w.rng, w.flag);
    HIP_TRY((exclusive_scan<int, int>(w.flag, w.ridx, n, w.scan_tmp, w.n_rows, st)));
    hipLaunchKernelGGL(k_rows_map, dim3(g), dim3(256), 0, st, d, keys, w.flag, w.ridx, n, w.row_cell, w.cell_row);
  }
  hipLaunchKernelGGL(k_copy_count, dim3(1), dim3(1), 0, st, w.n_rows, n_rows_dev, (int64_t)-1, (int*)nullptr);
  HIP_TRY(hipGetLastError());
  return LSIM_OK;
}

static int forward_rows_impl(lsim_ctx* c, const Dev& d, const Ws& w, int64_t n_seg, int64_t max_rows,
                             const lsim_fwd_out* out, hipStream_t st) {
  Tables tb{c->TM3, c->TM, c->TN, c->TNP, c->DM, c->DN, c->CTM, c->CTN};
  RowOut ro;
  ro.wfs = out->wfs;
  ro.adc = out->adc;
  ro.ticks = out->ticks;
  ro.tape = (FeeTape*)out->tape;
  ro.hit_cnt = w.hit_cnt;
  ro.pixel_ids = out->pixel_ids;
  hipLaunchKernelGGL(k_copy_count, dim3(1), dim3(1), 0, st, w.n_rows, out->n_rows, max_rows, out->status);
  if (max_rows > 0) {
    const int rc = launch_row_bands(d, w, max_rows, st);
    if (rc) return rc;
    hipEvent_t stop_ev;
    timer_begin(st, &stop_ev);
    switch (g_row_waves) {
      case 1: launch_rows<1>(d, tb, w, n_seg, max_rows, ro, st); break;
      case 2: launch_rows<2>(d, tb, w, n_seg, max_rows, ro, st); break;
      case 8: launch_rows<8>(d, tb, w, n_seg, max_rows, ro, st); break;
      default: launch_rows<4>(d, tb, w, n_seg, max_rows, ro, st); break;
    }
    HIP_TRY(hipGetLastError());
    if (stop_ev) HIP_TRY(hipEventRecord(stop_ev, st));
    HIP_TRY((exclusive_scan<int, int>(w.hit_cnt, w.hit_off, max_rows, w.scan_tmp, w.n_hits32, st)));
    if (out->hit_adc) {
      HitOut h{out->hit_adc, out->hit_x, out->hit_y, out->hit_z, out->hit_tick, out->hit_prob, out->hit_event,
               out->hit_pixel, out->hit_index};
      const int64_t tot = max_rows * d.H;
      hipLaunchKernelGGL(k_hits, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, d, out->pixel_ids, out->adc,
                         out->ticks, w.hit_off, out->n_rows, max_rows, (int64_t)0, h);
    }
  } else {
    fill<int>(w.n_hits32, 1, 0, st);
  }
  if (out->n_hits) hipLaunchKernelGGL(k_copy_count, dim3(1), dim3(1), 0, st, w.n_hits32, out->n_hits, (int64_t)-1, (int*)nullptr);
  HIP_TRY(hipGetLastError());
  return LSIM_OK;
}

static int check_out(const lsim_fwd_out* out) {
  if (!out) return fail(LSIM_E_ARG, "out is NULL");
  if (!out->n_rows || !out->pixel_ids || !out->adc || !out->ticks || !out->status)
    return fail(LSIM_E_ARG, "out: n_rows, pixel_ids, adc, ticks and status are required");
  if (out->hit_adc && (!out->hit_x || !out->hit_y || !out->hit_z || !out->hit_tick || !out->hit_prob || !out->hit_event ||
                       !out->hit_pixel || !out->hit_index))
    return fail(LSIM_E_ARG, "out: hit arrays must be all set or all NULL");
  return LSIM_OK;
}

int lsim_forward_rows(lsim_ctx* c, const lsim_params* p, int64_t n_seg, int32_t n_events, void* ws, size_t ws_bytes,
                      int64_t max_rows, const lsim_fwd_out* out, void* stream) {
  if (!c || !p || !ws) return fail(LSIM_E_ARG, "NULL argument");
  if (!c->has_lut) return fail(LSIM_E_STATE, "lsim_lut_set has not been called");
  int rc = check_out(out);
  if (rc) return rc;
  Dev d;
  rc = make_dev(c, p, n_events, d);
  if (rc) return rc;
  Ws w = layout(ws, n_seg, d.n_cells);
  layout_corrections(w, ws, n_seg, d.ne);
  if (w.bytes > ws_bytes) return fail(LSIM_E_SPACE, "workspace too small");
  // rows never exceed segments; the per-row workspace buffers hold max(n, 1) entries
  if (max_rows > (n_seg > 1 ? n_seg : 1)) max_rows = n_seg > 1 ? n_seg : 1;
  return forward_rows_impl(c, d, w, n_seg, max_rows, out, (hipStream_t)stream);
}

int lsim_forward(lsim_ctx* c, const lsim_params* p, const lsim_segments* seg, int32_t n_events, void* ws, size_t ws_bytes,
                 int64_t max_rows, const lsim_fwd_out* out, void* stream) {
  int rc = check_out(out);
  if (rc) return rc;
  if (!seg) return fail(LSIM_E_ARG, "segments is NULL");
  fill<int>(out->status, 1, 0, (hipStream_t)stream);
  rc = lsim_forward_prepare(c, p, seg, n_events, ws, ws_bytes, nullptr, out->status, stream);
  if (rc) return rc;
  return lsim_forward_rows(c, p, seg->n, n_events, ws, ws_bytes, max_rows, out, stream);
}

int lsim_negative_rows_query(lsim_ctx* c, size_t* ws_bytes, int64_t* max_ids) {
  if (!c || !ws_bytes || !max_ids) return fail(LSIM_E_ARG, "NULL argument");
  const int64_t span = (int64_t)c->geo.n_tpc * c->geo.n_pixels_x * c->geo.n_pixels_y;
  *ws_bytes = (size_t)((span + 31) >> 5) * sizeof(uint32_t);
  *max_ids = span;
  return LSIM_OK;
}

int lsim_negative_rows(lsim_ctx* c, const lsim_params* p, const lsim_segments* seg, void* ws, size_t ws_bytes,
                       int64_t* ids, int64_t max_ids, int64_t* n_ids, void* stream) {
  if (!c || !p || !ws || !n_ids || (max_ids > 0 && !ids)) return fail(LSIM_E_ARG, "NULL argument");
  if (max_ids < 0) return fail(LSIM_E_ARG, "max_ids must be >= 0");
  hipStream_t st = (hipStream_t)stream;
  Dev d;
  int rc = make_dev(c, p, 1, d);
  if (rc) return rc;
  rc = check_segments(seg, d.recomb);
  if (rc) return rc;
  const int64_t span = (int64_t)d.ntpc * d.nx * d.ny;
  const int64_t words = (span + 31) >> 5;
  if ((size_t)words * sizeof(uint32_t) > ws_bytes) return fail(LSIM_E_SPACE, "workspace too small");
  uint32_t* bits = (uint32_t*)ws;
  fill<uint32_t>(bits, words, 0u, st);
  if (seg->n > 0) {
    SegIn in = make_segin(seg);
    hipLaunchKernelGGL(k_neg_mark, dim3((unsigned)((seg->n + 255) / 256)), dim3(256), 0, st, d, in, bits);
  }
  hipLaunchKernelGGL(k_pad_list, dim3(1), dim3(kPadThreads), 0, st, bits, span, ids, max_ids, n_ids);
  HIP_TRY(hipGetLastError());
  return LSIM_OK;
}

int lsim_fee_query(lsim_ctx* c, int64_t n_rows, size_t* bytes) {
  if (!c || !bytes) return fail(LSIM_E_ARG, "NULL argument");
  Carve cv(nullptr);
  cv.take<int>(n_rows);
  cv.take<int>(n_rows);
  cv.take<int>(scan_tiles(n_rows) + 4);
  cv.take<int>(4);
  *bytes = cv.off + 256;
  return LSIM_OK;
}

}  // extern "C"

// The self-trigger of caller waveforms over a workspace of lsim_fee_query(n_rows) bytes at `ws`
// (a caller buffer, or the file-scope scratch of lsim_fee_forward_noalloc).  Every workspace
// value is written before it is read, so the workspace needs no clearing.
static int fee_forward_impl(lsim_ctx* c, const lsim_params* p, const float* wfs, const int64_t* pixel_ids,
                            int64_t n_rows, void* ws, size_t ws_bytes, const lsim_fwd_out* out, hipStream_t st) {
  if (!c || !p || !out || !out->adc || !out->ticks || !ws) return fail(LSIM_E_ARG, "NULL argument");
  if (n_rows < 0) return fail(LSIM_E_ARG, "n_rows < 0");
  if (n_rows > 0 && (!wfs || !pixel_ids)) return fail(LSIM_E_ARG, "wfs / pixel_ids is NULL");
  Dev d;
  int rc = make_dev(c, p, 1, d);
  if (rc) return rc;
  size_t need;
  lsim_fee_query(c, n_rows, &need);
  if (need > ws_bytes) return fail(LSIM_E_SPACE, "fee workspace too small");
  Carve cv(ws);
  int* cnt = cv.take<int>(n_rows);
  int* off = cv.take<int>(n_rows);
  int* tmp = cv.take<int>(scan_tiles(n_rows) + 4);
  int* tot = cv.take<int>(4);
  RowOut ro;
  ro.wfs = nullptr;
  ro.adc = out->adc;
  ro.ticks = out->ticks;
  ro.tape = (FeeTape*)out->tape;
  ro.hit_cnt = cnt;
  ro.pixel_ids = nullptr;
  if (n_rows > 0) {
    hipLaunchKernelGGL((k_fee_rows<256>), dim3((unsigned)n_rows), dim3(256), 0, st, d, wfs, pixel_ids, n_rows, ro);
    HIP_TRY((exclusive_scan<int, int>(cnt, off, n_rows, tmp, tot, st)));
    if (out->hit_adc) {
      HitOut h{out->hit_adc, out->hit_x, out->hit_y, out->hit_z, out->hit_tick, out->hit_prob, out->hit_event,
               out->hit_pixel, out->hit_index};
      const int64_t t = n_rows * d.H;
      hipLaunchKernelGGL(k_hits, dim3((unsigned)((t + 255) / 256)), dim3(256), 0, st, d, pixel_ids, out->adc, out->ticks, off,
                         (const int64_t*)nullptr, n_rows, n_rows, h);
    }
  } else {
    fill<int>(tot, 1, 0, st);
  }
  if (out->n_hits) hipLaunchKernelGGL(k_copy_count, dim3(1), dim3(1), 0, st, tot, out->n_hits, (int64_t)-1, (int*)nullptr);
  HIP_TRY(hipGetLastError());
  return LSIM_OK;
}

extern "C" {

int lsim_fee_forward(lsim_ctx* c, const lsim_params* p, const float* wfs, const int64_t* pixel_ids, int64_t n_rows,
                     void* ws, size_t ws_bytes, const lsim_fwd_out* out, void* stream) {
  return fee_forward_impl(c, p, wfs, pixel_ids, n_rows, ws, ws_bytes, out, (hipStream_t)stream);
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// lsim_fee_forward_noalloc: the workspace of lsim_fee_forward in a file-scope device array
// ------------------------------------------------------------------------------------------
namespace {

// lsim_fee_query's layout for LSIM_NOALLOC_MAX_ROWS rows: hit counts and offsets (one int per
// row each), the scan's tile sums and its total, each run 256-byte aligned.
constexpr int64_t kNoAllocRows = LSIM_NOALLOC_MAX_ROWS;
constexpr size_t kNoAllocBytes =
    (size_t)2 * kNoAllocRows * 4 + (size_t)((kNoAllocRows + kScanTile - 1) / kScanTile + 4) * 4 + 16 + 5 * 256;

// Each GPU the library runs on holds its own copy until the process ends.  Host code never
// passes this name to a kernel or a copy: it is a host-side stand-in, and the device address
// comes from hipGetSymbolAddress (scratch_on_current_device).
__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device.
int scratch_on_current_device(void** out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(LSIM_E_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(LSIM_E_HIP, "hipGetSymbolAddress returned NULL for the scratch array");
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return LSIM_OK;
}

}  // namespace

extern "C" {

int lsim_fee_forward_noalloc(lsim_ctx* c, const lsim_params* p, const float* wfs, const int64_t* pixel_ids,
                             int64_t n_rows, const lsim_fwd_out* out, void* stream) {
  if (!c) return fail(LSIM_E_ARG, "NULL argument");
  if (n_rows > kNoAllocRows)
    return fail(LSIM_E_SPACE, "n_rows " + std::to_string(n_rows) + " exceeds LSIM_NOALLOC_MAX_ROWS (" +
                                  std::to_string(kNoAllocRows) + ")");
  size_t need = 0;
  int rc = lsim_fee_query(c, n_rows, &need);
  if (rc) return rc;
  if (need > kNoAllocBytes) return fail(LSIM_E_SPACE, "scratch array smaller than the workspace layout");
  void* ws = nullptr;
  rc = scratch_on_current_device(&ws);
  if (rc) return rc;
  return fee_forward_impl(c, p, wfs, pixel_ids, n_rows, ws, kNoAllocBytes, out, (hipStream_t)stream);
}

int lsim_segment_records(lsim_ctx* c, int64_t n_seg, int32_t n_events, void* ws, size_t ws_bytes, void* rec_out,
                         uint32_t* keys_out, uint32_t* order_out, void* stream) {
  if (!c || !ws || !rec_out || !keys_out || !order_out) return fail(LSIM_E_ARG, "NULL argument");
  hipStream_t st = (hipStream_t)stream;
  Dev d;
  int rc = make_dev(c, nullptr, n_events, d);
  if (rc) return rc;
  Ws w = layout(ws, n_seg, d.n_cells);
  if (w.bytes > ws_bytes) return fail(LSIM_E_SPACE, "workspace too small");
  if (n_seg <= 0) return LSIM_OK;
  uint32_t *keys, *vals;
  sorted_buffers(w, d, keys, vals);
  HIP_TRY(hipMemcpyAsync(rec_out, w.rec, (size_t)n_seg * sizeof(SegRec), hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(keys_out, keys, (size_t)n_seg * 4, hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(order_out, vals, (size_t)n_seg * 4, hipMemcpyDeviceToDevice, st));
  return LSIM_OK;
}

int lsim_backward_query(lsim_ctx* c, int64_t n_seg, int64_t max_rows, size_t* bytes) {
  if (!c || !bytes) return fail(LSIM_E_ARG, "NULL argument");
  Carve cv(nullptr);
  cv.take<Pieces>(max_rows);
  cv.take<double>(max_rows);
  cv.take<double>(2 * max_rows * kNGrad);
  cv.take<AdjRec>(n_seg);
  *bytes = cv.off + 256;
  return LSIM_OK;
}

// Row count from which the adjoint's row kernels take the forward's work-balanced bands (rebuilt
// per backward, about 0.1 ms); below it, equal row counts per XCD.
constexpr int64_t kAdjBandRows = 32768;

static int backward_impl(lsim_ctx* c, const lsim_params* p, const lsim_segments* seg, int32_t n_events, void* ws,
                         size_t ws_bytes, int64_t max_rows, const lsim_fwd_out* fwd, const float* d_adc, const float* d_z,
                         const float* d_wfs, void* bwd_ws, size_t bwd_ws_bytes, double* d_params, void* stream) {
  if (!c || !p || !seg || !ws || !fwd || !fwd->tape || !d_adc || !bwd_ws || !d_params)
    return fail(LSIM_E_ARG, "NULL argument (the forward tape is required)");
  if (!c->has_lut) return fail(LSIM_E_STATE, "lsim_lut_set has not been called");
  hipStream_t st = (hipStream_t)stream;
  Dev d;
  int rc = make_dev(c, p, n_events, d);
  if (rc) return rc;
  rc = check_segments(seg, d.recomb);
  if (rc) return rc;
  const int64_t n = seg->n;
  Ws w = layout(ws, n, d.n_cells);
  if (w.bytes > ws_bytes) return fail(LSIM_E_SPACE, "workspace too small");
  if (max_rows > (n > 1 ? n : 1)) max_rows = n > 1 ? n : 1;
  size_t need;
  lsim_backward_query(c, n, max_rows, &need);
  if (need > bwd_ws_bytes) return fail(LSIM_E_SPACE, "backward workspace too small");
  Carve cv(bwd_ws);
  Pieces* pieces = cv.take<Pieces>(max_rows);
  double* dz = cv.take<double>(max_rows);
  double* partial = cv.take<double>(2 * max_rows * kNGrad);  // far pass rows, then near pass rows
  AdjRec* adj = cv.take<AdjRec>(n);
  Tables tb{c->TM3, c->TM, c->TN, c->TNP, c->DM, c->DN, c->CTM, c->CTN};
  SegIn in = make_segin(seg);
  uint32_t *keys, *vals;
  sorted_buffers(w, d, keys, vals);
  if (max_rows > 0) {
    // large launches: the rows' processing order and XCD bands, rebuilt as the forward builds them
    // (the schedule settings may have changed since; the result does not depend on them: the
    // partials are reduced by row); the rows no workgroup takes keep zero partials
    const bool banded = max_rows >= kAdjBandRows;
    if (banded) {
      const int rc2 = launch_row_bands(d, w, max_rows, st);
      if (rc2) return rc2;
    }
    fill<double>(partial, 2 * max_rows * kNGrad, 0.0, st);
    const int* band = banded ? w.band : nullptr;
    const int* order = banded ? row_order(w, d) : nullptr;
    hipLaunchKernelGGL(k_fee_adjoint, dim3((unsigned)((max_rows + kFeeRows - 1) / kFeeRows)), dim3(kFeeRows * kWave), 0, st,
                       d, (const FeeTape*)fwd->tape,
                       d_adc, d_z, fwd->n_rows, max_rows, pieces, dz);
    if (n > 0)
      hipLaunchKernelGGL(k_adj_records, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d, in, keys, vals, n, adj);
    const dim3 grid((unsigned)(banded ? kXcd * band_slots(max_rows) : max_rows)), block(kAdjWaves * kWave);
    double* pnear = partial + max_rows * kNGrad;
#define LSIM_ADJ_LAUNCH(NBV)                                                                                        \
  hipLaunchKernelGGL((k_rows_adj<NBV, kAdjWaves, false>), grid, block, 0, st, d, tb, w.rec, w.prs, adj, w.rng,       \
                     w.row_cell, w.n_rows, max_rows, band, order, pieces, partial);                               \
  hipLaunchKernelGGL((k_rows_adj<NBV, kAdjWaves, true>), grid, block, 0, st, d, tb, w.rec, w.prs, adj, w.rng,        \
                     w.row_cell, w.n_rows, max_rows, band, order, pieces, pnear)
#define LSIM_ADJ_DENSE(NBV)                                                                                       \
  hipLaunchKernelGGL((k_rows_adj_dense<NBV, kAdjWaves>), grid, block, 0, st, d, tb, w.rec, w.prs, adj, w.rng,       \
                     w.row_cell, w.n_rows, max_rows, band, order, pieces, d_wfs, partial)
    if (d_wfs) {
      if (d.nb == 5) {
        LSIM_ADJ_DENSE(5);
      } else if (d.nb == 3) {
        LSIM_ADJ_DENSE(3);
      } else {
        LSIM_ADJ_DENSE(1);
      }
    } else if (d.nb == 5) {
      LSIM_ADJ_LAUNCH(5);
    } else if (d.nb == 3) {
      LSIM_ADJ_LAUNCH(3);
    } else {
      LSIM_ADJ_LAUNCH(1);
    }
#undef LSIM_ADJ_LAUNCH
#undef LSIM_ADJ_DENSE
  }
  hipLaunchKernelGGL(k_grad_reduce, dim3(1), dim3(kReduceThreads), 0, st, d, partial, 2 * max_rows, dz, fwd->n_rows,
                     max_rows, d_params);
  HIP_TRY(hipGetLastError());
  return LSIM_OK;
}

int lsim_backward(lsim_ctx* c, const lsim_params* p, const lsim_segments* seg, int32_t n_events, void* ws, size_t ws_bytes,
                  int64_t max_rows, const lsim_fwd_out* fwd, const float* d_adc, const float* d_z, void* bwd_ws,
                  size_t bwd_ws_bytes, double* d_params, void* stream) {
  return backward_impl(c, p, seg, n_events, ws, ws_bytes, max_rows, fwd, d_adc, d_z, nullptr, bwd_ws, bwd_ws_bytes,
                       d_params, stream);
}

int lsim_backward_wfs(lsim_ctx* c, const lsim_params* p, const lsim_segments* seg, int32_t n_events, void* ws,
                      size_t ws_bytes, int64_t max_rows, const lsim_fwd_out* fwd, const float* d_adc, const float* d_z,
                      const float* d_wfs, void* bwd_ws, size_t bwd_ws_bytes, double* d_params, void* stream) {
  if (!d_wfs) return fail(LSIM_E_ARG, "NULL d_wfs");
  return backward_impl(c, p, seg, n_events, ws, ws_bytes, max_rows, fwd, d_adc, d_z, d_wfs, bwd_ws, bwd_ws_bytes,
                       d_params, stream);
}

int lsim_rbf_sums(const float* x, int64_t nx, const int32_t* lo, const int32_t* hi, const float* y, const float* c,
                  int64_t ny, float sigma, double* out, void* stream) {
  if (nx < 0 || ny < 0) return fail(LSIM_E_ARG, "negative size");
  if (!(sigma > 0.f)) return fail(LSIM_E_ARG, "sigma must be > 0");
  if (nx == 0) return LSIM_OK;
  if (!x || !lo || !hi || !out || (ny > 0 && (!y || !c))) return fail(LSIM_E_ARG, "NULL argument");
  if (nx >= ((int64_t)1 << 40)) return fail(LSIM_E_ARG, "too many points");
  const float inv = 1.f / (2.f * sigma * sigma);
  hipLaunchKernelGGL(k_rbf_sums<4>, dim3((unsigned)((nx + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     (const float4*)x, nx, lo, hi, (const float4*)y, (const float4*)c, inv, out);
  HIP_TRY(hipGetLastError());
  return LSIM_OK;
}

int lsim_param_electrons(lsim_ctx* c, const lsim_params* p, const lsim_segments* seg, int32_t mode, int32_t diffusion,
                         const float* normals, const double* bin_edges, int32_t n_edges, double* ex, double* ey,
                         double* et0, double* ene, double* esig, int32_t* eplane, int64_t* pair_pids, void* stream) {
  if (!c || !p || !seg) return fail(LSIM_E_ARG, "NULL argument");
  if (mode != 0 && mode != 1) return fail(LSIM_E_ARG, "mode must be 0 (transverse bins) or 1 (normal draws)");
  if (mode == 0 && (n_edges < 2 || n_edges > kParamMaxEdges || !bin_edges))
    return fail(LSIM_E_ARG, "transverse bins need 2 .. 17 bin edges");
  if (mode == 1 && seg->n > 0 && !normals) return fail(LSIM_E_ARG, "normal draws are NULL");
  Dev d;
  int rc = make_dev(c, p, 1, d);
  if (rc) return rc;
  rc = check_segments(seg, d.recomb);
  if (rc) return rc;
  if (seg->n == 0) return LSIM_OK;
  if (!ex || !ey || !et0 || !ene || !esig || !eplane || !pair_pids) return fail(LSIM_E_ARG, "NULL output");
  const SegIn in = make_segin(seg);
  ParamE e{ex, ey, et0, ene, esig, eplane};
  hipLaunchKernelGGL(k_param_electrons, dim3((unsigned)((seg->n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, d, in,
                     mode, diffusion, normals, bin_edges, n_edges, e, pair_pids);
  HIP_TRY(hipGetLastError());
  return LSIM_OK;
}

int lsim_param_rows(lsim_ctx* c, const lsim_params* p, const double* ex, const double* ey, const double* et0,
                    const double* ene, const double* esig, const int64_t* pixel_ids, int64_t n_rows,
                    const int32_t* pair_order, const int32_t* row_offsets, int32_t diffusion, float* wfs, void* stream) {
  if (!c || !p) return fail(LSIM_E_ARG, "NULL argument");
  if (n_rows < 0 || n_rows >= ((int64_t)1 << 31)) return fail(LSIM_E_ARG, "n_rows out of range");
  Dev d;
  int rc = make_dev(c, p, 1, d);
  if (rc) return rc;
  const int T = (int)(5.0 / d.ts) + 1;  // linspace(0, 5, int(5 / t_sampling) + 1)
  if (T < 2 || T > kParamMaxSamples) return fail(LSIM_E_ARG, "t_sampling must give 2 .. 63 current samples in 5 us");
  if (d.nout > kMaxOut) return fail(LSIM_E_ARG, "too many output ticks");
  if (n_rows == 0) return LSIM_OK;
  if (!ex || !ey || !et0 || !ene || !esig || !pixel_ids || !pair_order || !row_offsets || !wfs)
    return fail(LSIM_E_ARG, "NULL argument");
  const int K = (2 * d.nn + 1) * (2 * d.nn + 1);
  ParamE e{const_cast<double*>(ex), const_cast<double*>(ey), const_cast<double*>(et0), const_cast<double*>(ene),
           const_cast<double*>(esig), nullptr};
  hipLaunchKernelGGL(k_param_rows, dim3((unsigned)n_rows), dim3(kParamWaves * kWave), 0, (hipStream_t)stream, d, e,
                     pixel_ids, pair_order, row_offsets, K, T, diffusion, wfs);
  HIP_TRY(hipGetLastError());
  return LSIM_OK;
}

int lsim_param_backward_query(lsim_ctx* c, int64_t n_electrons, int64_t n_rows, size_t* bytes) {
  if (!c || !bytes) return fail(LSIM_E_ARG, "NULL argument");
  if (n_electrons < 0 || n_rows < 0) return fail(LSIM_E_ARG, "negative size");
  Carve cv(nullptr);
  cv.take<ParamAdj>(n_electrons);
  cv.take<Pieces>(n_rows);
  cv.take<double>(n_rows);
  cv.take<double>(n_rows * kNGrad);
  cv.take<int64_t>(1);
  *bytes = cv.off + 256;
  return LSIM_OK;
}

int lsim_param_backward(lsim_ctx* c, const lsim_params* p, const lsim_segments* seg, int32_t mode, int32_t diffusion,
                        const float* normals, const double* bin_edges, int32_t n_edges, const double* ex,
                        const double* ey, const double* et0, const double* ene, const double* esig,
                        const int64_t* pixel_ids, int64_t n_rows, const int32_t* pair_order, const int32_t* row_offsets,
                        const void* fee_tape, const float* d_adc, const float* d_z, const float* d_wfs, void* ws,
                        size_t ws_bytes, double* d_params, void* stream) {
  if (!c || !p || !seg || !d_params || !ws) return fail(LSIM_E_ARG, "NULL argument");
  if (mode != 0 && mode != 1) return fail(LSIM_E_ARG, "mode must be 0 (transverse bins) or 1 (normal draws)");
  if (mode == 0 && (n_edges < 2 || n_edges > kParamMaxEdges || !bin_edges))
    return fail(LSIM_E_ARG, "transverse bins need 2 .. 17 bin edges");
  if (mode == 1 && seg->n > 0 && !normals) return fail(LSIM_E_ARG, "normal draws are NULL");
  if (n_rows < 0 || n_rows >= ((int64_t)1 << 31)) return fail(LSIM_E_ARG, "n_rows out of range");
  Dev d;
  int rc = make_dev(c, p, 1, d);
  if (rc) return rc;
  rc = check_segments(seg, d.recomb);
  if (rc) return rc;
  const int T = (int)(5.0 / d.ts) + 1;
  if (T < 2 || T > kParamMaxSamples) return fail(LSIM_E_ARG, "t_sampling must give 2 .. 63 current samples in 5 us");
  if (d.nout > kMaxOut) return fail(LSIM_E_ARG, "too many output ticks");
  const int64_t per = mode == 0 ? (int64_t)(n_edges - 1) * (n_edges - 1) : 1;
  const int64_t n_el = seg->n * per;
  size_t need;
  lsim_param_backward_query(c, n_el, n_rows, &need);
  if (need > ws_bytes) return fail(LSIM_E_SPACE, "workspace too small");
  hipStream_t st = (hipStream_t)stream;
  Carve cv(ws);
  ParamAdj* adj = cv.take<ParamAdj>(n_el);
  Pieces* pieces = cv.take<Pieces>(n_rows);
  double* dz = cv.take<double>(n_rows);
  double* partial = cv.take<double>(n_rows * kNGrad);
  int64_t* n_rows_dev = cv.take<int64_t>(1);
  fill<int64_t>(n_rows_dev, 1, seg->n > 0 ? n_rows : 0, st);
  if (seg->n > 0 && n_rows > 0) {
    if (!ex || !ey || !et0 || !ene || !esig || !pixel_ids || !pair_order || !row_offsets || !fee_tape || !d_adc)
      return fail(LSIM_E_ARG, "NULL argument (the electrons, the rows, the readout tape and d_adc are required)");
    const int K = (2 * d.nn + 1) * (2 * d.nn + 1);
    hipLaunchKernelGGL(k_fee_adjoint, dim3((unsigned)((n_rows + kFeeRows - 1) / kFeeRows)), dim3(kFeeRows * kWave), 0, st, d,
                       (const FeeTape*)fee_tape, d_adc, d_z, (const int64_t*)n_rows_dev, n_rows, pieces, dz);
    hipLaunchKernelGGL(k_param_adj_electrons, dim3((unsigned)((seg->n + 255) / 256)), dim3(256), 0, st, d, make_segin(seg),
                       mode, diffusion, normals, bin_edges, n_edges, adj);
    ParamE e{const_cast<double*>(ex), const_cast<double*>(ey), const_cast<double*>(et0), const_cast<double*>(ene),
             const_cast<double*>(esig), nullptr};
    hipLaunchKernelGGL(k_param_rows_adj, dim3((unsigned)n_rows), dim3(kParamWaves * kWave), 0, st, d, e, adj, pixel_ids,
                       pair_order, row_offsets, K, T, diffusion, pieces, d_wfs, partial);
  }
  hipLaunchKernelGGL(k_grad_reduce, dim3(1), dim3(kReduceThreads), 0, st, d, partial, seg->n > 0 ? n_rows : 0, dz,
                     (const int64_t*)n_rows_dev, n_rows, d_params);
  HIP_TRY(hipGetLastError());
  return LSIM_OK;
}

float lsim_rbf_cull_radius(float sigma) {
  // exp(-d^2 / (2 sigma^2)) is exactly 0 in float32 once d^2 / (2 sigma^2) > 106; 1 % margin
  return 1.01f * sqrtf(106.f * 2.f * sigma * sigma);
}

int lsim_rbf_sums_grid(const float* x, const int32_t* x_event, int64_t nx, const float* y, const float* c,
                       const int64_t* y_key, int64_t ny, const float* grid, int32_t dims, float sigma, double* out,
                       void* stream) {
  if (nx < 0 || ny < 0) return fail(LSIM_E_ARG, "negative size");
  if (ny >= ((int64_t)1 << 31) || nx >= ((int64_t)1 << 40)) return fail(LSIM_E_ARG, "too many points");
  if (dims != 3 && dims != 4) return fail(LSIM_E_ARG, "dims must be 3 or 4");
  if (!(sigma > 0.f) || !std::isfinite(sigma)) return fail(LSIM_E_ARG, "sigma must be finite and > 0");
  if (nx == 0) return LSIM_OK;
  if (!x || !x_event || !grid || !out || (ny > 0 && (!y || !c || !y_key))) return fail(LSIM_E_ARG, "NULL argument");
  const float inv = 1.f / (2.f * sigma * sigma);
  const float rc = lsim_rbf_cull_radius(sigma);
  const dim3 g((unsigned)((nx + 255) / 256));
  if (dims == 3)
    hipLaunchKernelGGL(k_rbf_grid<3>, g, dim3(256), 0, (hipStream_t)stream, (const float4*)x, x_event, nx,
                       (const float4*)y, (const float4*)c, y_key, (int32_t)ny, grid, inv, rc, out);
  else
    hipLaunchKernelGGL(k_rbf_grid<4>, g, dim3(256), 0, (hipStream_t)stream, (const float4*)x, x_event, nx,
                       (const float4*)y, (const float4*)c, y_key, (int32_t)ny, grid, inv, rc, out);
  HIP_TRY(hipGetLastError());
  return LSIM_OK;
}

int lsim_nearest(const float* x, int64_t nx, const int32_t* lo, const int32_t* hi, const float* y, int64_t ny,
                 int64_t n_origin, int32_t* idx, float* dmin, void* stream) {
  if (nx < 0 || ny < 0 || n_origin < 0 || n_origin > nx) return fail(LSIM_E_ARG, "bad sizes");
  if (nx == 0) return LSIM_OK;
  if (!x || !lo || !hi || !idx || !dmin || (ny > 0 && !y)) return fail(LSIM_E_ARG, "NULL argument");
  if (nx >= ((int64_t)1 << 40) || ny >= ((int64_t)1 << 31)) return fail(LSIM_E_ARG, "too many points");
  hipLaunchKernelGGL(k_nearest, dim3((unsigned)((nx + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     (const float4*)x, nx, lo, hi, (const float4*)y, n_origin, idx, dmin);
  HIP_TRY(hipGetLastError());
  return LSIM_OK;
}

// ---- track chopping ----------------------------------------------------------------------------

static int chop_layout(const lsim_chop_layout* l, ChopLayout& L) {
  if (!l) return fail(LSIM_E_ARG, "NULL layout");
  L.F = l->n_fields;
  if (L.F < 1 || L.F > kChopMaxFields) return fail(LSIM_E_ARG, "n_fields must be in 1 .. 64");
  auto col = [&](int32_t c) { return c >= 0 && c < L.F; };
  for (int a = 0; a < 3; ++a) {
    L.s[a] = l->start[a];
    L.e[a] = l->end[a];
    L.m[a] = l->mid[a];
    if (!col(L.s[a]) || !col(L.e[a]) || !col(L.m[a])) return fail(LSIM_E_ARG, "position column out of range");
  }
  L.dE = l->dE;
  L.dx = l->dx;
  L.ev = l->event;
  if (!col(L.dE) || !col(L.dx)) return fail(LSIM_E_ARG, "dE / dx column out of range");
  if (L.ev != -1 && !col(L.ev)) return fail(LSIM_E_ARG, "event column out of range");
  if (l->n_zero < 0 || l->n_zero > kChopMaxCols || l->n_neg_one < 0 || l->n_neg_one > kChopMaxCols)
    return fail(LSIM_E_ARG, "at most 8 zero and 8 -1 columns");
  L.nzero = l->n_zero;
  L.nneg = l->n_neg_one;
  for (int c = 0; c < L.nzero; ++c)
    if (!col(L.zero[c] = l->zero[c])) return fail(LSIM_E_ARG, "zero column out of range");
  for (int c = 0; c < L.nneg; ++c)
    if (!col(L.neg[c] = l->neg_one[c])) return fail(LSIM_E_ARG, "-1 column out of range");
  return LSIM_OK;
}

static size_t chop_ws_bytes(int64_t n) { return (size_t)(n + scan_tiles(n) + 4) * sizeof(int64_t); }

static int chop_check(const float* tracks, int64_t n, double precision, size_t ws_bytes) {
  if (n < 0) return fail(LSIM_E_ARG, "n_tracks must be >= 0");
  if (n >= ((int64_t)1 << 40)) return fail(LSIM_E_ARG, "too many tracks");
  if (!(precision > 0) || !std::isfinite(precision)) return fail(LSIM_E_ARG, "precision must be finite and > 0");
  if (n > 0 && !tracks) return fail(LSIM_E_ARG, "NULL tracks");
  if (ws_bytes < chop_ws_bytes(n)) return fail(LSIM_E_SPACE, "chop workspace too small");
  return LSIM_OK;
}

int lsim_chop_query(int64_t n_tracks, size_t* ws_bytes) {
  if (n_tracks < 0 || !ws_bytes) return fail(LSIM_E_ARG, "bad arguments");
  *ws_bytes = chop_ws_bytes(n_tracks);
  return LSIM_OK;
}

int lsim_chop_count(const lsim_chop_layout* layout, const float* tracks, int64_t n_tracks, double precision, void* ws,
                    size_t ws_bytes, int64_t* info, void* stream) {
  ChopLayout L;
  int rc = chop_layout(layout, L);
  if (rc != LSIM_OK) return rc;
  if ((rc = chop_check(tracks, n_tracks, precision, ws_bytes)) != LSIM_OK) return rc;
  if (!ws || !info) return fail(LSIM_E_ARG, "NULL workspace or info");
  hipStream_t st = (hipStream_t)stream;
  int64_t* off = (int64_t*)ws;
  int64_t* tmp = off + n_tracks;
  hipLaunchKernelGGL((k_fill<int64_t>), dim3(1), dim3(64), 0, st, info, (int64_t)3, (int64_t)0);
  HIP_TRY(hipGetLastError());
  if (n_tracks == 0) return LSIM_OK;
  hipLaunchKernelGGL(k_chop_count, dim3((unsigned)((n_tracks + 255) / 256)), dim3(256), 0, st, L, tracks, n_tracks,
                     (float)precision, off, info);
  HIP_TRY(hipGetLastError());
  HIP_TRY((exclusive_scan<int64_t, int64_t>(off, off, n_tracks, tmp, info, st)));
  return LSIM_OK;
}

int lsim_chop_fill(const lsim_chop_layout* layout, const float* tracks, int64_t n_tracks, double precision,
                   const void* ws, size_t ws_bytes, int64_t n_steps, int64_t n_rows, const int64_t* src_events,
                   const int64_t* batch_events, int64_t n_batch_events, int64_t n_local_events, float* out,
                   int64_t* info, void* stream) {
  ChopLayout L;
  int rc = chop_layout(layout, L);
  if (rc != LSIM_OK) return rc;
  if ((rc = chop_check(tracks, n_tracks, precision, ws_bytes)) != LSIM_OK) return rc;
  if (n_steps < 0 || n_rows < n_steps) return fail(LSIM_E_ARG, "need 0 <= n_steps <= n_rows");
  if (n_steps > 0 && n_tracks == 0) return fail(LSIM_E_ARG, "n_steps > 0 with no tracks");
  if (n_steps < n_tracks) return fail(LSIM_E_ARG, "n_steps below the track count (every track has a step)");
  if (n_rows >= ((int64_t)1 << 40)) return fail(LSIM_E_ARG, "too many rows");
  if (src_events && (L.ev < 0 || n_batch_events < 0 || (n_batch_events > 0 && !batch_events)))
    return fail(LSIM_E_ARG, "event remap needs the event column and the batch event ids");
  if (n_local_events < -1) return fail(LSIM_E_ARG, "n_local_events must be >= -1");
  if (n_rows == 0) return LSIM_OK;
  if (!out || !info || (n_steps > 0 && !ws)) return fail(LSIM_E_ARG, "NULL output, info or workspace");
  const size_t lds = (size_t)kChopRows * L.F * sizeof(float);
  const dim3 grid((unsigned)((n_rows + kChopRows - 1) / kChopRows));
  if (((uintptr_t)out & 15) == 0)
    hipLaunchKernelGGL(k_chop_fill<true>, grid, dim3(kChopRows), lds, (hipStream_t)stream, L, tracks, n_tracks,
                       precision, (const int64_t*)ws, n_steps, n_rows, src_events, batch_events, n_batch_events,
                       n_local_events, out, info);
  else
    hipLaunchKernelGGL(k_chop_fill<false>, grid, dim3(kChopRows), lds, (hipStream_t)stream, L, tracks, n_tracks,
                       precision, (const int64_t*)ws, n_steps, n_rows, src_events, batch_events, n_batch_events,
                       n_local_events, out, info);
  HIP_TRY(hipGetLastError());
  return LSIM_OK;
}

// ---- probabilistic readout ---------------------------------------------------------------------

static size_t pf_flags_bytes(int S) { return (((size_t)(S + 1) * sizeof(int) + 255) / 256) * 256; }

// flags [S + 1] | per-row tables [P][4 T] | path tape: charges, log probabilities [S + 1][P][V] and
// selected ticks [S][P][V]
static size_t pf_ws_bytes(int64_t P, int T, int V, int S) {
  return pf_flags_bytes(S) + (size_t)P * 4 * T * sizeof(float) + (size_t)2 * (S + 1) * P * V * sizeof(float) +
         (size_t)S * P * V * sizeof(int);
}

struct PfWs {
  int* flags;
  float* tab;
  float* tc;
  float* tp;
  int* tk;
};

static PfWs pf_layout(void* ws, int64_t P, int T, int V, int S) {
  PfWs w;
  w.flags = (int*)ws;
  w.tab = (float*)((char*)ws + pf_flags_bytes(S));
  w.tc = w.tab + (size_t)P * 4 * T;
  w.tp = w.tc + (size_t)(S + 1) * P * V;
  w.tk = (int*)(w.tp + (size_t)(S + 1) * P * V);
  return w;
}

static int pf_check(int64_t n_rows, int32_t n_ticks, int32_t n_paths, int32_t n_hits, int32_t interval, float noise_sigma,
                    float t_sampling) {
  if (n_rows < 0) return fail(LSIM_E_ARG, "n_rows must be >= 0");
  if (n_ticks < 2 || n_ticks > kPfMaxTicks) return fail(LSIM_E_ARG, "n_ticks must be in 2 .. 2048");
  if (n_paths < 1 || n_paths > kPfMaxPaths || n_paths > n_ticks - 1)
    return fail(LSIM_E_ARG, "n_paths must be in 1 .. min(32, n_ticks - 1)");
  if (n_hits < 1 || n_hits > 64) return fail(LSIM_E_ARG, "n_hits must be in 1 .. 64");
  if (interval < 0) return fail(LSIM_E_ARG, "interval must be >= 0");
  if (!(noise_sigma > 0.f) || !(t_sampling > 0.f)) return fail(LSIM_E_ARG, "noise_sigma and t_sampling must be > 0");
  if (n_rows >= ((int64_t)1 << 31)) return fail(LSIM_E_ARG, "too many rows");
  return LSIM_OK;
}

int lsim_prob_fee_query(int64_t n_rows, int32_t n_ticks, int32_t n_paths, int32_t n_hits, size_t* ws_bytes) {
  if (n_rows < 0 || n_ticks < 2 || n_paths < 1 || n_hits < 1 || !ws_bytes) return fail(LSIM_E_ARG, "bad arguments");
  *ws_bytes = pf_ws_bytes(n_rows, n_ticks, n_paths, n_hits);
  return LSIM_OK;
}

int lsim_prob_fee(const float* wfs, int64_t n_rows, int32_t n_ticks, int32_t n_paths, int32_t n_hits, int32_t interval,
                  float threshold, float noise_sigma, float t_sampling, float stop_threshold, void* ws,
                  size_t ws_bytes, float* log_prob, float* charge, void* stream) {
  int rc = pf_check(n_rows, n_ticks, n_paths, n_hits, interval, noise_sigma, t_sampling);
  if (rc) return rc;
  if (!(stop_threshold > 0.f)) return fail(LSIM_E_ARG, "stop_threshold must be > 0");
  if (ws_bytes < pf_ws_bytes(n_rows, n_ticks, n_paths, n_hits)) return fail(LSIM_E_SPACE, "workspace too small");
  if (n_rows == 0) return LSIM_OK;
  if (!wfs || !ws || !log_prob || !charge) return fail(LSIM_E_ARG, "NULL argument");
  PfDev d;
  d.T = n_ticks;
  d.V = n_paths;
  d.S = n_hits;
  d.interval = interval;
  d.theta = threshold;
  d.zs = 1.f / noise_sigma;
  d.ts = t_sampling;
  d.log_stop = logf(stop_threshold);
  hipStream_t st = (hipStream_t)stream;
  const PfWs w = pf_layout(ws, n_rows, n_ticks, n_paths, n_hits);
  hipLaunchKernelGGL((k_fill<int>), dim3(1), dim3(64), 0, st, w.flags, (int64_t)(n_hits + 1), 0);
  hipLaunchKernelGGL((k_fill<int>), dim3(1), dim3(64), 0, st, w.flags, (int64_t)1, 1);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_pf_prep, dim3((unsigned)n_rows), dim3(kPfThreads), 0, st, d, wfs, n_rows, w.tab);
  HIP_TRY(hipGetLastError());
  for (int s = 0; s < n_hits; ++s) {
    hipLaunchKernelGGL(k_pf_step, dim3((unsigned)n_rows), dim3(kPfThreads), 0, st, d, s, (const float*)w.tab, n_rows,
                       w.tc, w.tp, w.tk, w.flags, log_prob, charge);
    HIP_TRY(hipGetLastError());
  }
  return LSIM_OK;
}

int lsim_prob_fee_backward(int64_t n_rows, int32_t n_ticks, int32_t n_paths, int32_t n_hits, int32_t interval,
                           float threshold, float noise_sigma, float t_sampling, const void* ws, size_t ws_bytes,
                           const float* log_prob, const float* d_log_prob, const float* d_charge, float* d_wfs,
                           void* stream) {
  int rc = pf_check(n_rows, n_ticks, n_paths, n_hits, interval, noise_sigma, t_sampling);
  if (rc) return rc;
  if (ws_bytes < pf_ws_bytes(n_rows, n_ticks, n_paths, n_hits)) return fail(LSIM_E_SPACE, "workspace too small");
  if (n_rows == 0) return LSIM_OK;
  if (!ws || !log_prob || !d_log_prob || !d_charge || !d_wfs) return fail(LSIM_E_ARG, "NULL argument");
  PfDev d;
  d.T = n_ticks;
  d.V = n_paths;
  d.S = n_hits;
  d.interval = interval;
  d.theta = threshold;
  d.zs = 1.f / noise_sigma;
  d.ts = t_sampling;
  d.log_stop = 0.f;
  const PfWs w = pf_layout(const_cast<void*>(ws), n_rows, n_ticks, n_paths, n_hits);
  hipLaunchKernelGGL(k_pf_adj, dim3((unsigned)n_rows), dim3(kPfThreads), 0, (hipStream_t)stream, d,
                     (const float*)w.tab, n_rows, (const float*)w.tc, (const float*)w.tp, (const int*)w.tk,
                     (const int*)w.flags, log_prob, d_log_prob, d_charge, d_wfs);
  HIP_TRY(hipGetLastError());
  return LSIM_OK;
}

// ---- soft-DTW -----------------------------------------------------------------------------------

__global__ void k_sdtw_value(const float* R, int64_t n, int64_t m, float* value) {
  if (threadIdx.x == 0 && blockIdx.x == 0) value[0] = R[(n - 1) * m + (m - 1)];
}

static size_t sdtw_ws_bytes(int64_t n, int64_t m) {
  const int64_t nTJ = (m + kDtwTile - 1) / kDtwTile;
  return (size_t)(2 * n * m + n * nTJ) * sizeof(float);
}

int lsim_sdtw_query(int64_t n, int64_t m, size_t* ws_bytes) {
  if (n < 1 || m < 1 || !ws_bytes) return fail(LSIM_E_ARG, "bad arguments");
  // R and E take 8 bytes per cell: 2^34 cells = 137 GB, half of one GPU's HBM
  if (n >= ((int64_t)1 << 31) || m >= ((int64_t)1 << 31) || n * m > ((int64_t)1 << 34))
    return fail(LSIM_E_ARG, "sequences too long (n * m above 2^34 cells)");
  *ws_bytes = sdtw_ws_bytes(n, m);
  return LSIM_OK;
}

int lsim_sdtw(const float* x, int64_t n, const float* y, int64_t m, float gamma, void* ws, size_t ws_bytes,
              float* value, double* grad_x, void* stream) {
  size_t need = 0;
  int rc = lsim_sdtw_query(n, m, &need);
  if (rc != LSIM_OK) return rc;
  if (!(gamma > 0.f)) return fail(LSIM_E_ARG, "gamma must be > 0");
  if (!x || !y || !ws || !value) return fail(LSIM_E_ARG, "NULL argument");
  if (ws_bytes < need) return fail(LSIM_E_SPACE, "soft-DTW workspace too small");
  hipStream_t st = (hipStream_t)stream;
  float* R = (float*)ws;
  float* E = R + n * m;
  float* part = E + n * m;
  const int nTI = (int)((n + kDtwTile - 1) / kDtwTile), nTJ = (int)((m + kDtwTile - 1) / kDtwTile);
  const unsigned diag = (unsigned)(nTI < nTJ ? nTI : nTJ);
  for (int k = 0; k < nTI + nTJ - 1; ++k) {
    hipLaunchKernelGGL(k_sdtw_fwd, dim3(diag), dim3(kDtwTile), 0, st, x, n, y, m, gamma, k, R);
    HIP_TRY(hipGetLastError());
  }
  hipLaunchKernelGGL(k_sdtw_value, dim3(1), dim3(64), 0, st, (const float*)R, n, m, value);
  HIP_TRY(hipGetLastError());
  if (!grad_x) return LSIM_OK;
  for (int k = nTI + nTJ - 2; k >= 0; --k) {
    hipLaunchKernelGGL(k_sdtw_bwd, dim3(diag), dim3(kDtwTile), 0, st, x, n, y, m, gamma, k, (const float*)R, E, part);
    HIP_TRY(hipGetLastError());
  }
  hipLaunchKernelGGL(k_sdtw_grad, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (const float*)part, n, nTJ,
                     grad_x);
  HIP_TRY(hipGetLastError());
  return LSIM_OK;
}

}  // extern "C"
