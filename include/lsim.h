/*
 * lsim.h -- C ABI of the MI355X (gfx950) implementation of the larnd-sim-jax LUT simulation
 * chain: quenching -> drift -> segment-to-pixel -> LUT induced current -> pixel x tick
 * waveforms -> self-trigger ADC, and the adjoint of that chain for 7 detector parameters.
 *
 * Conventions
 *   - Every entry point returns 0 on success or a negative LSIM_E_* code; lsim_last_error()
 *     returns a thread-local message for the last failure.  Nothing throws across the ABI.
 *   - All array arguments are device pointers owned by the caller unless stated otherwise; the
 *     context owns only the tables derived from the response LUT.
 *   - `stream` is a hipStream_t passed as void*.  All work is stream ordered; no entry point
 *     below lsim_forward_prepare synchronises the host, so lsim_forward / lsim_backward can be
 *     captured into a hipGraph.
 *   - Results are bitwise reproducible for fixed inputs (no float atomics anywhere).
 *
 * Reference interfaces replaced (larnd-sim-jax, paths relative to the repository root):
 *   lsim_ctx_create / lsim_lut_set : load_detector_properties + load_lut
 *                                    (src/larndsim/consts_jax.py:220-385, 387-449)
 *   lsim_forward_prepare          : simulate_drift_new + the unique main-pixel set
 *                                    (src/larndsim/sim_jax.py:374-453, 716-725)
 *   lsim_forward_rows             : simulate_signals + get_adc_values + digitize + parse_output
 *                                    (src/larndsim/sim_jax.py:141-286, 738-769, 619-647;
 *                                     src/larndsim/fee_jax.py:55-71, 168-279)
 *   lsim_forward                  : simulate_wfs + simulate_stochastic, i.e. the body of
 *                                    LUTSimulation.predict (optimize/strategies.py:91-105)
 *   lsim_fee_forward              : simulate_stochastic on caller waveforms (sim_jax.py:738-769)
 *   lsim_backward                 : the reverse-mode derivative jax.value_and_grad takes through
 *                                    LUTSimulation.predict (optimize/fit_params.py:731)
 */
#ifndef LSIM_H
#define LSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LSIM_OK 0
#define LSIM_E_ARG (-1)       /* invalid argument / unsupported configuration */
#define LSIM_E_HIP (-2)       /* a HIP runtime call failed */
#define LSIM_E_STATE (-3)     /* call order violated (e.g. no LUT set) */
#define LSIM_E_SPACE (-4)     /* workspace too small */

#define LSIM_MAX_TPC 8
#define LSIM_MAX_TEMPLATES 128
#define LSIM_MAX_ADC 16

/* status bits written to lsim_fwd_out.status (device int32) */
#define LSIM_STATUS_ROWS_OVERFLOW 1   /* more pixel rows than max_rows: extra rows skipped */
#define LSIM_STATUS_BAD_EVENT 2       /* eventID >= n_events: segment ignored */

typedef struct lsim_ctx lsim_ctx;

/* Detector geometry (consts_jax.py:357-383).  tpc_borders[t][axis][0/1] in cm; for the z
 * axis entry 0 is the anode and entry 1 the cathode. */
typedef struct {
    int32_t n_tpc;
    int32_t n_pixels_x, n_pixels_y;
    double pixel_pitch;
    double tpc_borders[LSIM_MAX_TPC][3][2];
} lsim_geometry;

/* Static simulation configuration (the non-differentiable Params fields that shape arrays). */
typedef struct {
    int32_t number_pix_neighbors;        /* n: (2n+1)^2 neighbour pixels */
    int32_t nb_tran_diff_bins;           /* 1, 3 or 5 */
    int32_t nb_sampling_bins_per_pixel;  /* 10 */
    int32_t signal_length;               /* L: samples per window (150) */
    int32_t n_ticks;                     /* array ticks incl. the garbage column (2001) */
    int32_t max_adc_values;              /* H <= LSIM_MAX_ADC (10) */
    int32_t adc_interval;                /* self-trigger integration ticks (18) */
    int32_t n_templates;                 /* longitudinal-diffusion templates (100) */
    int32_t long_diff_extent;            /* Gaussian half width in ticks (20) */
    double long_diff_template[LSIM_MAX_TEMPLATES]; /* template sigmas in ticks, ascending */
} lsim_config;

/* Per-call parameters (Params fields read by the path).  recombination_mode: 1 BOX, 2 BIRKS,
 * 3 ELLIPSOID.  vdrift_eff / dvdrift_deField are get_vdrift(eField) and its derivative,
 * computed by the caller (the `vdrift` field itself is never read, as in the reference). */
typedef struct {
    double eField, Ab, kb, lifetime, vdrift, long_diff, tran_diff;
    double lArDensity, alpha, beta, R_param, MeVToElectrons;
    double vdrift_eff, dvdrift_deField;
    double shift_x, shift_y, shift_z, size_margin, t_sampling;
    double threshold, gain, v_pedestal, v_cm, v_ref, adc_counts;
    int32_t recombination_mode;
    int32_t pad_;
} lsim_params;

/* Track segments.  Element k of a column is at ptr[k * stride]: stride 1 is SoA, stride F reads
 * the reference's (N, F) float32 row-major array in place (pointers = base + column index).
 * z_start, z_end and dx are read only for ELLIPSOID recombination (may be NULL otherwise). */
typedef struct {
    const float *x, *y, *z, *dE, *dEdx, *event_id;
    const float *z_start, *z_end, *dx;
    int64_t stride;
    int64_t n;
} lsim_segments;

/* Forward outputs, all caller-allocated device buffers sized for `max_rows` pixel rows. */
typedef struct {
    int64_t *n_rows;     /* [1]  number of valid pixel rows P (sorted unique main pixels) */
    int64_t *pixel_ids;  /* [max_rows] pixel id of each row, ascending */
    float *wfs;          /* [max_rows][n_ticks-1] waveforms (NULL: not stored) */
    float *adc;          /* [max_rows][H] digitised ADC counts (noise off) */
    int32_t *ticks;      /* [max_rows][H] hit ticks (n_ticks-3 = no hit) */
    void *tape;          /* [max_rows * lsim_tape_bytes()] self-trigger decisions for the
                            adjoint (NULL: not stored) */
    int64_t *n_hits;     /* [1]  number of compacted hits (parse_output) */
    float *hit_adc, *hit_x, *hit_y, *hit_z, *hit_tick, *hit_prob; /* [max_rows*H] */
    int64_t *hit_event, *hit_pixel;                                /* [max_rows*H] */
    int32_t *hit_index;  /* [max_rows*H] row*H + k of each compacted hit */
    int32_t *status;     /* [1]  LSIM_STATUS_* bits */
} lsim_fwd_out;

const char *lsim_last_error(void);
int lsim_version(void);
size_t lsim_tape_bytes(void);

int lsim_ctx_create(lsim_ctx **out, const lsim_geometry *geo, const lsim_config *cfg);
int lsim_ctx_destroy(lsim_ctx *ctx);

/* Build the template tables from the raw (nx, ny, nt) float32 LUT on the device. */
int lsim_lut_set(lsim_ctx *ctx, const float *raw_lut_dev, int32_t nx, int32_t ny, int32_t nt, void *stream);

/* Workspace bytes for n_seg segments of n_events events, and an upper bound on pixel rows. */
int lsim_forward_query(lsim_ctx *ctx, int64_t n_seg, int32_t n_events, size_t *ws_bytes, int64_t *max_rows_bound);

/* Phase 1: per-segment quantities, sort by main pixel, pixel-row enumeration.  Writes the row
 * count to n_rows_dev (device int64).  The workspace keeps the state phase 2 and the backward
 * read. */
int lsim_forward_prepare(lsim_ctx *ctx, const lsim_params *p, const lsim_segments *seg, int32_t n_events,
                         void *ws, size_t ws_bytes, int64_t *n_rows_dev, int32_t *status_dev, void *stream);

/* Phase 2: waveforms of every row, self-trigger, digitisation, hit compaction.  n_seg and
 * n_events must match the prepare call that filled `ws`. */
int lsim_forward_rows(lsim_ctx *ctx, const lsim_params *p, int64_t n_seg, int32_t n_events, void *ws,
                      size_t ws_bytes, int64_t max_rows, const lsim_fwd_out *out, void *stream);

/* Both phases in one stream-ordered call (graph capturable). */
int lsim_forward(lsim_ctx *ctx, const lsim_params *p, const lsim_segments *seg, int32_t n_events, void *ws,
                 size_t ws_bytes, int64_t max_rows, const lsim_fwd_out *out, void *stream);

/* Negative ids of the reference's pixel set (src/larndsim/sim_jax.py:716-721: sorted unique main-pixel
 * ids plus an appended -1).  A padding segment (eventID -1) on the pixel plane has the negative id
 * pixel2id packs for event -1 (src/larndsim/detsim_jax.py:230-244); a segment whose main pixel is
 * off the plane has id -1.  The reference keeps a row for each distinct one; they carry no charge
 * and no hits.  lsim_negative_rows_query gives the workspace bytes and the largest possible count
 * (n_tpc * n_pixels_x * n_pixels_y); lsim_negative_rows writes the distinct negative main-pixel ids
 * (not the appended -1) in ascending order to ids (at most max_ids of them) and their count to
 * *n_ids (device int64).  Segments with eventID < -1 are skipped. */
int lsim_negative_rows_query(lsim_ctx *ctx, size_t *ws_bytes, int64_t *max_ids);
int lsim_negative_rows(lsim_ctx *ctx, const lsim_params *p, const lsim_segments *seg, void *ws, size_t ws_bytes,
                       int64_t *ids, int64_t max_ids, int64_t *n_ids, void *stream);

/* Workspace bytes of lsim_fee_forward for n_rows rows. */
int lsim_fee_query(lsim_ctx *ctx, int64_t n_rows, size_t *ws_bytes);

/* Self-trigger + digitisation + compaction of caller waveforms wfs[n_rows][n_ticks-1] with pixel
 * ids pixel_ids[n_rows] (rows with id < 0 produce no hits).  out->n_rows/pixel_ids/wfs are
 * ignored; the adc/ticks/hit arrays must hold n_rows rows; out->tape is written if not NULL. */
int lsim_fee_forward(lsim_ctx *ctx, const lsim_params *p, const float *wfs, const int64_t *pixel_ids, int64_t n_rows,
                     void *ws, size_t ws_bytes, const lsim_fwd_out *out, void *stream);

/* lsim_fee_forward with no workspace argument and no allocation of any kind: its scratch (hit
 * counts, offsets, scan partials) lives in a file-scope device array that each GPU holds until
 * the process ends (about 8.4 MB).  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under hipGraph capture, serving loops, harnesses that forbid
 * hipMalloc).  At most LSIM_NOALLOC_MAX_ROWS rows (LSIM_E_SPACE beyond).  Calls share their GPU's
 * scratch, so calls that overlap on one GPU (other streams, other threads) race; calls ordered on
 * one stream do not.  Results equal lsim_fee_forward's bit for bit. */
#define LSIM_NOALLOC_MAX_ROWS (1 << 20)
int lsim_fee_forward_noalloc(lsim_ctx *ctx, const lsim_params *p, const float *wfs, const int64_t *pixel_ids,
                             int64_t n_rows, const lsim_fwd_out *out, void *stream);

/* Copy the per-segment records of the last prepare call (in main-pixel order) for inspection:
 * rec_out[n_seg * 80 bytes] = {int k0, uxy; float n_e, f; int bx, by, idx; float a, b, c, wx[5], wy[5]}
 * (uxy = (bx mod 10) | (by mod 10) << 16),
 * keys_out[n_seg] = cell key (>= number of cells for dropped segments), order_out[n_seg] =
 * input index of each record. */
int lsim_segment_records(lsim_ctx *ctx, int64_t n_seg, int32_t n_events, void *ws, size_t ws_bytes, void *rec_out,
                         uint32_t *keys_out, uint32_t *order_out, void *stream);

/* Measurement: while enabled, hipEvents are recorded on the launch stream around every launch
 * of the waveform kernel; lsim_timing_read waits for them and returns the summed kernel time
 * (ms) and the number of launches since lsim_timing(1).  Not for use under graph capture. */
int lsim_timing(int enable);
int lsim_timing_read(double *total_ms, int64_t *launches);

/* Tuning: waves per pixel-row workgroup of the waveform kernel (1, 2, 4 or 8; default 4). */
int lsim_set_row_waves(int waves);

/* Tuning: the fixed cost of one pixel row in units of segment-row pairs (default 256), used to cut
 * the rows into per-XCD bands of equal estimated work for the waveform kernel.  Only
 * the schedule changes: results are the same for every value. */
int lsim_set_row_cost(int pairs);

/* Scheduling knob of the waveform kernel, no counterpart in the reference (like the two above): the
 * pixel rows are processed tile by tile, `cells` x `cells` pixel cells of one event plane per tile
 * (default 16; 1 = plain row order), so that the rows running together on one XCD share their
 * segments in its L2.  Changes only the order rows are computed in, never their values. */
int lsim_set_row_tile(int cells);

/* Backward workspace bytes for n_seg segments and max_rows rows. */
int lsim_backward_query(lsim_ctx *ctx, int64_t n_seg, int64_t max_rows, size_t *bwd_ws_bytes);

/* Gradient of sum(d_adc * adc) + sum(d_z * hit_z) with respect to
 * (Ab, kb, eField, lifetime, vdrift, long_diff, tran_diff), written to d_params[7] (device
 * float64).  d_adc and d_z are [max_rows][H] cotangents on the digitised ADC grid and on the
 * hit z grid (d_z may be NULL).  Requires the workspace and the tape of the matching forward
 * call.  Integer decisions (bins, pixels, ticks, template index, self-trigger crossings) are
 * held fixed, as reverse-mode AD through the reference does. */
int lsim_backward(lsim_ctx *ctx, const lsim_params *p, const lsim_segments *seg, int32_t n_events, void *ws,
                  size_t ws_bytes, int64_t max_rows, const lsim_fwd_out *fwd, const float *d_adc, const float *d_z,
                  void *bwd_ws, size_t bwd_ws_bytes, double *d_params, void *stream);

/* As lsim_backward, plus a cotangent on the waveforms themselves: the gradient of
 * sum(d_adc * adc) + sum(d_z * hit_z) + sum(d_wfs * wfs), d_wfs [max_rows][nout] float32 (device),
 * the rows and ticks of the forward's wfs output (the reference differentiates simulate_wfs's
 * waveforms as well, sim_jax.py:689-736).  The row cotangent becomes a step function with one
 * piece per tick, so this costs more than lsim_backward. */
int lsim_backward_wfs(lsim_ctx *ctx, const lsim_params *p, const lsim_segments *seg, int32_t n_events, void *ws,
                      size_t ws_bytes, int64_t max_rows, const lsim_fwd_out *fwd, const float *d_adc, const float *d_z,
                      const float *d_wfs, void *bwd_ws, size_t bwd_ws_bytes, double *d_params, void *stream);

/* Fit losses: event-blocked Gaussian-kernel sums of the hit-cloud MMD in mse_adc
 * (src/larndsim/losses_jax.py:13-39 rbf_kernel / mmd, :58-80 mse_adc) and of the intensity of
 * nll_loss (:329-361), without forming the kernel matrices:
 * out[i][k] = sum_{j = lo[i]}^{hi[i]-1} exp(-|x_i - y_j|^2 / (2 sigma^2)) c[j][k].
 * x [nx][4] and y [ny][4] float32 4-D points (3-D callers set the fourth coordinate to 0), c [ny][4]
 * float32 coefficients, lo/hi [nx]
 * int32 with 0 <= lo[i] <= hi[i] <= ny (the points of x_i's event in y, both clouds sorted by
 * event), out [nx][4] float64.  Distances and exponentials in float32, sums in float64. */
int lsim_rbf_sums(const float *x, int64_t nx, const int32_t *lo, const int32_t *hi, const float *y, const float *c,
                  int64_t ny, float sigma, double *out, void *stream);

/* The same sums over the whole event of each x_i, walking only the cells of a uniform grid that can
 * hold a non-zero float32 kernel value (the reference forms the full kernel matrix, losses_jax.py:13-39;
 * the pairs left out are those whose exp(-d^2 / (2 sigma^2)) is exactly 0 in float32, so the sums are
 * those of every pair of the event, in another order).
 * grid [9] float32 device array: lo (x, y, z), cell size h (x, y, z), cell counts G (x, y, z) as floats;
 * the cell of a point is clamp(floor((v - lo) / h), 0, G - 1) per axis.  y and c [ny][4] sorted by
 * y_key [ny] int64 = ((event * Gx + cx) * Gy + cy) * Gz + cz, ascending.  x [nx][4] with x_event [nx]
 * int32, any order (sorting x by the same key keeps each wave's points close).  dims 3 (the fourth
 * coordinate is ignored) or 4.  Non-finite coordinates are outside the contract.  out [nx][4] float64. */
int lsim_rbf_sums_grid(const float *x, const int32_t *x_event, int64_t nx, const float *y, const float *c,
                       const int64_t *y_key, int64_t ny, const float *grid, int32_t dims, float sigma, double *out,
                       void *stream);

/* Parametrized current model (simulate_parametrized, src/larndsim/sim_jax.py:339-372), step 1: the
 * electrons of the drifted segments and their pixel pairs.  mode 0: the transverse bins of
 * apply_tran_diff (detsim_jax.py:356-373) on bin_edges [n_edges] float64 (device), n_edges - 1 bins per
 * axis, (n_edges - 1)^2 electrons per segment; mode 1: one electron per segment spread by the caller's
 * standard-normal draws normals [n][3] float32 (mc_diff, detsim_jax.py:377-399; z is spread only when
 * diffusion == 0, i.e. the current model carries none).  Outputs, E = n * electrons per segment, in
 * (segment, sub-electron) order: ex, ey (cm), et0 (drift time to the anode, us), ene (e-), esig (the
 * longitudinal spread in time, us) float64 [E], eplane int32 [E], and pair_pids int64 [E][(2n+1)^2], the
 * pixel ids around each electron's pixel in (a, b) row-major order, -1 outside the plane
 * (get_pixels, detsim_jax.py:483-496).  Float64 throughout. */
int lsim_param_electrons(lsim_ctx *c, const lsim_params *p, const lsim_segments *seg, int32_t mode, int32_t diffusion,
                         const float *normals, const double *bin_edges, int32_t n_edges, double *ex, double *ey,
                         double *et0, double *ene, double *esig, int32_t *eplane, int64_t *pair_pids, void *stream);

/* Step 2: the pixel waveforms wfs [n_rows][n_ticks - 1] float32 of the sorted pixel set pixel_ids [n_rows]
 * (current_mc, detsim_jax.py:623-639, with current_model / current_model_diff :544-622, and
 * accumulate_signals_parametrized :209-228): row r adds, for the pairs pair_order[row_offsets[r] ..
 * row_offsets[r + 1]) (pair index = electron * (2n+1)^2 + offset, in ascending order within a row), the
 * electron's charge times the induced current at the int(5 / t_sampling) + 1 sample times of
 * linspace(0, 5) us, at waveform ticks t0_tick - samples + k (ticks outside the waveform are dropped, as
 * the reference's column 0 is).  diffusion != 0: the exponentially modified Gaussian tails of
 * current_model_diff with the electron's spread.  Rows with a negative id stay zero.  Float64 sums in a
 * fixed order (deterministic).  The sample count must be 2 .. 63 (LSIM_E_ARG otherwise). */
int lsim_param_rows(lsim_ctx *c, const lsim_params *p, const double *ex, const double *ey, const double *et0,
                    const double *ene, const double *esig, const int64_t *pixel_ids, int64_t n_rows,
                    const int32_t *pair_order, const int32_t *row_offsets, int32_t diffusion, float *wfs, void *stream);

/* Reverse-mode derivative of the parametrized model (what jax.value_and_grad takes through
 * ParametrizedSimulation.predict, optimize/strategies.py:136-149 -> simulate_parametrized,
 * src/larndsim/sim_jax.py:339-372): d loss / d (Ab, kb, eField, lifetime, vdrift, long_diff, tran_diff)
 * into d_params [7] float64 (device), from cotangents on the readout grid d_adc [n_rows][H] and
 * d_z [n_rows][H] (may be NULL) and on the waveforms d_wfs [n_rows][n_ticks - 1] (may be NULL).
 * The arguments are those of the forward: the segments, mode, diffusion, normals and bin_edges given
 * to lsim_param_electrons and the electrons it wrote; the rows given to lsim_param_rows; fee_tape, the
 * tape lsim_fee_forward wrote for those rows' waveforms (out->tape).  Pixel ids, the whole-tick split
 * of the arrival time and the readout's decisions are held at their forward values.  Ticks whose
 * cotangent is exactly zero contribute nothing (the reference propagates a NaN of an overflowing
 * pair there).  Deterministic.  lsim_param_backward_query: workspace bytes for E electrons and
 * n_rows rows. */
int lsim_param_backward_query(lsim_ctx *c, int64_t n_electrons, int64_t n_rows, size_t *ws_bytes);
int lsim_param_backward(lsim_ctx *c, const lsim_params *p, const lsim_segments *seg, int32_t mode, int32_t diffusion,
                        const float *normals, const double *bin_edges, int32_t n_edges, const double *ex,
                        const double *ey, const double *et0, const double *ene, const double *esig,
                        const int64_t *pixel_ids, int64_t n_rows, const int32_t *pair_order, const int32_t *row_offsets,
                        const void *fee_tape, const float *d_adc, const float *d_z, const float *d_wfs, void *ws,
                        size_t ws_bytes, double *d_params, void *stream);

/* The distance beyond which lsim_rbf_sums_grid leaves pairs out (1.01 sqrt(212) sigma); callers size
 * the grid's cells from it (any cell size is correct; about a third of it culls well). */
float lsim_rbf_cull_radius(float sigma);

/* Fit loss: event-blocked nearest neighbour of the chamfer distance (losses_jax.py:195-229
 * chamfer_distance_3d_og, :231-249 chamfer_3d): for each x_i, the index j in [lo[i], hi[i]) of the
 * y point at the smallest squared distance (dx^2 + dy^2) + dz^2 (float32; the first j on ties, as
 * argmin) and that distance.  x [nx][4], y [ny][4] float32 (x, y, z, unused), lo/hi as for
 * lsim_rbf_sums.  The first n_origin queries also consider the zero-weight padding point at the
 * origin that the reference's pad_size padding adds after the real points: when it is strictly
 * nearer, idx[i] = -1.  No candidate: idx[i] = -1, dmin[i] = +inf.  idx [nx] int32, dmin [nx] float32. */
int lsim_nearest(const float *x, int64_t nx, const int32_t *lo, const int32_t *hi, const float *y, int64_t ny,
                 int64_t n_origin, int32_t *idx, float *dmin, void *stream);

/* ---- Track chopping and batch padding (the data step before the path) ------------------------ */

/* Column indices of the float32 (N, n_fields) track array (the reference's `fields.index`). */
typedef struct {
  int32_t n_fields;                     /* F: floats per row, 1 .. 64 */
  int32_t start[3], end[3], mid[3];     /* x_start y_start z_start, x_end y_end z_end, x y z */
  int32_t dE, dx;
  int32_t event;                        /* eventID column, or -1 */
  int32_t n_zero, zero[8];              /* set to 0 on invalidated rows (n_electrons, dE, dEdx, dx, long_diff, tran_diff) */
  int32_t n_neg_one, neg_one[8];        /* set to -1 on invalidated rows (trackID, pixel_plane) */
} lsim_chop_layout;

/* Workspace bytes of lsim_chop_count / lsim_chop_fill for n_tracks track rows. */
int lsim_chop_query(int64_t n_tracks, size_t *ws_bytes);

/* Step counts of chop_tracks (optimize/dataio.py:63-106, `nsteps = max(ceil(length / precision), 1)`,
 * float32 length) and their exclusive scan, kept in the workspace.  info [3] device int64:
 * info[0] = total output rows, info[1] = tracks whose length is not finite or needs more than
 * 2^31 - 1 steps (counted as one step; the reference fails on them), info[2] = 0. */
int lsim_chop_count(const lsim_chop_layout *layout, const float *tracks, int64_t n_tracks, double precision,
                    void *ws, size_t ws_bytes, int64_t *info, void *stream);

/* chop_tracks (optimize/dataio.py:63-106) into out [n_rows][F] float32, bit-identical to the
 * reference under NumPy 2 rounding, followed by the batch padding of TracksDataset.pad_batch
 * (:357-373): rows n_steps .. n_rows - 1 are zero rows invalidated as _invalidate_rows (:347-355)
 * does (eventID -1, zero columns 0, -1 columns -1).  n_steps must be info[0] of the matching
 * lsim_chop_count call and n_rows >= n_steps.
 * Optional event remap (remap_event_ids_to_local, :47-61): when src_events [n_tracks] is not NULL,
 * each track's eventID becomes the index of src_events[i] in the sorted batch_events
 * [n_batch_events]; ids not found are written as -1 and counted in info[2] (the reference raises).
 * When n_local_events >= 0, rows whose eventID e >= 0 has trunc(e) >= n_local_events are
 * invalidated (pad_batch with a batch index); -1 disables it. */
int lsim_chop_fill(const lsim_chop_layout *layout, const float *tracks, int64_t n_tracks, double precision,
                   const void *ws, size_t ws_bytes, int64_t n_steps, int64_t n_rows, const int64_t *src_events,
                   const int64_t *batch_events, int64_t n_batch_events, int64_t n_local_events, float *out,
                   int64_t *info, void *stream);

/* ---- Probabilistic self-trigger readout (the step after the path in the llhd fits) ------------- */

/* Workspace bytes of lsim_prob_fee. */
int lsim_prob_fee_query(int64_t n_rows, int32_t n_ticks, int32_t n_paths, int32_t n_hits, size_t *ws_bytes);

/* get_adc_values_average_noise_vmap (src/larndsim/fee_jax.py:334-461): the beam search of n_paths
 * reset paths per waveform row with Gaussian reset noise noise_sigma (e-), n_hits steps.
 * wfs [n_rows][n_ticks] float32 (current per tick; charge = wfs * t_sampling).  Outputs
 * log_prob and charge [n_rows][n_hits][n_ticks - 1] float32: the log probability that hit s
 * starts at tick t and its expected charge (e-).  interval = round((3 + ADC_HOLD_DELAY)
 * CLOCK_CYCLE / t_sampling) ticks.  As in the reference, all rows stop together once no row's
 * paths keep a total probability above stop_threshold; stopped steps hold -1000 and 0.
 * n_ticks <= 2048, n_paths <= 32.  The workspace keeps the per-row tables and the paths of every
 * step for lsim_prob_fee_backward. */
int lsim_prob_fee(const float *wfs, int64_t n_rows, int32_t n_ticks, int32_t n_paths, int32_t n_hits, int32_t interval,
                  float threshold, float noise_sigma, float t_sampling, float stop_threshold, void *ws,
                  size_t ws_bytes, float *log_prob, float *charge, void *stream);

/* The reverse-mode derivative of lsim_prob_fee (what jax.grad takes through
 * get_adc_values_average_noise_vmap, fee_jax.py:334-461, for the llhd and probabilistic loss
 * strategies, optimize/strategies.py:107-134, 182-567): from cotangents d_log_prob and d_charge
 * [n_rows][n_hits][n_ticks - 1] of its two outputs, d_wfs [n_rows][n_ticks] float32.  The path
 * selections (top_k, under stop_gradient in the reference) and the stop are held fixed.  ws and
 * log_prob are those of the lsim_prob_fee call being differentiated, same sizes and settings. */
int lsim_prob_fee_backward(int64_t n_rows, int32_t n_ticks, int32_t n_paths, int32_t n_hits, int32_t interval,
                           float threshold, float noise_sigma, float t_sampling, const void *ws, size_t ws_bytes,
                           const float *log_prob, const float *d_log_prob, const float *d_charge, float *d_wfs,
                           void *stream);

/* ---- Soft-DTW fit losses ------------------------------------------------------------------------- */

/* Workspace bytes of lsim_sdtw for sequences of n and m values (n * m <= 2^34; 8 bytes per cell). */
int lsim_sdtw_query(int64_t n, int64_t m, size_t *ws_bytes);

/* Soft dynamic time warping between x [n] and y [m] (float32, device) with the squared-Euclidean
 * ground cost, as SoftDTW(gamma).pairwise (src/larndsim/softdtw_jax.py:205-275) computes it for the
 * sdtw losses (losses_jax.py:309-327): value [1] float32 (device).  When grad_x is not NULL it
 * receives d value / d x [n] (float64, device).  Deterministic: a tiled anti-diagonal wavefront,
 * one launch per tile diagonal, fixed-order gradient sums. */
int lsim_sdtw(const float *x, int64_t n, const float *y, int64_t m, float gamma, void *ws, size_t ws_bytes,
              float *value, double *grad_x, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* LSIM_H */
