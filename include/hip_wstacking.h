/*
 * hip_wstacking.h - C ABI of the MI355X (gfx950) w-stacking gridder/degridder.
 *
 * The entry points mirror the call protocol of the reference core
 * WStackingGridder<T> (gridding/wstackinggridder.h:44-84):
 *
 *   invert:  hws_create -> hws_prepare_w_layers -> per pass
 *            { hws_start_inversion_pass; hws_add_data*; hws_finish_inversion_pass }
 *            -> hws_finalize_image
 *   predict: hws_create -> hws_prepare_w_layers -> hws_initialize_prediction
 *            -> per pass { hws_start_prediction_pass; hws_sample_data* }
 *
 * plus one-call device-resident forms (hws_invert_device / hws_predict_device)
 * that run every pass, all channels and all images of a batch in one sequence
 * of launches, with the uv-tile binning of the visibilities kept in a plan
 * that invert and predict share.
 *
 * Conventions (identical to the reference):
 *   - images are row-major, x fastest; pixel (x, y) has
 *     l = (W/2 - x) * dl + l_shift and m = (y - H/2) * dm + m_shift;
 *   - uvw are in metres, one triple per row; wavelengths in metres per
 *     channel (the value BandData::ChannelWavelength(ch) returns);
 *   - visibilities are complex64 (interleaved float re, im) laid out
 *     [row][channel][image];
 *   - a context grids n_images images that share uvw (e.g. the correlations
 *     XX, XY, YX, YY); bit i of complex_mask marks image i as a complex image
 *     (the reference's SetIsComplex, used for XY/YX,
 *     scheduling/griddingtaskmanager.cpp:106-107);
 *   - predicted samples that no pass covers, or that fall off the grid, are
 *     NaN (gridding/wstackinggridder.cpp:626-627, 677-678, 685-687).
 *
 * Every entry point returns an hws_status; on failure hws_last_error() gives a
 * message. Nothing throws across this boundary. A context is used by one
 * thread at a time; several contexts may run concurrently on one device.
 */
#ifndef HIP_WSTACKING_H
#define HIP_WSTACKING_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum hws_status {
  HWS_OK = 0,
  HWS_ERR_INVALID_ARGUMENT = 1,
  HWS_ERR_HIP = 2,
  HWS_ERR_OUT_OF_MEMORY = 3,
  HWS_ERR_UNSUPPORTED = 4,
  HWS_ERR_BAD_STATE = 5
} hws_status;

/* Values of GriddingKernelMode (gridding/gridmode.h:6-44). */
typedef enum hws_kernel_mode {
  HWS_KERNEL_NEAREST_NEIGHBOUR = 0,
  HWS_KERNEL_KAISER_BESSEL = 1,
  HWS_KERNEL_KAISER_BESSEL_NO_SINC = 2,
  HWS_KERNEL_RECTANGULAR = 3,
  HWS_KERNEL_GAUSSIAN = 4,
  HWS_KERNEL_GAUSSIAN_NO_SINC = 5,
  HWS_KERNEL_BLACKMAN_NUTTALL = 6,
  HWS_KERNEL_BLACKMAN_NUTTALL_NO_SINC = 7,
  HWS_KERNEL_BLACKMAN_HARRIS = 8 /* rejected, as in the reference */
} hws_kernel_mode;

typedef enum hws_precision {
  /* One launch per channel and image; uvw scaling, sample position, layer
   * index and image-domain phase in FP32; FP32 kernel tables; every tap is a
   * global FP32 atomic. The plain GPU form. */
  HWS_PRECISION_PLAIN_FP32 = 0,
  /* Batched uv-tile binned gridding with LDS subgrids; FP64 uvw scaling,
   * position, layer index and phase arguments (as the reference computes
   * them); FP32 kernel tables; FP64 sin/cos for the w-correction; Kahan
   * compensated FP32 layer->image sums. */
  HWS_PRECISION_REFERENCE = 1,
  /* As REFERENCE, but FP16 kernel tables (hi + lo pair), FP32 accumulation,
   * and w-correction phases reduced modulo one turn in FP64 before an FP32
   * sincospi. */
  HWS_PRECISION_MIXED = 2
} hws_precision;

typedef struct hws_config {
  size_t width, height;         /* even, 7-smooth, <= 16384 */
  double pixel_size_x, pixel_size_y; /* radians */
  size_t kernel_size;           /* odd, <= 15 */
  size_t oversampling;          /* <= 1023 */
  int kernel_mode;              /* hws_kernel_mode */
  int precision;                /* hws_precision */
  int device;                   /* HIP device ordinal */
  size_t n_images;              /* 1..4 */
  uint32_t complex_mask;        /* bit i: image i is complex */
  double l_shift, m_shift;      /* SetDenormalPhaseCentre */
  /* 1: multiply visibilities by exp(-2 pi i (u l_shift + v m_shift)/lambda)
   * before gridding and by the conjugate after degridding, as WSMSGridder
   * does (gridding/wsmsgridder.cpp:185-190, 328-332). 0: only the n' shift
   * of WStackingGridder::SetDenormalPhaseCentre. */
  int rotate_visibilities;
  /* uv tile edge for binning; 0 = the edge closest to 16 that divides the
   * image (16: the matrix-core grid kernel; 8: the VALU cell kernel; other
   * edges: the private-tile kernel over origin-sorted records). */
  size_t tile_size;
  /* Records per work item of the binned grid/degrid kernels: a (layer, tile)
   * bin with more records is split into several items (whose tiles are
   * summed in a fixed order afterwards). 0 = chosen from the sample count. */
  size_t work_chunk;
} hws_config;

typedef struct hws_gridder hws_gridder;

/* Counters of the last invert/predict on a context. */
typedef struct hws_stats {
  uint64_t samples_in;        /* (row, channel) samples handed in */
  uint64_t samples_gridded;   /* accepted on a layer of a pass and on the grid */
  uint64_t samples_off_grid;  /* rejected by -W/2 < x <= W/2, -H/2 < y <= H/2 */
  uint64_t samples_off_layer; /* w outside [minW - dw/2, maxW + dw/2] */
  uint64_t bins_used;         /* non-empty (layer, tile) bins */
  uint64_t chunks;            /* work items of the binned grid/degrid kernels */
  double ms_bin, ms_grid, ms_fft, ms_project, ms_total; /* device times */
  /* Per image kind (0: real images, 1: complex images), over all layers of the
   * plan: (layer, uv tile) bins holding records, and tiles a degridding
   * subgrid reads. The sparse layer transforms read (backward) or write
   * (forward) only these tiles in their first (second) pass. */
  uint64_t tiles_total;
  uint64_t tiles_occupied[2];
  uint64_t tiles_needed[2];
  /* (layer, tile row) pairs over all layers of the plan, and those with a
   * tile holding records: the rows the backward column pass reads. */
  uint64_t tile_rows_total;
  uint64_t tile_rows_live[2];
  /* (layer, tile column) pairs, and those with a tile a degridding subgrid
   * reads: the columns the forward column pass transforms. */
  uint64_t tile_cols_total;
  uint64_t tile_cols_needed[2];
  /* Weighted contexts (hws_set_weighting), per image, over the samples of the
   * plan whose w-layer the image grids: the counters MSGridderBase keeps
   * (gridding/msgridderbase.h:197-213, msgridderbase.cpp:640-659). */
  double total_weight[4];
  double visibility_weight_sum[4];
  double max_gridded_weight[4];
  uint64_t gridded_visibility_count[4];
} hws_stats;

/* Replaces: WStackingGridder<T>::WStackingGridder (gridding/wstackinggridder.h:120-122,
 * cpp:16-41) together with SetGridMode/SetIsComplex/SetDenormalPhaseCentre
 * (h:470-508). */
hws_status hws_create(const hws_config* config, hws_gridder** out);
/* Replaces: ~WStackingGridder (cpp:53-61). */
void hws_destroy(hws_gridder* g);
/* Message for the last failed call on g (or a global message if g is NULL). */
const char* hws_last_error(const hws_gridder* g);
const char* hws_status_string(hws_status s);

/* Replaces: WStackingGridder::PrepareWLayers (h:158, cpp:63-116). max_mem is a
 * device-memory budget in bytes for the layers of one pass (0 = 75% of the
 * free device memory). Returns the pass count in *n_passes. Layer ranges per
 * pass follow layerRangeStart (h:566-568). */
hws_status hws_prepare_w_layers(hws_gridder* g, size_t n_w_layers,
                                double max_mem, double min_w, double max_w,
                                size_t* n_passes);
/* Replaces: WToLayer (h:182-193) for image kind is_complex. Returns
 * SIZE_MAX where the reference casts a negative value to size_t. */
size_t hws_w_to_layer(const hws_gridder* g, double w_lambda, int is_complex);
/* Replaces: LayerToW (h:202-212). */
double hws_layer_to_w(const hws_gridder* g, size_t layer, int is_complex);
/* Replaces: layerRangeStart (h:566-568): layers [*start, *end) of a pass. */
hws_status hws_pass_layer_range(const hws_gridder* g, size_t pass,
                                size_t* start, size_t* end);

/* Replaces: StartInversionPass (h:311, cpp:127-137). */
hws_status hws_start_inversion_pass(hws_gridder* g, size_t pass);
/* Replaces: AddData (h:288-289, cpp:971-982) for a block of rows, i.e.
 * AddDataSample (cpp:517-597) for every (row, channel) with
 * u = u_m / wavelength etc. Host pointers; vis must be pre-weighted.
 * Each call plans its own block of rows, which replaces a plan made by
 * hws_plan_device (the *_planned calls then return HWS_ERR_BAD_STATE until
 * hws_plan_device is called again). */
hws_status hws_add_data(hws_gridder* g, size_t nrow, size_t nchan,
                        const double* uvw_m, const double* wavelengths,
                        const float* vis);
/* Replaces: AddDataSample (h:300-301) for n samples whose uvw are already in
 * wavelengths: uvw_lambda [n][3], vis [n][n_images]. */
hws_status hws_add_samples(hws_gridder* g, size_t n, const double* uvw_lambda,
                           const float* vis);
/* Replaces: FinishInversionPass (h:319, cpp:323-336). */
hws_status hws_finish_inversion_pass(hws_gridder* g);
/* Replaces: FinalizeImage (h:326, cpp:691-715) + RealImageFloat /
 * ImaginaryImageFloat (h:412-435, cpp:996-1022). real: [n_images][H][W];
 * imag: [n_images][H][W] or NULL (filled for complex images only). */
hws_status hws_finalize_image(hws_gridder* g, double multiplication_factor,
                              float* real, float* imag);

/* Replaces: InitializePrediction (h:341-357, cpp:829-849). real/imag as in
 * hws_finalize_image; imag is required iff some image is complex. */
hws_status hws_initialize_prediction(hws_gridder* g, const float* real,
                                     const float* imag);
/* Replaces: StartPredictionPass (h:366, cpp:139-157). */
hws_status hws_start_prediction_pass(hws_gridder* g, size_t pass);
/* Replaces: SampleData (h:385-386, cpp:984-994) for a block of rows; vis_out
 * [nrow][nchan][n_images]; NaN where this pass predicts nothing. Replaces a
 * plan made by hws_plan_device, as hws_add_data does. */
hws_status hws_sample_data(hws_gridder* g, size_t nrow, size_t nchan,
                           const double* uvw_m, const double* wavelengths,
                           float* vis_out);
/* Replaces: SampleDataSample (h:400-401) for n samples in wavelengths. */
hws_status hws_sample_samples(hws_gridder* g, size_t n,
                              const double* uvw_lambda, float* vis_out);

/* Replaces: GetGriddedUVLayer (h:518-520): copies layer `layer` (relative to
 * the current pass) of image `image` as [H][W] complex64 to host memory. */
hws_status hws_get_gridded_uv_layer(hws_gridder* g, size_t image,
                                    size_t layer, float* out);
/* Replaces: the static GetKernel (h:530-531, cpp:387-423). */
hws_status hws_get_kernel(int kernel_mode, size_t oversampling, size_t size,
                          double* kernel);

/* ---- device-resident batched forms (no reference counterpart as single
 * calls; they run the same protocol as above over every pass) ----
 * All pointers are device pointers. Work is ordered on the stream set with
 * hws_set_stream (default: a non-blocking stream the context owns; NULL
 * selects HIP's null stream); calls return once the work is enqueued, except
 * where a host-visible result is needed.
 * The context must have been prepared with hws_prepare_w_layers. */
hws_status hws_set_stream(hws_gridder* g, void* hip_stream);
/* Bin (row, channel) samples by (w-layer, uv tile) for both image kinds; the
 * plan is kept in the context and used by the *_planned calls below, until
 * hws_prepare_w_layers, hws_add_data or hws_sample_data replaces it. The
 * plan reads d_uvw_m and d_wavelengths again at every planned call. */
hws_status hws_plan_device(hws_gridder* g, size_t nrow, size_t nchan,
                           const double* d_uvw_m, const double* d_wavelengths);
/* Whole invert over all passes: real/imag [n_images][H][W] device pointers,
 * imag may be NULL when no image is complex. */
hws_status hws_invert_planned_device(hws_gridder* g, const float* d_vis,
                                     double multiplication_factor,
                                     float* d_real, float* d_imag);
/* Whole predict over all passes: d_vis_out [nrow][nchan][n_images]. */
hws_status hws_predict_planned_device(hws_gridder* g, const float* d_real,
                                      const float* d_imag, float* d_vis_out);
/* hws_plan_device followed by the planned call. */
hws_status hws_invert_device(hws_gridder* g, size_t nrow, size_t nchan,
                             const double* d_uvw_m,
                             const double* d_wavelengths, const float* d_vis,
                             double multiplication_factor, float* d_real,
                             float* d_imag);
hws_status hws_predict_device(hws_gridder* g, size_t nrow, size_t nchan,
                              const double* d_uvw_m,
                              const double* d_wavelengths, const float* d_real,
                              const float* d_imag, float* d_vis_out);
/* Counters and device timings of the last invert/predict. */
hws_status hws_get_stats(const hws_gridder* g, hws_stats* out);
/* Wait for all work of the context. */
hws_status hws_synchronize(hws_gridder* g);

/* ---- imaging weights on the device ----
 * An ImageWeights (structures/imageweights.h:19-129) whose grid lives on one
 * device, gridded from device-resident rows, and the weighting a context
 * applies to the visibilities it grids. */
typedef enum hws_weight_mode {
  HWS_WEIGHT_NATURAL = 0,
  HWS_WEIGHT_UNIFORM = 1,
  HWS_WEIGHT_BRIGGS = 2
} hws_weight_mode;

typedef struct hws_weights_config {
  int mode;                 /* hws_weight_mode */
  double briggs_robustness; /* R of Briggs weighting */
  size_t width, height;     /* image size; the grid is width x height / 2 */
  double pixel_scale_x, pixel_scale_y; /* radians */
  int weights_as_taper;     /* non-zero visibility weights count as 1 */
  double super_weight;      /* grid cells are super_weight pixels wide */
  int device;               /* HIP device ordinal */
} hws_weights_config;

typedef struct hws_weights hws_weights;

/* Replaces: ImageWeights::ImageWeights (structures/imageweights.cpp:33-49):
 * a zeroed (width/super_weight) x (height/super_weight)/2 grid, sizes rounded
 * and made even as the reference makes them. */
hws_status hws_weights_create(const hws_weights_config* config, hws_weights** out);
void hws_weights_destroy(hws_weights* w);
/* Replaces: ImageWeights::Grid(MSProvider&, const BandData&)
 * (structures/imageweights.cpp:75-111) for rows in device memory: every
 * weight d_weights[row][chan][pol] (NULL: 1) is added to the cell of
 * (u, v) = uvw / lambda in the v >= 0 half plane and to the total. Sums are
 * bit-equal to the reference's row-order sums whenever they are exact in FP64
 * (e.g. unit or few-digit weights), within FP64 rounding otherwise. Enqueued
 * on hip_stream; may be called once per MeasurementSet. */
hws_status hws_weights_grid_device(hws_weights* w, void* hip_stream, size_t nrow,
                                   size_t nchan, size_t npol, const double* d_uvw_m,
                                   const double* d_wavelengths, const float* d_weights);
/* Replaces: ImageWeights::FinishGridding (structures/imageweights.cpp:113-144):
 * natural 1 / uniform 1 / sum / Briggs 1 / (1 + sum s^2). Synchronises
 * hip_stream; fails with HWS_ERR_BAD_STATE when called twice. */
hws_status hws_weights_finish(hws_weights* w, void* hip_stream);
/* Replaces: ImageWeights::Width/Height (structures/imageweights.h:65-66). */
hws_status hws_weights_shape(const hws_weights* w, size_t* width, size_t* height);
/* Replaces: ImageWeights::GetGrid (structures/imageweights.h:60): the grid
 * ([height/2][width] doubles, host memory; NULL to skip) and the total of the
 * gridded weights (NULL to skip). Synchronises the device. */
hws_status hws_weights_get_grid(const hws_weights* w, double* grid, double* total_sum);

/* Replaces: the weighting of MSGridderBase::readAndWeightVisibilities
 * (gridding/msgridderbase.cpp:563-589, 640-659) on the planned device path:
 * every visibility is gridded times (visibility weight x imaging weight), the
 * visibility weight d_vis_weights[row][chan][image] (NULL: 1) taken as given
 * (mode 0), squared (1) or as 1 where non-zero (2), the imaging weight
 * GetWeight(u, v) of `weights` (finished, same device; NULL detaches). Each
 * invert then multiplies image i by multiplication_factor /
 * total_weight[i] (WSMSGridder finalizes with 1 / totalWeight,
 * gridding/wsmsgridder.cpp:436), and hws_get_stats reports the sums over the
 * samples whose w-layer the image grids. Applies to hws_plan_device and the
 * *_planned_device calls and graphs; the host-pointer protocol grids its rows
 * as given. The weights and d_vis_weights must outlive the context's use of
 * them. */
hws_status hws_set_weighting(hws_gridder* g, const hws_weights* weights,
                             const float* d_vis_weights, int visibility_weighting_mode);

/* ---- staged host rows (a host caller of the planned device path) ----
 * The MSGridderBase plugin reads and weights its rows on the host and stages
 * them here batch by batch: each batch is copied into one of two pinned host
 * buffers and uploaded asynchronously while the caller reads the next one.
 * hws_stage_end plans the staged samples once; hws_invert_staged and
 * hws_predict_staged then run every pass on the device, and convert between
 * the context's (inversion) size and the caller's image and trimmed sizes on
 * the device (hws_resample_device, hws_trim_device, hws_untrim_device), as
 * WSMSGridder::Invert and ::Predict do (gridding/wsmsgridder.cpp:446-484,
 * 510-551). No single reference counterpart: together they replace
 * WSMSGridder's gridMeasurementSet / predictMeasurementSet row loops
 * (gridding/wsmsgridder.cpp:130-253, 255-365) with their lanes and passes.
 *
 * hws_stage_begin reserves device room for max_rows rows of nchan channels
 * (wavelengths: host, metres) with (with_visibilities != 0) or without
 * visibilities; hws_stage_rows appends nrow rows: uvw_m [nrow][3] host, vis
 * [nrow][nchan][n_images] complex64 host (pre-weighted; NULL without
 * visibilities). */
hws_status hws_stage_begin(hws_gridder* g, size_t max_rows, size_t nchan,
                           const double* wavelengths, int with_visibilities);
hws_status hws_stage_rows(hws_gridder* g, size_t nrow, const double* uvw_m, const float* vis);
hws_status hws_stage_end(hws_gridder* g);
/* Invert of the staged visibilities, times multiplication_factor (and each
 * image by its own weight sum when a weighting is attached), resampled to
 * image_width x image_height when that differs from the context size, then
 * trimmed to trim_width x trim_height: real (imag for complex images, may be
 * NULL) host [n_images][trim_height][trim_width]. Synchronous. */
hws_status hws_invert_staged(hws_gridder* g, double multiplication_factor,
                             size_t image_width, size_t image_height, size_t trim_width,
                             size_t trim_height, float* real, float* imag);
/* Predict at the staged rows from host models of trim_width x trim_height
 * (untrimmed to image_width x image_height, resampled to the context size):
 * vis_out host [rows][nchan][n_images] complex64, NaN where not predicted.
 * Synchronous. */
hws_status hws_predict_staged(hws_gridder* g, const float* real, const float* imag,
                              size_t trim_width, size_t trim_height, size_t image_width,
                              size_t image_height, float* vis_out);
/* The same image ends for the host-pointer protocol: hws_finalize_image
 * followed by the resampling and trim; hws_initialize_prediction preceded by
 * the untrim and resampling. */
hws_status hws_finalize_image_resampled(hws_gridder* g, double multiplication_factor,
                                        size_t image_width, size_t image_height,
                                        size_t trim_width, size_t trim_height, float* real,
                                        float* imag);
hws_status hws_initialize_prediction_resampled(hws_gridder* g, const float* real,
                                               const float* imag, size_t trim_width,
                                               size_t trim_height, size_t image_width,
                                               size_t image_height);

/* ---- graph capture of the device-resident forms (no reference
 * counterpart) ----
 * hws_graph_capture_invert / _predict capture the whole
 * hws_invert_planned_device / hws_predict_planned_device sequence (clear,
 * grid or image -> layers, transforms, corrections, degrid) of a planned
 * context in a batched precision with its layers in one pass into an
 * executable HIP graph, so a
 * major cycle replays it with one launch. The graph keeps the pointers given
 * here and the context's plan buffers: it is valid until the context is
 * planned again (hws_plan_device, hws_add_data, hws_sample_data, the staged
 * calls), re-prepared, given another weighting (hws_set_weighting) or
 * destroyed; hws_graph_launch then returns HWS_ERR_BAD_STATE rather than
 * replay stale pointers. Phase timings (hws_get_stats) are not recorded by
 * graph launches. Captures run on the context's own stream;
 * hws_graph_launch enqueues a replay on hip_stream (NULL: the null stream). */
typedef struct hws_graph hws_graph;
hws_status hws_graph_capture_invert(hws_gridder* g, const float* d_vis,
                                    double multiplication_factor, float* d_real,
                                    float* d_imag, hws_graph** out);
hws_status hws_graph_capture_predict(hws_gridder* g, const float* d_real,
                                     const float* d_imag, float* d_vis_out,
                                     hws_graph** out);
/* The same for n contexts at once (e.g. the facets of one major cycle, each
 * context with its own image size and phase centre): one graph that runs
 * every context's invert (predict) in order, replayed with one launch. All
 * contexts must be on one device, planned and single-pass; arrays hold one
 * entry per context (d_imag may be NULL, or hold NULL entries). */
hws_status hws_graph_capture_invert_batch(hws_gridder* const* gridders, size_t n,
                                          const float* const* d_vis,
                                          const double* multiplication_factors,
                                          float* const* d_real, float* const* d_imag,
                                          hws_graph** out);
hws_status hws_graph_capture_predict_batch(hws_gridder* const* gridders, size_t n,
                                           const float* const* d_real,
                                           const float* const* d_imag,
                                           float* const* d_vis_out, hws_graph** out);
hws_status hws_graph_launch(hws_graph* graph, void* hip_stream);
void hws_graph_destroy(hws_graph* graph);

/* ---- FP64 direct Fourier transform (no context) ----
 * Pixel (x, y) of the width x height image has l = (W/2 - x) dl + l_shift,
 * m = (y - H/2) dm + m_shift, n' = sqrt(1 - l^2 - m^2) - 1 (0 outside the
 * horizon); phi = u l + v m + w n' with uvw in wavelengths (uvw_m / lambda).
 * Samples are the (row, channel) pairs; sample (r, c) of d_vis is the
 * complex64 at index (r * nchan + c) * vis_stride (vis_stride = n_images
 * and an offset pointer select one image of a [row][chan][image] array).
 * Work is enqueued on hip_stream (NULL: the null stream) of `device`.
 *
 * Replaces: DirectMSGridder<double>::Invert / gridSample /
 * initializeSqrtLMLookupTable (gridding/directmsgridder.cpp:16-74, 76-114,
 * 175-196): d_image[y * W + x] = mult * sum Re(V exp(-2 pi i phi)), FP64.
 * Returns once the transform is enqueued. */
hws_status hws_direct_invert_device(int device, void* hip_stream, size_t width,
                                    size_t height, double dl, double dm,
                                    double l_shift, double m_shift, size_t nrow,
                                    size_t nchan, const double* d_uvw_m,
                                    const double* d_wavelengths,
                                    const float* d_vis, size_t vis_stride,
                                    double multiplication_factor,
                                    double* d_image);
/* The adjoint of hws_direct_invert_device (with mult = 1), which the
 * reference leaves unimplemented (DirectMSGridder::Predict throws,
 * gridding/directmsgridder.cpp:167-171): sample = sum over the non-zero
 * pixels of I(x, y) exp(+2 pi i phi), summed in FP64, stored as complex64.
 * Allocates scratch for the non-zero pixel list and returns after the
 * transform has completed. */
hws_status hws_direct_predict_device(int device, void* hip_stream, size_t width,
                                     size_t height, double dl, double dm,
                                     double l_shift, double m_shift,
                                     const float* d_image, size_t nrow,
                                     size_t nchan, const double* d_uvw_m,
                                     const double* d_wavelengths,
                                     float* d_vis_out, size_t vis_stride);

/* Host-pointer forms of the two transforms above, for host callers (the
 * DirectFT plugin gridder): arguments as there with vis_stride 1, the image
 * FP64 [height][width] host memory (invert) or float32 (predict), uvw,
 * wavelengths and visibilities host memory. They upload, transform and
 * download on the null stream of `device` and return when done. */
hws_status hws_direct_invert(int device, size_t width, size_t height, double dl, double dm,
                             double l_shift, double m_shift, size_t nrow, size_t nchan,
                             const double* uvw_m, const double* wavelengths, const float* vis,
                             double multiplication_factor, double* image);
hws_status hws_direct_predict(int device, size_t width, size_t height, double dl, double dm,
                              double l_shift, double m_shift, const float* image, size_t nrow,
                              size_t nchan, const double* uvw_m, const double* wavelengths,
                              float* vis_out);

/* ---- small inversion and trimming on the device (no context) ----
 * Images are float32 [n_images][height][width] device arrays; work is ordered
 * on hip_stream of `device`. Sizes are even, at most 65536.
 *
 * Replaces: schaapcommon::fft::Resampler::Resample as WSMSGridder applies it
 * for small inversion (gridding/wsmsgridder.cpp:452-474 interpolates the
 * inverted in_w x in_h image up to the image size, :521-551 decimates the
 * model down to the inversion size): the real-to-complex transform of each
 * image, the low-frequency block both sizes share copied into the output's
 * half spectrum, the complex-to-real transform and 1 / (in_w in_h),
 * evaluated in FP64 as two separable resampling matrices. Allocates scratch
 * and returns after the resampling has completed. */
hws_status hws_resample_device(int device, void* hip_stream, const float* d_in, size_t in_w,
                               size_t in_h, float* d_out, size_t out_w, size_t out_h,
                               size_t n_images);
/* Replaces: aocommon::Image::Trim (gridding/wsmsgridder.cpp:476-484): the
 * trim_width x trim_height centre of each image (offset ((w - tw)/2,
 * (h - th)/2)). Returns once enqueued. */
hws_status hws_trim_device(int device, void* hip_stream, const float* d_in, size_t width,
                           size_t height, float* d_out, size_t trim_width, size_t trim_height,
                           size_t n_images);
/* Replaces: aocommon::Image::Untrim (gridding/wsmsgridder.cpp:516-520): each
 * trimmed image placed at the centre of a zeroed width x height image.
 * Returns once enqueued. */
hws_status hws_untrim_device(int device, void* hip_stream, const float* d_in, size_t trim_width,
                             size_t trim_height, float* d_out, size_t width, size_t height,
                             size_t n_images);

/* hws_direct_predict_device without allocation: no counterpart in the
 * reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc): the same
 * kernels and results as hws_direct_predict_device, but their scratch (the
 * non-zero pixel list and block counts) is a fixed array the library keeps on
 * each device until the process ends, so images are limited to 2^20 pixels
 * (32 MiB of scratch per device; larger images are rejected with
 * HWS_ERR_INVALID_ARGUMENT). Nothing is allocated or freed and the call
 * returns once the transform is enqueued on hip_stream. Calls on one device
 * share that array: calls on one stream are ordered, calls that overlap in
 * time on one device (other streams or threads) race. */
hws_status hws_direct_predict_device_noalloc(int device, void* hip_stream, size_t width,
                                             size_t height, double dl, double dm,
                                             double l_shift, double m_shift,
                                             const float* d_image, size_t nrow, size_t nchan,
                                             const double* d_uvw_m,
                                             const double* d_wavelengths, float* d_vis_out,
                                             size_t vis_stride);

#ifdef __cplusplus
}
#endif

#endif /* HIP_WSTACKING_H */
