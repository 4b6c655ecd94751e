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
 * u = u_m / wavelength etc. Host pointers; vis must be pre-weighted. */
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
 * [nrow][nchan][n_images]; NaN where this pass predicts nothing. */
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
 * plan is kept in the context and used by the *_planned calls below. */
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

/* ---- graph capture of the device-resident forms (no reference
 * counterpart) ----
 * hws_graph_capture_invert / _predict capture the whole
 * hws_invert_planned_device / hws_predict_planned_device sequence (clear,
 * grid or image -> layers, transforms, corrections, degrid) of a planned
 * context in a batched precision with its layers in one pass into an
 * executable HIP graph, so a
 * major cycle replays it with one launch. The graph keeps the pointers given
 * here and the context's plan buffers: it is valid until the context is
 * planned again, re-prepared or destroyed. Phase timings (hws_get_stats) are
 * not recorded by graph launches. Captures run on the context's own stream;
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
