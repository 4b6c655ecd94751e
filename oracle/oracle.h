/*
 * ORACLE / TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker for the HIP w-stacking core, never the thing
 * that is measured as the product or shipped with it. Only tests/, the
 * smoke() check in __graft_entry__.py and the cpu_baseline leg of bench.py may
 * load it. It holds:
 *   - a CPU restatement of the reference WStackingGridder<float|double>
 *     (gridding/wstackinggridder.{h,cpp}), with a project-written FFT in place
 *     of FFTW and a direct cosine sum in place of FFTW's REDFT01;
 *   - a restatement of the WSMSGridder / MSGridderBase host pipeline
 *     (weighting, w-limits, suggested w-layer count, layer-routed worker
 *     lanes, ordered prediction writer) used as the timed CPU baseline;
 *   - the FP64 direct Fourier transform of DirectMSGridder<double>
 *     (gridding/directmsgridder.cpp:16-114) and its exact adjoint (predict),
 *     the accuracy oracle every result's error is stated against;
 *   - ImageWeights (structures/imageweights.{h,cpp}).
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   orc_make_kernel        closed form k1d[mid] = 1, symmetry, sinc*KB identity
 *   orc_wsg_* layer map    round-trip |LayerToW(WToLayer(w)) - w| <= dw/2, pass
 *                          partition covers [0, nW) exactly once
 *   orc_wsg_add/sample     single-sample footprints equal kx (x) ky with wrap;
 *                          NN deposits a delta; adjoint identity grid/degrid
 *   FFT                    textbook DFT on small sizes, round trip = N * id
 *   kernel correction      direct cosine sum vs an independent DCT-III via FFT
 *   invert end to end      FP64 DFT, PSF peak = 1, float vs double
 *   predict end to end     FP64 DFT predict and dot-product adjoint test
 *   DFT invert/predict     brute-force small cases, adjoint identity
 *   ImageWeights           closed-form sums on synthetic grids
 * Bit-level FFTW results and the reference's thread-order image sum are not
 * reproducible without the reference build (parity unpinned at bit level).
 */
#ifndef WSCLEAN_AMD_ORACLE_H
#define WSCLEAN_AMD_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- kernels (gridding/wstackinggridder.cpp:387-515) ---- */
int orc_make_kernel(int mode, size_t oversampling, size_t size, double* kernel);
double orc_bessel0(double x, double precision);

/* ---- WStackingGridder restatement; prec 0 = float, 1 = double ---- */
void* orc_wsg_create(int prec, size_t width, size_t height, double dl,
                     double dm, size_t nthreads, size_t kernel_size,
                     size_t oversampling);
void orc_wsg_destroy(void* g);
int orc_wsg_set_grid_mode(void* g, int mode);
void orc_wsg_set_is_complex(void* g, int is_complex);
void orc_wsg_set_denormal_phase_centre(void* g, double dl, double dm);
void orc_wsg_prepare_w_layers(void* g, size_t nw, double max_mem, double min_w,
                              double max_w);
size_t orc_wsg_n_passes(void* g);
size_t orc_wsg_n_fft_threads(void* g);
size_t orc_wsg_w_to_layer(void* g, double w);
double orc_wsg_layer_to_w(void* g, size_t layer);
int orc_wsg_is_in_layer_range(void* g, double w);
int orc_wsg_is_in_layer_range2(void* g, double w1, double w2);
void orc_wsg_start_inversion_pass(void* g, size_t pass);
void orc_wsg_add_data_sample(void* g, float re, float im, double u, double v,
                             double w);
/* n samples: uvw [n*3] in wavelengths, vis [n*2] interleaved floats */
void orc_wsg_add_samples(void* g, size_t n, const double* uvw,
                         const float* vis);
void orc_wsg_finish_inversion_pass(void* g);
void orc_wsg_finalize_image(void* g, double mult);
/* copies the (real, imag) image as double; imag may be NULL */
void orc_wsg_get_image(void* g, double* real, double* imag);
/* copies layer (relative to the current pass) as interleaved doubles */
void orc_wsg_get_layer(void* g, size_t layer, double* out);
void orc_wsg_initialize_prediction(void* g, const float* real,
                                   const float* imag);
void orc_wsg_start_prediction_pass(void* g, size_t pass);
void orc_wsg_sample_data_sample(void* g, double u, double v, double w,
                                float* out2);
void orc_wsg_sample_samples(void* g, size_t n, const double* uvw, float* out);
/* D[k] for k = 0..n/2 (n/2+1 values) */
void orc_wsg_kernel_correction(void* g, size_t axis, double* out);
void orc_wsg_sqrt_lm_table(void* g, double* out);

/* ---- ImageWeights ---- */
void* orc_iw_create(int mode, double robust, size_t width, size_t height,
                    double dl, double dm, int weights_as_taper,
                    double super_weight);
void orc_iw_destroy(void* iw);
void orc_iw_grid(void* iw, size_t nrow, size_t nchan, size_t npol,
                 const double* uvw_m, const double* wavelengths,
                 const float* weights);
void orc_iw_finish(void* iw);
double orc_iw_get_weight(void* iw, double u, double v);
/* weights per (row, channel) for uvw_m/wavelengths */
void orc_iw_lookup(void* iw, size_t nrow, size_t nchan, const double* uvw_m,
                   const double* wavelengths, float* out);
size_t orc_iw_width(void* iw);
size_t orc_iw_height(void* iw);
void orc_iw_grid_values(void* iw, double* out);

/* ---- WSMSGridder pipeline (CPU baseline); prec 0 float, 1 double ---- */
typedef struct orc_ms_settings {
  size_t width, height, trim_width, trim_height;
  double dl, dm;
  size_t kernel_size, oversampling;
  int grid_mode;
  size_t nw_layers;
  double nw_factor;
  double l_shift, m_shift;
  int is_complex;
  size_t ncpus;
  double memory;
  double w_limit;
  int psf_mode;
  int small_inversion;
} orc_ms_settings;

typedef struct orc_invert_stats {
  double total_weight, visibility_weight_sum, max_gridded_weight;
  size_t gridded_count;
  double min_w, max_w;
  size_t nw, n_passes;
} orc_invert_stats;

/* data: [nrow*nchan*2] floats; weights [nrow*nchan] or NULL; iw may be NULL
 * (natural weights of 1 everywhere are then used only if iw==NULL). */
int orc_ms_invert(int prec, const orc_ms_settings* s, size_t nrow,
                  size_t nchan, const double* uvw_m, const double* wavelengths,
                  const float* data, const float* weights, void* iw,
                  float* real_out, float* imag_out, orc_invert_stats* stats);
int orc_ms_predict(int prec, const orc_ms_settings* s, size_t nrow,
                   size_t nchan, const double* uvw_m,
                   const double* wavelengths, const float* weights, void* iw,
                   const float* real, const float* imag, float* out,
                   orc_invert_stats* stats);
size_t orc_suggested_w_grid_size(const orc_ms_settings* s, double min_w,
                                 double max_w);
void orc_calculate_fft_size(size_t size, double pixel_scale, double beam_size,
                            size_t* minimal, size_t* optimal);

/* ---- FP64 DFT oracle ---- */
void orc_dft_invert(size_t width, size_t height, double dl, double dm,
                    double l_shift, double m_shift, size_t n,
                    const double* uvw, const float* vis, double factor,
                    double* image, size_t nthreads);
void orc_dft_predict(size_t width, size_t height, double dl, double dm,
                     double l_shift, double m_shift, size_t n,
                     const double* uvw, const double* image, double* vis,
                     size_t nthreads);

/* ---- FFT (for tests of the FFT itself) ---- */
void orc_fft2d(size_t width, size_t height, int sign, double* data);

/* msgridderbase.cpp:200-286: min/max |w| (lambda) over the samples with a
 * non-zero visibility weight (weights NULL: all 1) and imaging weight that
 * fall inside the uv window floor(x) > -W/2, ceil(x) < W/2 (likewise y), and
 * the longest such baseline in lambda; 0/0 when no sample qualifies. */
void orc_calculate_w_limits(const orc_ms_settings* s, size_t nrow, size_t nchan,
                            const double* uvw_m, const double* wavelengths,
                            const float* weights, void* iw, double* min_w,
                            double* max_w, double* max_baseline_uvw);
/* FFT resampling of a real image (schaapcommon::fft::Resampler as
 * WSMSGridder uses it for small inversion), FP64. */
void orc_resample(const double* in, size_t in_w, size_t in_h, double* out, size_t out_w,
                  size_t out_h);
/* Centre-keeping trim (w x h -> tw x th) and untrim with zero fill. */
void orc_trim_image(const float* in, size_t w, size_t h, size_t tw, size_t th,
                    float* out);
void orc_untrim_image(const float* in, size_t tw, size_t th, size_t w, size_t h,
                      float* out);

#ifdef __cplusplus
}
#endif

#endif
