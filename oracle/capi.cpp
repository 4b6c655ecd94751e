// ORACLE / TEST INFRASTRUCTURE ONLY (see oracle.h): extern "C" wrappers.
#include "oracle.h"

#include "msref.h"
#include "wsref.h"

#include <complex>
#include <cstring>
#include <exception>
#include <memory>

using namespace oracle;

namespace {
struct Handle {
  int prec;
  std::unique_ptr<WStackingRef<float>> f;
  std::unique_ptr<WStackingRef<double>> d;
};

template <typename Fn>
auto dispatch(void* g, Fn fn) {
  Handle* h = static_cast<Handle*>(g);
  if (h->prec == 0) return fn(*h->f);
  return fn(*h->d);
}

GridderSettings toSettings(const orc_ms_settings* s) {
  GridderSettings o;
  o.width = s->width;
  o.height = s->height;
  o.trimWidth = s->trim_width;
  o.trimHeight = s->trim_height;
  o.pixelSizeX = s->dl;
  o.pixelSizeY = s->dm;
  o.kernelSize = s->kernel_size;
  o.overSampling = s->oversampling;
  o.gridMode = KernelMode(s->grid_mode);
  o.nWLayers = s->nw_layers;
  o.nWLayersFactor = s->nw_factor;
  o.lShift = s->l_shift;
  o.mShift = s->m_shift;
  o.isComplex = s->is_complex != 0;
  o.nCpus = s->ncpus;
  o.memory = s->memory;
  o.wLimit = s->w_limit;
  o.psfMode = s->psf_mode != 0;
  o.smallInversion = s->small_inversion != 0;
  return o;
}
}  // namespace

extern "C" {

int orc_make_kernel(int mode, size_t oversampling, size_t size, double* kernel) {
  return MakeKernel1D(KernelMode(mode), kernel, oversampling, size) ? 0 : -1;
}
double orc_bessel0(double x, double precision) { return Bessel0(x, precision); }

void* orc_wsg_create(int prec, size_t width, size_t height, double dl,
                     double dm, size_t nthreads, size_t kernel_size,
                     size_t oversampling) {
  Handle* h = new Handle();
  h->prec = prec;
  if (prec == 0)
    h->f.reset(new WStackingRef<float>(width, height, dl, dm, nthreads,
                                       kernel_size, oversampling));
  else
    h->d.reset(new WStackingRef<double>(width, height, dl, dm, nthreads,
                                        kernel_size, oversampling));
  return h;
}
void orc_wsg_destroy(void* g) { delete static_cast<Handle*>(g); }
int orc_wsg_set_grid_mode(void* g, int mode) {
  try {
    dispatch(g, [&](auto& x) {
      x.SetGridMode(KernelMode(mode));
      return 0;
    });
    return 0;
  } catch (std::exception&) {
    return -1;
  }
}
void orc_wsg_set_is_complex(void* g, int c) {
  dispatch(g, [&](auto& x) {
    x.SetIsComplex(c != 0);
    return 0;
  });
}
void orc_wsg_set_denormal_phase_centre(void* g, double dl, double dm) {
  dispatch(g, [&](auto& x) {
    x.SetDenormalPhaseCentre(dl, dm);
    return 0;
  });
}
void orc_wsg_prepare_w_layers(void* g, size_t nw, double max_mem, double min_w,
                              double max_w) {
  dispatch(g, [&](auto& x) {
    x.PrepareWLayers(nw, max_mem, min_w, max_w);
    return 0;
  });
}
size_t orc_wsg_n_passes(void* g) {
  return dispatch(g, [&](auto& x) { return x.NPasses(); });
}
size_t orc_wsg_n_fft_threads(void* g) {
  return dispatch(g, [&](auto& x) { return x.NFFTThreads(); });
}
size_t orc_wsg_w_to_layer(void* g, double w) {
  return dispatch(g, [&](auto& x) { return x.WToLayer(w); });
}
double orc_wsg_layer_to_w(void* g, size_t layer) {
  return dispatch(g, [&](auto& x) { return x.LayerToW(layer); });
}
int orc_wsg_is_in_layer_range(void* g, double w) {
  return dispatch(g, [&](auto& x) { return int(x.IsInLayerRange(w)); });
}
int orc_wsg_is_in_layer_range2(void* g, double w1, double w2) {
  return dispatch(g, [&](auto& x) { return int(x.IsInLayerRange(w1, w2)); });
}
void orc_wsg_start_inversion_pass(void* g, size_t pass) {
  dispatch(g, [&](auto& x) {
    x.StartInversionPass(pass);
    return 0;
  });
}
void orc_wsg_add_data_sample(void* g, float re, float im, double u, double v,
                             double w) {
  dispatch(g, [&](auto& x) {
    x.AddDataSample(std::complex<float>(re, im), u, v, w);
    return 0;
  });
}
void orc_wsg_add_samples(void* g, size_t n, const double* uvw,
                         const float* vis) {
  dispatch(g, [&](auto& x) {
    for (size_t i = 0; i != n; ++i)
      x.AddDataSample(std::complex<float>(vis[2 * i], vis[2 * i + 1]),
                      uvw[3 * i], uvw[3 * i + 1], uvw[3 * i + 2]);
    return 0;
  });
}
void orc_wsg_finish_inversion_pass(void* g) {
  dispatch(g, [&](auto& x) {
    x.FinishInversionPass();
    return 0;
  });
}
void orc_wsg_finalize_image(void* g, double mult) {
  dispatch(g, [&](auto& x) {
    x.FinalizeImage(mult);
    return 0;
  });
}
void orc_wsg_get_image(void* g, double* real, double* imag) {
  dispatch(g, [&](auto& x) {
    const size_t n = x.Width() * x.Height();
    for (size_t i = 0; i != n; ++i) real[i] = x.RealImage()[i];
    if (imag && x.IsComplex())
      for (size_t i = 0; i != n; ++i) imag[i] = x.ImaginaryImage()[i];
    return 0;
  });
}
void orc_wsg_get_layer(void* g, size_t layer, double* out) {
  dispatch(g, [&](auto& x) {
    const size_t n = x.Width() * x.Height();
    const auto* p = x.GetGriddedUVLayer(layer);
    for (size_t i = 0; i != n; ++i) {
      out[2 * i] = p[i].real();
      out[2 * i + 1] = p[i].imag();
    }
    return 0;
  });
}
void orc_wsg_initialize_prediction(void* g, const float* real,
                                   const float* imag) {
  dispatch(g, [&](auto& x) {
    x.InitializePrediction(real, imag);
    return 0;
  });
}
void orc_wsg_start_prediction_pass(void* g, size_t pass) {
  dispatch(g, [&](auto& x) {
    x.StartPredictionPass(pass);
    return 0;
  });
}
void orc_wsg_sample_data_sample(void* g, double u, double v, double w,
                                float* out2) {
  dispatch(g, [&](auto& x) {
    std::complex<float> val;
    x.SampleDataSample(val, u, v, w);
    out2[0] = val.real();
    out2[1] = val.imag();
    return 0;
  });
}
void orc_wsg_sample_samples(void* g, size_t n, const double* uvw, float* out) {
  dispatch(g, [&](auto& x) {
    for (size_t i = 0; i != n; ++i) {
      std::complex<float> val;
      x.SampleDataSample(val, uvw[3 * i], uvw[3 * i + 1], uvw[3 * i + 2]);
      out[2 * i] = val.real();
      out[2 * i + 1] = val.imag();
    }
    return 0;
  });
}
void orc_wsg_kernel_correction(void* g, size_t axis, double* out) {
  dispatch(g, [&](auto& x) {
    const std::vector<double> d = x.KernelCorrection(axis);
    std::memcpy(out, d.data(), d.size() * sizeof(double));
    return 0;
  });
}
void orc_wsg_sqrt_lm_table(void* g, double* out) {
  dispatch(g, [&](auto& x) {
    const auto& t = x.SqrtLMTable();
    for (size_t i = 0; i != t.size(); ++i) out[i] = t[i];
    return 0;
  });
}

void* orc_iw_create(int mode, double robust, size_t width, size_t height,
                    double dl, double dm, int weights_as_taper,
                    double super_weight) {
  return new ImageWeightsRef(WeightingMode(mode), robust, width, height, dl, dm,
                             weights_as_taper != 0, super_weight);
}
void orc_iw_destroy(void* iw) { delete static_cast<ImageWeightsRef*>(iw); }
void orc_iw_grid(void* iw, size_t nrow, size_t nchan, size_t npol,
                 const double* uvw_m, const double* wavelengths,
                 const float* weights) {
  static_cast<ImageWeightsRef*>(iw)->Grid(nrow, nchan, npol, uvw_m, wavelengths,
                                          weights);
}
void orc_iw_finish(void* iw) { static_cast<ImageWeightsRef*>(iw)->FinishGridding(); }
double orc_iw_get_weight(void* iw, double u, double v) {
  return static_cast<ImageWeightsRef*>(iw)->GetWeight(u, v);
}
void orc_iw_lookup(void* iw, size_t nrow, size_t nchan, const double* uvw_m,
                   const double* wavelengths, float* out) {
  const ImageWeightsRef* w = static_cast<ImageWeightsRef*>(iw);
  for (size_t r = 0; r != nrow; ++r)
    for (size_t c = 0; c != nchan; ++c)
      out[r * nchan + c] = float(w->GetWeight(uvw_m[3 * r] / wavelengths[c],
                                              uvw_m[3 * r + 1] / wavelengths[c]));
}
size_t orc_iw_width(void* iw) { return static_cast<ImageWeightsRef*>(iw)->Width(); }
size_t orc_iw_height(void* iw) { return static_cast<ImageWeightsRef*>(iw)->Height(); }
void orc_iw_grid_values(void* iw, double* out) {
  const auto& g = static_cast<ImageWeightsRef*>(iw)->GridValues();
  std::memcpy(out, g.data(), g.size() * sizeof(double));
}

int orc_ms_invert(int prec, const orc_ms_settings* s, size_t nrow,
                  size_t nchan, const double* uvw_m, const double* wavelengths,
                  const float* data, const float* weights, void* iw,
                  float* real_out, float* imag_out, orc_invert_stats* stats) {
  try {
    GridderSettings gs = toSettings(s);
    MsView ms;
    ms.nrow = nrow;
    ms.nchan = nchan;
    ms.uvw_m = uvw_m;
    ms.wavelengths = wavelengths;
    ms.data = reinterpret_cast<const std::complex<float>*>(data);
    ms.weights = weights;
    std::unique_ptr<ImageWeightsRef> natural;
    const ImageWeightsRef* w = static_cast<ImageWeightsRef*>(iw);
    if (!w) {
      natural.reset(new ImageWeightsRef(WeightingMode::Natural, 0, gs.width,
                                        gs.height, gs.pixelSizeX,
                                        gs.pixelSizeY, false, 1.0));
      natural->Grid(nrow, nchan, 1, uvw_m, wavelengths, weights);
      natural->FinishGridding();
      w = natural.get();
    }
    InvertResult r = prec == 0 ? InvertMs<float>(gs, ms, *w)
                               : InvertMs<double>(gs, ms, *w);
    std::memcpy(real_out, r.real.data(), r.real.size() * sizeof(float));
    if (imag_out && !r.imag.empty())
      std::memcpy(imag_out, r.imag.data(), r.imag.size() * sizeof(float));
    if (stats) {
      stats->total_weight = r.totalWeight;
      stats->visibility_weight_sum = r.visibilityWeightSum;
      stats->max_gridded_weight = r.maxGriddedWeight;
      stats->gridded_count = r.griddedVisibilityCount;
      stats->min_w = r.minW;
      stats->max_w = r.maxW;
      stats->nw = r.actualWGridSize;
      stats->n_passes = r.nPasses;
    }
    return 0;
  } catch (std::exception&) {
    return -1;
  }
}

int orc_ms_predict(int prec, const orc_ms_settings* s, size_t nrow,
                   size_t nchan, const double* uvw_m,
                   const double* wavelengths, const float* weights, void* iw,
                   const float* real, const float* imag, float* out,
                   orc_invert_stats* stats) {
  try {
    GridderSettings gs = toSettings(s);
    MsView ms;
    ms.nrow = nrow;
    ms.nchan = nchan;
    ms.uvw_m = uvw_m;
    ms.wavelengths = wavelengths;
    ms.weights = weights;
    std::unique_ptr<ImageWeightsRef> natural;
    const ImageWeightsRef* w = static_cast<ImageWeightsRef*>(iw);
    if (!w) {
      natural.reset(new ImageWeightsRef(WeightingMode::Natural, 0, gs.width,
                                        gs.height, gs.pixelSizeX,
                                        gs.pixelSizeY, false, 1.0));
      natural->Grid(nrow, nchan, 1, uvw_m, wavelengths, weights);
      natural->FinishGridding();
      w = natural.get();
    }
    double minW = 0, maxW = 0;
    size_t nPasses = 0;
    auto* o = reinterpret_cast<std::complex<float>*>(out);
    if (prec == 0)
      PredictMs<float>(gs, ms, *w, real, imag, o, &minW, &maxW, &nPasses);
    else
      PredictMs<double>(gs, ms, *w, real, imag, o, &minW, &maxW, &nPasses);
    if (stats) {
      std::memset(stats, 0, sizeof(*stats));
      stats->min_w = minW;
      stats->max_w = maxW;
      stats->n_passes = nPasses;
    }
    return 0;
  } catch (std::exception&) {
    return -1;
  }
}

size_t orc_suggested_w_grid_size(const orc_ms_settings* s, double min_w,
                                 double max_w) {
  return SuggestedWGridSize(toSettings(s), min_w, max_w);
}
void orc_calculate_fft_size(size_t size, double pixel_scale, double beam_size,
                            size_t* minimal, size_t* optimal) {
  CalculateFFTSize(size, pixel_scale, beam_size, *minimal, *optimal);
}

void orc_dft_invert(size_t width, size_t height, double dl, double dm,
                    double l_shift, double m_shift, size_t n,
                    const double* uvw, const float* vis, double factor,
                    double* image, size_t nthreads) {
  DftInvert(width, height, dl, dm, l_shift, m_shift, n, uvw,
            reinterpret_cast<const std::complex<float>*>(vis), factor, image,
            nthreads);
}
void orc_dft_predict(size_t width, size_t height, double dl, double dm,
                     double l_shift, double m_shift, size_t n,
                     const double* uvw, const double* image, double* vis,
                     size_t nthreads) {
  DftPredict(width, height, dl, dm, l_shift, m_shift, n, uvw, image,
             reinterpret_cast<std::complex<double>*>(vis), nthreads);
}

void orc_fft2d(size_t width, size_t height, int sign, double* data) {
  Fft2D<double> fft(width, height, sign);
  fft.Transform(reinterpret_cast<std::complex<double>*>(data));
}

void orc_calculate_w_limits(const orc_ms_settings* s, size_t nrow, size_t nchan,
                            const double* uvw_m, const double* wavelengths,
                            const float* weights, void* iw, double* min_w,
                            double* max_w, double* max_baseline_uvw) {
  GridderSettings gs = toSettings(s);
  MsView ms;
  ms.nrow = nrow;
  ms.nchan = nchan;
  ms.uvw_m = uvw_m;
  ms.wavelengths = wavelengths;
  ms.weights = weights;
  CalculateWLimits(gs, ms, *static_cast<ImageWeightsRef*>(iw), *min_w, *max_w,
                   *max_baseline_uvw);
}

void orc_resample(const double* in, size_t in_w, size_t in_h, double* out, size_t out_w,
                  size_t out_h) {
  std::vector<double> r = Resample(in, in_w, in_h, out_w, out_h);
  std::memcpy(out, r.data(), r.size() * sizeof(double));
}

void orc_trim_image(const float* in, size_t w, size_t h, size_t tw, size_t th,
                    float* out) {
  std::vector<float> v(in, in + w * h);
  std::vector<float> r = trimImage(v, w, h, tw, th);
  std::memcpy(out, r.data(), r.size() * sizeof(float));
}

void orc_untrim_image(const float* in, size_t tw, size_t th, size_t w, size_t h,
                      float* out) {
  std::vector<float> r = untrimImage(in, tw, th, w, h);
  std::memcpy(out, r.data(), r.size() * sizeof(float));
}

}  // extern "C"
