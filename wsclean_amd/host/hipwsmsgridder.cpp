#include "hipwsmsgridder.h"

#include <cmath>
#include <stdexcept>
#include <string>

namespace wsclean_amd {

HipWSMSGridder::HipWSMSGridder(const Settings& settings, const Resources& resources)
    : MSGridderBase(settings),
      resources_(resources),
      kernelSize_(settings.antialiasingKernelSize),
      overSampling_(settings.overSamplingFactor) {}

void HipWSMSGridder::check(hws_status s, const hws_gridder* g) {
  if (s != HWS_OK)
    throw std::runtime_error(std::string("MI355X w-stacking gridder: ") +
                             hws_status_string(s) + ": " + hws_last_error(g));
}

void HipWSMSGridder::makeContext(Context& ctx) const {
  hws_config c{};
  c.width = ActualInversionWidth();
  c.height = ActualInversionHeight();
  c.pixel_size_x = ActualPixelSizeX();
  c.pixel_size_y = ActualPixelSizeY();
  c.kernel_size = kernelSize_;
  c.oversampling = overSampling_;
  c.kernel_mode = int(GetGridMode());
  c.precision = int(settings_.hipPrecision);
  c.device = settings_.hipDevice;
  c.n_images = 1;
  c.complex_mask = IsComplex() ? 1u : 0u;
  c.l_shift = LShift();
  c.m_shift = MShift();
  // WSMSGridder rotates samples when the phase centre is shifted
  // (gridding/wsmsgridder.cpp:185-190, 328-332)
  c.rotate_visibilities = HasDenormalPhaseCentre() ? 1 : 0;
  c.tile_size = settings_.hipTileSize;
  check(hws_create(&c, &ctx.g), nullptr);
}

size_t HipWSMSGridder::getSuggestedWGridSize() const {
  size_t wWidth, wHeight;
  if (settings_.widthForNWCalculation != 0 || settings_.heightForNWCalculation != 0) {
    wWidth = settings_.widthForNWCalculation;
    wHeight = settings_.heightForNWCalculation;
  } else {
    wWidth = TrimWidth();
    wHeight = TrimHeight();
  }
  const double maxL = wWidth * PixelSizeX() * 0.5 + std::fabs(LShift()),
               maxM = wHeight * PixelSizeY() * 0.5 + std::fabs(MShift()),
               lmSq = maxL * maxL + maxM * maxM;
  const double cMinW = IsComplex() ? -maxW_ : minW_;
  const double radiansForAllLayers = (lmSq < 1.0)
                                         ? 2 * M_PI * (maxW_ - cMinW) * (1.0 - std::sqrt(1.0 - lmSq))
                                         : 2 * M_PI * (maxW_ - cMinW);
  size_t suggested = size_t(std::ceil(radiansForAllLayers * settings_.nWLayersFactor));
  if (suggested == 0) suggested = 1;
  if (suggested < resources_.NCpus()) {
    const double memoryRequired = double(resources_.NCpus()) * double(sizeof(float)) *
                                  double(ActualInversionWidth() * ActualInversionHeight());
    if (4.0 * memoryRequired < double(resources_.Memory())) suggested = resources_.NCpus();
  }
  return suggested;
}

void HipWSMSGridder::Invert() {
  std::vector<MSData> msDataVector;
  initializeMSDataVector(msDataVector);
  Context ctx;
  makeContext(ctx);
  check(hws_prepare_w_layers(ctx.g, ActualWGridSize(), settings_.hipMaxLayerMemory, minW_,
                             maxW_, &nPasses_),
        ctx.g);

  // Read and weight every row once (readAndWeightVisibilities<1>); a sample
  // counts towards the weight sums when its w falls on a layer of some pass,
  // which is what the per-pass selection of gridMeasurementSet adds up to
  // (gridding/wsmsgridder.cpp:173-178).
  resetVisibilityCounters();
  struct Rows {
    std::vector<double> uvw, wl;
    std::vector<std::complex<float>> data;
  };
  std::vector<Rows> rows(msDataVector.size());
  for (size_t i = 0; i != msDataVector.size(); ++i) {
    MSProvider& provider = *msDataVector[i].msProvider;
    const size_t nchan = provider.NChannels();
    Rows& r = rows[i];
    for (size_t ch = 0; ch != nchan; ++ch) r.wl.push_back(provider.ChannelWavelength(ch));
    std::vector<std::complex<float>> rowData(nchan);
    std::vector<float> weights(nchan);
    std::unique_ptr<bool[]> selected(new bool[nchan]);
    std::unique_ptr<MSReader> reader = provider.MakeReader();
    while (reader->CurrentRowAvailable()) {
      double uvw[3];
      reader->ReadMeta(uvw[0], uvw[1], uvw[2]);
      for (size_t ch = 0; ch != nchan; ++ch)
        selected[ch] =
            hws_w_to_layer(ctx.g, uvw[2] / r.wl[ch], IsComplex() ? 1 : 0) < ActualWGridSize();
      readAndWeightVisibilities(*reader, provider, uvw, rowData.data(), weights.data(),
                                selected.get());
      r.uvw.insert(r.uvw.end(), uvw, uvw + 3);
      r.data.insert(r.data.end(), rowData.begin(), rowData.end());
      reader->NextInputRow();
    }
  }
  for (size_t pass = 0; pass != nPasses_; ++pass) {
    check(hws_start_inversion_pass(ctx.g, pass), ctx.g);
    for (const Rows& r : rows) {
      const size_t nrow = r.uvw.size() / 3;
      check(hws_add_data(ctx.g, nrow, r.wl.size(), r.uvw.data(), r.wl.data(),
                         reinterpret_cast<const float*>(r.data.data())),
            ctx.g);
    }
    check(hws_finish_inversion_pass(ctx.g), ctx.g);
  }
  Image real(ActualInversionWidth(), ActualInversionHeight());
  Image imag;
  if (IsComplex()) imag = Image(ActualInversionWidth(), ActualInversionHeight());
  check(hws_finalize_image(ctx.g, 1.0 / totalWeight(), real.Data(),
                           IsComplex() ? imag.Data() : nullptr),
        ctx.g);
  realImage_ = std::move(real);
  imaginaryImage_ = std::move(imag);
  if (TrimWidth() != ImageWidth() || TrimHeight() != ImageHeight()) {
    realImage_ = realImage_.Trim(TrimWidth(), TrimHeight());
    if (IsComplex()) imaginaryImage_ = imaginaryImage_.Trim(TrimWidth(), TrimHeight());
  }
}

void HipWSMSGridder::Predict(std::vector<Image>&& images) {
  if (images.size() != 2 && IsComplex())
    throw std::runtime_error("Missing imaginary in complex prediction");
  if (images.size() != 1 && !IsComplex())
    throw std::runtime_error("Imaginary specified in non-complex prediction");
  std::vector<MSData> msDataVector;
  initializeMSDataVector(msDataVector);
  Context ctx;
  makeContext(ctx);
  check(hws_prepare_w_layers(ctx.g, ActualWGridSize(), settings_.hipMaxLayerMemory, minW_,
                             maxW_, &nPasses_),
        ctx.g);
  if (TrimWidth() != ImageWidth() || TrimHeight() != ImageHeight())
    for (Image& image : images) image = image.Untrim(ImageWidth(), ImageHeight());
  check(hws_initialize_prediction(ctx.g, images[0].Data(),
                                  images.size() == 2 ? images[1].Data() : nullptr),
        ctx.g);
  for (size_t pass = 0; pass != nPasses_; ++pass) {
    check(hws_start_prediction_pass(ctx.g, pass), ctx.g);
    for (size_t i = 0; i != msDataVector.size(); ++i) {
      MSProvider& provider = *msDataVector[i].msProvider;
      provider.ResetWritePosition();
      const size_t nchan = provider.NChannels();
      std::vector<double> wl(nchan), uvw;
      for (size_t ch = 0; ch != nchan; ++ch) wl[ch] = provider.ChannelWavelength(ch);
      std::unique_ptr<MSReader> reader = provider.MakeReader();
      while (reader->CurrentRowAvailable()) {
        double u, v, w;
        reader->ReadMeta(u, v, w);
        uvw.push_back(u);
        uvw.push_back(v);
        uvw.push_back(w);
        reader->NextInputRow();
      }
      const size_t nrow = uvw.size() / 3;
      std::vector<std::complex<float>> out(nrow * nchan);
      check(hws_sample_data(ctx.g, nrow, nchan, uvw.data(), wl.data(),
                            reinterpret_cast<float*>(out.data())),
            ctx.g);
      // rows in order; samples of other passes are NaN and skipped
      for (size_t r = 0; r != nrow; ++r)
        writeVisibilities(provider, msDataVector[i].msIndex, &out[r * nchan]);
    }
  }
}

}  // namespace wsclean_amd
