// ORACLE / TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// CPU restatement of the host logic around the w-stacking core:
//  - ImageWeights (structures/imageweights.h:19-129, imageweights.cpp:33-144)
//  - MSGridderBase::calculateWLimits / calculateOverallMetaData /
//    readAndWeightVisibilities / rotateVisibilities
//    (gridding/msgridderbase.cpp:200-286, 370-446, 515-660, 692-704)
//  - WSMSGridder::Invert / Predict and their thread pipelines
//    (gridding/wsmsgridder.cpp:77-128, 130-365, 367-586)
//  - DirectMSGridder<double> (gridding/directmsgridder.cpp:16-114, 175-196)
//    and its exact adjoint for prediction (the reference throws there,
//    gridding/directmsgridder.cpp:168-172).
#ifndef ORACLE_MSREF_H
#define ORACLE_MSREF_H

#include "wsref.h"

#include <complex>
#include <cstddef>
#include <vector>

namespace oracle {

// structures/weightmode.h:12
enum class WeightingMode : int { Natural = 0, Uniform = 1, Briggs = 2 };

class ImageWeightsRef {
 public:
  ImageWeightsRef(WeightingMode mode, double briggsRobustness, size_t width,
                  size_t height, double pixelScaleX, double pixelScaleY,
                  bool weightsAsTaper, double superWeight);
  // structures/imageweights.cpp:75-111 for one in-memory measurement set:
  // uvw_m [nrow*3], wavelengths [nchan], weights [nrow*nchan*npol].
  void Grid(size_t nrow, size_t nchan, size_t npol, const double* uvw_m,
            const double* wavelengths, const float* weights);
  void FinishGridding();  // imageweights.cpp:113-144
  double GetWeight(double u, double v) const;  // imageweights.h:33, 96-104
  const std::vector<double>& GridValues() const { return grid_; }
  size_t Width() const { return width_; }
  size_t Height() const { return height_; }

 private:
  void uvToXY(double u, double v, int& x, int& y) const;
  bool isWithinLimits(int x, int y) const;
  WeightingMode mode_;
  double robust_;
  size_t width_, height_;
  double scaleX_, scaleY_;
  std::vector<double> grid_;
  double totalSum_ = 0.0;
  bool weightsAsTaper_;
};

// One in-memory measurement set holding a single correlation (the one being
// imaged), rows in reading order.
struct MsView {
  size_t nrow = 0, nchan = 0;
  const double* uvw_m = nullptr;        // [nrow][3]
  const double* wavelengths = nullptr;  // [nchan]
  const std::complex<float>* data = nullptr;  // [nrow][nchan]
  const float* weights = nullptr;             // [nrow][nchan]
};

struct GridderSettings {
  size_t width = 0, height = 0;          // ActualInversionWidth/Height
  size_t trimWidth = 0, trimHeight = 0;  // 0 = no trim
  double pixelSizeX = 0, pixelSizeY = 0;
  size_t kernelSize = 7, overSampling = 63;
  KernelMode gridMode = KernelMode::KaiserBessel;
  size_t nWLayers = 0;  // 0 = use the suggestion (msgridderbase.cpp:443-445)
  double nWLayersFactor = 1.0;
  double lShift = 0, mShift = 0;
  bool isComplex = false;
  size_t nCpus = 1;
  double memory = 0;  // bytes; PrepareWLayers receives 0.6 of it
  double wLimit = 0;  // percent
  bool psfMode = false;
  // small inversion (msgridderbase.cpp:409-439): width/height are then the
  // image size, and the inversion runs at the size that samples the
  // theoretical beam by two pixels, resampled to the image size
  bool smallInversion = false;
};

struct InvertResult {
  std::vector<float> real, imag;  // trimWidth x trimHeight
  double totalWeight = 0, visibilityWeightSum = 0, maxGriddedWeight = 0;
  size_t griddedVisibilityCount = 0;
  double minW = 0, maxW = 0;
  size_t actualWGridSize = 0, nPasses = 0;
};

// Double-precision (T=double) or single (T=float) instantiation of the
// reference pipeline; weights may be null (unit weights).
template <typename T>
InvertResult InvertMs(const GridderSettings& s, const MsView& ms,
                      const ImageWeightsRef& imageWeights);

// Writes predicted visibilities into out[nrow][nchan]. Samples that no pass
// predicts keep their initial value (the NaN-means-skip contract,
// msproviders/msprovider.cpp:456-657).
template <typename T>
void PredictMs(const GridderSettings& s, const MsView& ms,
               const ImageWeightsRef& imageWeights, const float* real,
               const float* imag, std::complex<float>* out, double* minW,
               double* maxW, size_t* nPassesOut);

// w-limits pass, msgridderbase.cpp:200-286 (one MS, single correlation).
void CalculateWLimits(const GridderSettings& s, const MsView& ms,
                      const ImageWeightsRef& w, double& minW, double& maxW,
                      double& maxBaselineUVW);

// Centre-keeping trim of a w x h image to tw x th, and its inverse with zero
// fill (aocommon::Image::Trim/Untrim as WSMSGridder uses them,
// gridding/wsmsgridder.cpp:476-484, 516-520): trimmed pixel (x, y) is pixel
// (x + (w - tw)/2, y + (h - th)/2) of the full image.
std::vector<float> trimImage(const std::vector<float>& in, size_t w, size_t h,
                             size_t tw, size_t th);
std::vector<float> untrimImage(const float* in, size_t tw, size_t th, size_t w,
                               size_t h);

// The FFT resampler WSMSGridder applies for small inversion
// (schaapcommon::fft::Resampler, gridding/wsmsgridder.cpp:452-474, 521-551;
// schaapcommon is not vendored in the reference tree): real-to-complex
// transform of the in_w x in_h image, the low-frequency block shared by both
// sizes copied into an out_w x out_h half spectrum (columns 0 .. min(in_w,
// out_w)/2, rows 0 .. min(in_h, out_h)/2 - 1 and the last min(in_h,
// out_h)/2 rows), complex-to-real transform of that, times
// 1 / (in_w in_h). FP64 throughout.
std::vector<double> Resample(const double* in, size_t inW, size_t inH, size_t outW,
                             size_t outH);

// wsmsgridder.cpp:77-128
size_t SuggestedWGridSize(const GridderSettings& s, double minW, double maxW);

// math/calculatefftsize.h:22-32
void CalculateFFTSize(size_t size, double pixelScale, double beamSize,
                      size_t& minimalSize, size_t& optimalSize);

// FP64 direct Fourier transform, gridding/directmsgridder.cpp:16-114. uvw in
// wavelengths, vis already weighted. Output image is width x height with the
// reference remap and multiplied by factor (1/totalWeight in the reference).
void DftInvert(size_t width, size_t height, double dl, double dm,
               double lShift, double mShift, size_t n, const double* uvw,
               const std::complex<float>* vis, double factor, double* image,
               size_t nThreads);
// Exact adjoint of DftInvert's sum (without the factor):
// V_k = sum_p I_p exp(+2 pi i (u l_p + v m_p + w n'_p)).
void DftPredict(size_t width, size_t height, double dl, double dm,
                double lShift, double mShift, size_t n, const double* uvw,
                const double* image, std::complex<double>* vis,
                size_t nThreads);

}  // namespace oracle

#endif
