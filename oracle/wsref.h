// ORACLE / TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// CPU restatement of the reference w-stacking gridder, WStackingGridder<T>
// (gridding/wstackinggridder.h:92, gridding/wstackinggridder.cpp). Every method
// cites the reference lines it follows. FFTW is replaced by the project-written
// FFT in cpufft.h and the REDFT01 (DCT-III) by its defining cosine sum, so the
// results agree with the reference to FP rounding, not bit for bit.
#ifndef ORACLE_WSREF_H
#define ORACLE_WSREF_H

#include "cpufft.h"

#include <complex>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace oracle {

// gridding/gridmode.h:6-44
enum class KernelMode : int {
  NearestNeighbour = 0,
  KaiserBessel = 1,
  KaiserBesselWithoutSinc = 2,
  Rectangular = 3,
  Gaussian = 4,
  GaussianWithoutSinc = 5,
  BlackmanNuttall = 6,
  BlackmanNuttallWithoutSinc = 7,
  BlackmanHarris = 8
};

// gridding/wstackinggridder.cpp:387-423 (GetKernel) and the kernel builders
// at :425-502, :504-515. Returns false for BlackmanHarris (the reference throws).
bool MakeKernel1D(KernelMode mode, double* kernel, size_t overSampling,
                  size_t size);

// I0 power series, gridding/wstackinggridder.cpp:504-515.
double Bessel0(double x, double precision);

template <typename T>
class WStackingRef {
 public:
  // gridding/wstackinggridder.cpp:16-41
  WStackingRef(size_t width, size_t height, double pixelSizeX,
               double pixelSizeY, size_t fftThreadCount, size_t kernelSize,
               size_t overSamplingFactor);

  void SetGridMode(KernelMode mode);  // gridding/wstackinggridder.h:470-475
  void SetIsComplex(bool c) { isComplex_ = c; }
  void SetDenormalPhaseCentre(double dl, double dm) {
    lShift_ = dl;
    mShift_ = dm;
  }

  void PrepareWLayers(size_t nWLayers, double maxMem, double minW,
                      double maxW);  // cpp:63-116

  // h:182-193; a |w| below minW - dw/2 gives a negative rounded value; the
  // reference casts it to size_t (undefined behaviour). Here it is returned
  // as SIZE_MAX, which every caller treats as "not in any pass".
  size_t WToLayer(double w) const;
  double LayerToW(size_t layer) const;  // h:202-212
  size_t NPasses() const { return nPasses_; }
  size_t NWLayers() const { return nW_; }
  bool IsInLayerRange(double w) const;              // h:228-232
  bool IsInLayerRange(double w1, double w2) const;  // h:247-260
  size_t LayerRangeStart(size_t i) const { return (nW_ * i) / nPasses_; }
  size_t CurrentPass() const { return pass_; }
  double MinW() const { return minW_; }
  double MaxW() const { return maxW_; }
  bool IsComplex() const { return isComplex_; }

  void StartInversionPass(size_t pass);  // cpp:127-137
  void AddDataSample(std::complex<float> sample, double u, double v,
                     double w);  // cpp:517-597
  void FinishInversionPass();     // cpp:323-336
  void FinalizeImage(double multiplicationFactor);  // cpp:691-715

  void InitializePrediction(const float* real,
                            const float* imag);  // h:341-357, cpp:829-849
  void StartPredictionPass(size_t pass);         // cpp:139-157
  void SampleDataSample(std::complex<float>& value, double u, double v,
                        double w) const;  // cpp:599-689

  const std::complex<T>* GetGriddedUVLayer(size_t i) const {
    return layers_[i].data();
  }
  const T* RealImage() const { return image_[0].data(); }
  const T* ImaginaryImage() const { return imageImag_[0].data(); }
  const std::vector<double>& Kernel1D() const { return k1d_; }
  const std::vector<T>& SqrtLMTable() const { return sqrtLM_; }
  size_t Width() const { return width_; }
  size_t Height() const { return height_; }
  size_t NFFTThreads() const { return nFFTThreads_; }

  // The kernel-correction factor D[k] (the REDFT01 of the half kernel,
  // cpp:784-789) evaluated for k = 0..n/2 of an axis of n pixels.
  std::vector<double> KernelCorrection(size_t n) const;

 private:
  void makeKernels();  // cpp:338-385
  void initializeSqrtLMLookupTable();             // cpp:851-872
  void initializeSqrtLMLookupTableForSampling();  // cpp:909-932
  template <bool IsComplexImpl>
  void projectOnImageAndCorrect(const std::complex<T>* source, double w,
                                size_t threadIndex);  // cpp:874-907
  template <bool IsComplexImpl>
  void copyImageToLayerAndInverseCorrect(std::complex<T>* dest,
                                         double w);  // cpp:934-969
  template <bool Inverse>
  void correctImageForKernel(T* image) const;  // cpp:717-827
  void finalizeImage(double mult, std::vector<std::vector<T>>& data);
  void initializePrediction(const float* in, std::vector<std::vector<T>>& data);

  size_t width_, height_;
  double pixelSizeX_, pixelSizeY_;
  size_t nW_ = 0, nPasses_ = 0, pass_ = 0;
  double minW_ = 0, maxW_ = 0, lShift_ = 0, mShift_ = 0;
  bool isComplex_ = false;
  KernelMode gridMode_ = KernelMode::KaiserBessel;
  size_t overSampling_, kernelSize_;
  std::vector<double> k1d_;
  std::vector<std::vector<T>> kernels_;
  std::vector<std::vector<std::complex<T>>> layers_;
  std::vector<std::vector<T>> image_, imageImag_;
  std::vector<T> sqrtLM_;
  size_t nFFTThreads_;
};

}  // namespace oracle

#endif
