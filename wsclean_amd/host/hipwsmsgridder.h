// HipWSMSGridder: the MI355X w-stacking gridder behind the MSGridderBase
// plugin interface, in the role WSMSGridder (gridding/wsmsgridder.h:20-88)
// plays for the CPU WStackingGridder<float>. It reads and weights the
// visibilities on the host exactly as MSGridderBase does and hands them to
// the HIP core through the C ABI of include/hip_wstacking.h; every C-ABI
// failure becomes a std::runtime_error carrying hws_last_error().
#ifndef WSCLEAN_AMD_HOST_HIPWSMSGRIDDER_H
#define WSCLEAN_AMD_HOST_HIPWSMSGRIDDER_H

#include "msgridderbase.h"

#include "../../include/hip_wstacking.h"

namespace wsclean_amd {

class HipWSMSGridder final : public MSGridderBase {
 public:
  HipWSMSGridder(const Settings& settings, const Resources& resources);
  ~HipWSMSGridder() noexcept override = default;

  void Invert() override;                               // wsmsgridder.cpp:367-486
  void Predict(std::vector<Image>&& images) override;  // wsmsgridder.cpp:488-586
  std::vector<Image> ResultImages() override {
    if (IsComplex()) return {std::move(realImage_), std::move(imaginaryImage_)};
    return {std::move(realImage_)};
  }
  void FreeImagingData() override {}

  size_t AntialiasingKernelSize() const { return kernelSize_; }
  size_t OverSamplingFactor() const { return overSampling_; }
  size_t NPasses() const { return nPasses_; }

 protected:
  size_t getSuggestedWGridSize() const override;  // wsmsgridder.cpp:77-128

 private:
  struct Context {
    hws_gridder* g = nullptr;
    ~Context() { hws_destroy(g); }
  };
  void makeContext(Context& ctx) const;
  static void check(hws_status s, const hws_gridder* g);

  const Resources resources_;
  size_t kernelSize_, overSampling_;
  size_t nPasses_ = 0;
  Image realImage_, imaginaryImage_;
};

}  // namespace wsclean_amd

#endif
