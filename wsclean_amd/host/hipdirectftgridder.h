// HipDirectFTGridder: the FP64 direct Fourier transform behind the
// MSGridderBase plugin interface, in the role DirectMSGridder<double>
// (gridding/directmsgridder.{h,cpp}) plays on the CPU, registered for
// GridderType::DirectFT (scheduling/griddingtaskmanager.cpp:164-174). It
// images at the trimmed size with no kernel and no w-quantisation: every
// channel of every row is selected, weighted as readAndWeightVisibilities
// weights it, and summed exactly per pixel on the device
// (hws_direct_invert). Its Predict is the exact adjoint (hws_direct_predict),
// which the reference leaves unimplemented (directmsgridder.cpp:167-171).
#ifndef WSCLEAN_AMD_HOST_HIPDIRECTFTGRIDDER_H
#define WSCLEAN_AMD_HOST_HIPDIRECTFTGRIDDER_H

#include "msgridderbase.h"

#include "../../include/hip_wstacking.h"

#include <vector>

class HipDirectFTGridder final : public MSGridderBase {
 public:
  HipDirectFTGridder(const Settings& settings, const Resources& resources)
      : MSGridderBase(settings), _resources(resources), _hipDevice(settings.hipDevice) {}

  void Invert() final;
  void Predict(std::vector<aocommon::Image>&& images) final;
  std::vector<aocommon::Image> ResultImages() final { return {std::move(_image)}; }

 protected:
  size_t getSuggestedWGridSize() const final { return 1; }

 private:
  void check(hws_status s) const;
  const Resources _resources;
  const int _hipDevice;
  aocommon::Image _image;
};

#endif
