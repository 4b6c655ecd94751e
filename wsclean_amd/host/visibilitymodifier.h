// GainMode and GetGainMode (gridding/visibilitymodifier.h:25-54): which
// entries of a direction-dependent gain matrix correct the visibilities a
// gridder reads and writes. Facet beams and gain solutions are not part of
// this build, so the mode is carried through the plugin interface unused.
#ifndef WSCLEAN_AMD_HOST_VISIBILITYMODIFIER_H
#define WSCLEAN_AMD_HOST_VISIBILITYMODIFIER_H

#include "banddata.h"

#include <cstddef>
#include <stdexcept>

enum class GainMode { kXX, kYY, kDiagonal, kFull };

inline GainMode GetGainMode(aocommon::PolarizationEnum polarization,
                            size_t n_visibility_polarizations) {
  if (n_visibility_polarizations != 1) throw std::runtime_error("Not yet implemented");
  switch (polarization) {
    case aocommon::Polarization::XX:
      return GainMode::kXX;
    case aocommon::Polarization::YY:
      return GainMode::kYY;
    default:
      return GainMode::kDiagonal;
  }
}

#endif
