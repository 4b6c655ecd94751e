#include "hipdirectftgridder.h"

#include <complex>
#include <memory>
#include <stdexcept>
#include <string>

void HipDirectFTGridder::check(hws_status s) const {
  if (s != HWS_OK)
    throw std::runtime_error(std::string("MI355X direct FT gridder: ") + hws_status_string(s) +
                             ": " + hws_last_error(nullptr));
}

// directmsgridder.cpp:16-74 with invertMeasurementSet (:123-165): all
// channels selected, the image normalised by 1 / totalWeight.
void HipDirectFTGridder::Invert() {
  std::vector<MSData> msDataVector;
  initializeMSDataVector(msDataVector);
  resetVisibilityCounters();
  const size_t width = TrimWidth(), height = TrimHeight();
  std::vector<double> sum(width * height, 0.0), part(width * height);
  const GainMode gainMode = GetGainMode(Polarization(), 1);
  for (MSData& msData : msDataVector) {
    StartMeasurementSet(msData, false);
    const aocommon::BandData band = msData.SelectedBand();
    const size_t nchan = band.ChannelCount();
    std::vector<double> wl(nchan), uvw;
    for (size_t ch = 0; ch != nchan; ++ch) wl[ch] = band.ChannelWavelength(ch);
    std::vector<std::complex<float>> rowData(nchan), modelBuffer(nchan), data;
    std::vector<float> weightBuffer(nchan);
    std::unique_ptr<bool[]> isSelected(new bool[nchan]);
    for (size_t ch = 0; ch != nchan; ++ch) isSelected[ch] = true;
    InversionRow row;
    row.data = rowData.data();
    std::unique_ptr<MSReader> reader = msData.msProvider->MakeReader();
    while (reader->CurrentRowAvailable()) {
      reader->ReadMeta(row.uvw[0], row.uvw[1], row.uvw[2]);
      row.rowId = reader->RowId();
      readAndWeightVisibilities<1>(*reader, msData.antennaNames, row, band, weightBuffer.data(),
                                   modelBuffer.data(), isSelected.get(), gainMode);
      uvw.insert(uvw.end(), row.uvw, row.uvw + 3);
      data.insert(data.end(), rowData.begin(), rowData.end());
      reader->NextInputRow();
    }
    check(hws_direct_invert(_hipDevice, width, height, PixelSizeX(), PixelSizeY(), LShift(),
                            MShift(), uvw.size() / 3, nchan, uvw.data(), wl.data(),
                            reinterpret_cast<const float*>(data.data()), 1.0, part.data()));
    for (size_t i = 0; i != sum.size(); ++i) sum[i] += part[i];
  }
  _image = aocommon::Image(width, height);
  const double wFactor = 1.0 / totalWeight();
  for (size_t i = 0; i != sum.size(); ++i) _image[i] = float(sum[i] * wFactor);
}

// The adjoint of Invert without its normalisation: V = sum_p I_p
// exp(+2 pi i (u l_p + v m_p + w n'_p)), written to every row.
void HipDirectFTGridder::Predict(std::vector<aocommon::Image>&& images) {
  if (images.size() != 1)
    throw std::runtime_error("the direct FT gridder predicts from one real image");
  std::vector<MSData> msDataVector;
  initializeMSDataVector(msDataVector);
  const size_t width = TrimWidth(), height = TrimHeight();
  if (images[0].Width() != width || images[0].Height() != height)
    throw std::runtime_error("model image size differs from the trimmed image size");
  const GainMode gainMode = GetGainMode(Polarization(), 1);
  for (MSData& msData : msDataVector) {
    const aocommon::BandData band = msData.SelectedBand();
    const size_t nchan = band.ChannelCount();
    std::vector<double> wl(nchan), uvw;
    for (size_t ch = 0; ch != nchan; ++ch) wl[ch] = band.ChannelWavelength(ch);
    std::unique_ptr<MSReader> reader = msData.msProvider->MakeReader();
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
    check(hws_direct_predict(_hipDevice, width, height, PixelSizeX(), PixelSizeY(), LShift(),
                             MShift(), images[0].Data(), nrow, nchan, uvw.data(), wl.data(),
                             reinterpret_cast<float*>(out.data())));
    msData.msProvider->ResetWritePosition();
    StartMeasurementSet(msData, true);
    for (size_t r = 0; r != nrow; ++r)
      writeVisibilities<1>(*msData.msProvider, msData.antennaNames, band, &out[r * nchan],
                           gainMode);
  }
}
