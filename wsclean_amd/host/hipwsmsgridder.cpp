#include "hipwsmsgridder.h"

#include <algorithm>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>

namespace {
// rows read and weighted per staged batch
constexpr size_t kBatchRows = 4096;
}  // namespace

HipWSMSGridder::HipWSMSGridder(const Settings& settings, const Resources& resources)
    : MSGridderBase(settings),
      _resources(resources),
      _antialiasingKernelSize(settings.antialiasingKernelSize),
      _overSamplingFactor(settings.overSamplingFactor),
      _hipPrecision(int(settings.hipPrecision)),
      _hipDevice(settings.hipDevice),
      _hipTileSize(int(settings.hipTileSize)),
      _hipMaxLayerMemory(settings.hipMaxLayerMemory) {}

HipWSMSGridder::~HipWSMSGridder() noexcept { releaseContext(); }

void HipWSMSGridder::check(hws_status s) const {
  if (s != HWS_OK)
    throw std::runtime_error(std::string("MI355X w-stacking gridder: ") + hws_status_string(s) +
                             ": " + hws_last_error(_context));
}

void HipWSMSGridder::releaseContext() {
  hws_destroy(_context);
  _context = nullptr;
}

// The context at the inversion size (small inversion makes it smaller than
// the image), holding one image: a task grids one correlation.
void HipWSMSGridder::makeContext() {
  releaseContext();
  hws_config c{};
  c.width = ActualInversionWidth();
  c.height = ActualInversionHeight();
  c.pixel_size_x = ActualPixelSizeX();
  c.pixel_size_y = ActualPixelSizeY();
  c.kernel_size = _antialiasingKernelSize;
  c.oversampling = _overSamplingFactor;
  c.kernel_mode = int(GetGridMode());
  c.precision = _hipPrecision;
  c.device = _hipDevice;
  c.n_images = 1;
  c.complex_mask = IsComplex() ? 1u : 0u;
  c.l_shift = LShift();
  c.m_shift = MShift();
  // WSMSGridder rotates samples when the phase centre is shifted
  // (gridding/wsmsgridder.cpp:185-190, 328-332)
  c.rotate_visibilities = HasDenormalPhaseCentre() ? 1 : 0;
  c.tile_size = size_t(_hipTileSize);
  hws_gridder* g = nullptr;
  const hws_status s = hws_create(&c, &g);
  if (s != HWS_OK)
    throw std::runtime_error(std::string("MI355X w-stacking gridder: ") + hws_status_string(s) +
                             ": " + hws_last_error(nullptr));
  _context = g;
  check(hws_prepare_w_layers(_context, ActualWGridSize(), _hipMaxLayerMemory, _minW, _maxW,
                             &_nPasses));
}

size_t HipWSMSGridder::getSuggestedWGridSize() const {
  size_t wWidth, wHeight;
  if (_settings.widthForNWCalculation != 0 || _settings.heightForNWCalculation != 0) {
    wWidth = _settings.widthForNWCalculation;
    wHeight = _settings.heightForNWCalculation;
  } else {
    wWidth = TrimWidth();
    wHeight = TrimHeight();
  }
  const double maxL = wWidth * PixelSizeX() * 0.5 + std::fabs(LShift()),
               maxM = wHeight * PixelSizeY() * 0.5 + std::fabs(MShift()),
               lmSq = maxL * maxL + maxM * maxM;
  const double cMinW = IsComplex() ? -_maxW : _minW;
  const double radiansForAllLayers =
      (lmSq < 1.0) ? 2 * M_PI * (_maxW - cMinW) * (1.0 - std::sqrt(1.0 - lmSq))
                   : 2 * M_PI * (_maxW - cMinW);
  size_t suggested = size_t(std::ceil(radiansForAllLayers * _settings.nWLayersFactor));
  if (suggested == 0) suggested = 1;
  if (suggested < _resources.NCpus()) {
    const double memoryRequired = double(_resources.NCpus()) * double(sizeof(float)) *
                                  double(ActualInversionWidth() * ActualInversionHeight());
    if (4.0 * memoryRequired < double(_resources.Memory())) suggested = _resources.NCpus();
  }
  return suggested;
}

bool HipWSMSGridder::sameBand(const std::vector<MSData>& msDataVector) const {
  for (const MSData& d : msDataVector)
    if (!(d.SelectedBand() == msDataVector.front().SelectedBand())) return false;
  return true;
}

void HipWSMSGridder::Invert() {
  std::vector<MSData> msDataVector;
  initializeMSDataVector(msDataVector);
  makeContext();
  resetVisibilityCounters();
  _stagedPath = sameBand(msDataVector);
  if (_stagedPath)
    invertStaged(msDataVector);
  else
    invertHostProtocol(msDataVector);
}

// Rows are read, selected and weighted as gridMeasurementSet does
// (gridding/wsmsgridder.cpp:130-222): a row whose channels all fall outside
// the w-layers is skipped, a channel outside them is deselected so that it
// adds nothing to the weight sums. The device grids every pass at once, so a
// channel is selected when its w falls on any layer.
void HipWSMSGridder::invertStaged(std::vector<MSData>& msDataVector) {
  const aocommon::BandData band = msDataVector.front().SelectedBand();
  const size_t nchan = band.ChannelCount();
  size_t maxRows = 0;
  for (const MSData& d : msDataVector) maxRows += d.matchingRows;
  std::vector<double> wl(nchan);
  for (size_t ch = 0; ch != nchan; ++ch) wl[ch] = band.ChannelWavelength(ch);
  check(hws_stage_begin(_context, maxRows, nchan, wl.data(), 1));
  const int kind = IsComplex() ? 1 : 0;
  const size_t nW = ActualWGridSize();
  std::vector<double> uvwBatch;
  std::vector<std::complex<float>> dataBatch;
  uvwBatch.reserve(kBatchRows * 3);
  dataBatch.reserve(kBatchRows * nchan);
  std::vector<std::complex<float>> modelBuffer(nchan), rowData(nchan);
  std::vector<float> weightBuffer(nchan);
  std::unique_ptr<bool[]> isSelected(new bool[nchan]);
  const GainMode gainMode = GetGainMode(Polarization(), 1);
  auto flush = [&] {
    check(hws_stage_rows(_context, uvwBatch.size() / 3, uvwBatch.data(),
                         reinterpret_cast<const float*>(dataBatch.data())));
    uvwBatch.clear();
    dataBatch.clear();
  };
  for (MSData& msData : msDataVector) {
    StartMeasurementSet(msData, false);
    InversionRow row;
    row.data = rowData.data();
    std::unique_ptr<MSReader> reader = msData.msProvider->MakeReader();
    while (reader->CurrentRowAvailable()) {
      double u, v, w;
      reader->ReadMeta(u, v, w);
      bool any = false;
      for (size_t ch = 0; ch != nchan; ++ch) {
        isSelected[ch] = hws_w_to_layer(_context, w / band.ChannelWavelength(ch), kind) < nW;
        any = any || isSelected[ch];
      }
      if (any) {
        row.uvw[0] = u;
        row.uvw[1] = v;
        row.uvw[2] = w;
        row.rowId = reader->RowId();
        readAndWeightVisibilities<1>(*reader, msData.antennaNames, row, band,
                                     weightBuffer.data(), modelBuffer.data(), isSelected.get(),
                                     gainMode);
        uvwBatch.insert(uvwBatch.end(), row.uvw, row.uvw + 3);
        dataBatch.insert(dataBatch.end(), rowData.begin(), rowData.end());
        ++msData.totalRowsProcessed;
        if (uvwBatch.size() == kBatchRows * 3) flush();
      }
      reader->NextInputRow();
    }
  }
  flush();
  check(hws_stage_end(_context));
  _realImage = aocommon::Image(TrimWidth(), TrimHeight());
  _imaginaryImage = IsComplex() ? aocommon::Image(TrimWidth(), TrimHeight()) : aocommon::Image();
  check(hws_invert_staged(_context, 1.0 / totalWeight(), ImageWidth(), ImageHeight(),
                          TrimWidth(), TrimHeight(), _realImage.Data(),
                          IsComplex() ? _imaginaryImage.Data() : nullptr));
}

// Bands that differ between measurement sets: each set's rows are gridded
// pass by pass through the host-pointer protocol (hws_add_data), as
// WStackingGridder takes them.
void HipWSMSGridder::invertHostProtocol(std::vector<MSData>& msDataVector) {
  struct Rows {
    std::vector<double> uvw, wl;
    std::vector<std::complex<float>> data;
  };
  std::vector<Rows> rows(msDataVector.size());
  const int kind = IsComplex() ? 1 : 0;
  const size_t nW = ActualWGridSize();
  const GainMode gainMode = GetGainMode(Polarization(), 1);
  for (size_t i = 0; i != msDataVector.size(); ++i) {
    MSData& msData = msDataVector[i];
    const aocommon::BandData band = msData.SelectedBand();
    const size_t nchan = band.ChannelCount();
    Rows& r = rows[i];
    for (size_t ch = 0; ch != nchan; ++ch) r.wl.push_back(band.ChannelWavelength(ch));
    std::vector<std::complex<float>> rowData(nchan), modelBuffer(nchan);
    std::vector<float> weightBuffer(nchan);
    std::unique_ptr<bool[]> isSelected(new bool[nchan]);
    StartMeasurementSet(msData, false);
    InversionRow row;
    row.data = rowData.data();
    std::unique_ptr<MSReader> reader = msData.msProvider->MakeReader();
    while (reader->CurrentRowAvailable()) {
      reader->ReadMeta(row.uvw[0], row.uvw[1], row.uvw[2]);
      row.rowId = reader->RowId();
      for (size_t ch = 0; ch != nchan; ++ch)
        isSelected[ch] = hws_w_to_layer(_context, row.uvw[2] / r.wl[ch], kind) < nW;
      readAndWeightVisibilities<1>(*reader, msData.antennaNames, row, band, weightBuffer.data(),
                                   modelBuffer.data(), isSelected.get(), gainMode);
      r.uvw.insert(r.uvw.end(), row.uvw, row.uvw + 3);
      r.data.insert(r.data.end(), rowData.begin(), rowData.end());
      reader->NextInputRow();
    }
  }
  for (size_t pass = 0; pass != _nPasses; ++pass) {
    check(hws_start_inversion_pass(_context, pass));
    for (const Rows& r : rows)
      check(hws_add_data(_context, r.uvw.size() / 3, r.wl.size(), r.uvw.data(), r.wl.data(),
                         reinterpret_cast<const float*>(r.data.data())));
    check(hws_finish_inversion_pass(_context));
  }
  _realImage = aocommon::Image(TrimWidth(), TrimHeight());
  _imaginaryImage = IsComplex() ? aocommon::Image(TrimWidth(), TrimHeight()) : aocommon::Image();
  check(hws_finalize_image_resampled(_context, 1.0 / totalWeight(), ImageWidth(), ImageHeight(),
                                     TrimWidth(), TrimHeight(), _realImage.Data(),
                                     IsComplex() ? _imaginaryImage.Data() : nullptr));
}

void HipWSMSGridder::Predict(std::vector<aocommon::Image>&& images) {
  // gridding/wsmsgridder.cpp:489-492
  if (images.size() != 2 && IsComplex())
    throw std::runtime_error("Missing imaginary in complex prediction");
  if (images.size() != 1 && !IsComplex())
    throw std::runtime_error("Imaginary specified in non-complex prediction");
  std::vector<MSData> msDataVector;
  initializeMSDataVector(msDataVector);
  for (const aocommon::Image& image : images)
    if (image.Width() != TrimWidth() || image.Height() != TrimHeight())
      throw std::runtime_error("model image size differs from the trimmed image size");
  makeContext();
  _stagedPath = sameBand(msDataVector);
  if (_stagedPath)
    predictStaged(msDataVector, images);
  else
    predictHostProtocol(msDataVector, images);
}

// predictMeasurementSet (gridding/wsmsgridder.cpp:255-365): every row's uvw
// is read, every row written back in row order; samples no pass predicts are
// NaN and left alone by the write.
void HipWSMSGridder::predictStaged(std::vector<MSData>& msDataVector,
                                   std::vector<aocommon::Image>& images) {
  const aocommon::BandData band = msDataVector.front().SelectedBand();
  const size_t nchan = band.ChannelCount();
  size_t maxRows = 0;
  for (const MSData& d : msDataVector) maxRows += d.matchingRows;
  std::vector<double> wl(nchan);
  for (size_t ch = 0; ch != nchan; ++ch) wl[ch] = band.ChannelWavelength(ch);
  check(hws_stage_begin(_context, maxRows, nchan, wl.data(), 0));
  std::vector<double> uvwBatch;
  uvwBatch.reserve(kBatchRows * 3);
  std::vector<size_t> rowsPerSet;
  for (MSData& msData : msDataVector) {
    size_t n = 0;
    std::unique_ptr<MSReader> reader = msData.msProvider->MakeReader();
    while (reader->CurrentRowAvailable()) {
      double uvw[3];
      reader->ReadMeta(uvw[0], uvw[1], uvw[2]);
      uvwBatch.insert(uvwBatch.end(), uvw, uvw + 3);
      if (uvwBatch.size() == kBatchRows * 3) {
        check(hws_stage_rows(_context, kBatchRows, uvwBatch.data(), nullptr));
        uvwBatch.clear();
      }
      ++n;
      reader->NextInputRow();
    }
    rowsPerSet.push_back(n);
    msData.totalRowsProcessed += n;
  }
  check(hws_stage_rows(_context, uvwBatch.size() / 3, uvwBatch.data(), nullptr));
  check(hws_stage_end(_context));
  size_t totalRows = 0;
  for (size_t n : rowsPerSet) totalRows += n;
  std::vector<std::complex<float>> out(totalRows * nchan);
  check(hws_predict_staged(_context, images[0].Data(),
                           images.size() == 2 ? images[1].Data() : nullptr, TrimWidth(),
                           TrimHeight(), ImageWidth(), ImageHeight(),
                           reinterpret_cast<float*>(out.data())));
  const GainMode gainMode = GetGainMode(Polarization(), 1);
  size_t offset = 0;
  for (size_t i = 0; i != msDataVector.size(); ++i) {
    MSData& msData = msDataVector[i];
    msData.msProvider->ResetWritePosition();
    StartMeasurementSet(msData, true);
    for (size_t r = 0; r != rowsPerSet[i]; ++r)
      writeVisibilities<1>(*msData.msProvider, msData.antennaNames, band,
                           &out[(offset + r) * nchan], gainMode);
    offset += rowsPerSet[i];
  }
}

void HipWSMSGridder::predictHostProtocol(std::vector<MSData>& msDataVector,
                                         std::vector<aocommon::Image>& images) {
  check(hws_initialize_prediction_resampled(_context, images[0].Data(),
                                            images.size() == 2 ? images[1].Data() : nullptr,
                                            TrimWidth(), TrimHeight(), ImageWidth(),
                                            ImageHeight()));
  const GainMode gainMode = GetGainMode(Polarization(), 1);
  for (size_t pass = 0; pass != _nPasses; ++pass) {
    check(hws_start_prediction_pass(_context, pass));
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
      check(hws_sample_data(_context, nrow, nchan, uvw.data(), wl.data(),
                            reinterpret_cast<float*>(out.data())));
      msData.msProvider->ResetWritePosition();
      StartMeasurementSet(msData, true);
      // rows in order; samples of other passes are NaN and skipped
      for (size_t r = 0; r != nrow; ++r)
        writeVisibilities<1>(*msData.msProvider, msData.antennaNames, band, &out[r * nchan],
                             gainMode);
    }
  }
}
