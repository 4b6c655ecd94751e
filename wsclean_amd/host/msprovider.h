// Visibility access interfaces of the reference (msproviders/msprovider.h:44-151,
// msproviders/msreaders/msreader.h:16-73), reduced to what the w-stacking
// path calls, plus an in-memory provider holding one correlation of one
// observation in place of a casacore MeasurementSet.
#ifndef WSCLEAN_AMD_HOST_MSPROVIDER_H
#define WSCLEAN_AMD_HOST_MSPROVIDER_H

#include <cmath>
#include <complex>
#include <cstddef>
#include <memory>
#include <vector>

namespace wsclean_amd {

class MSReader;

class MSProvider {
 public:
  struct MetaData {
    double uInM, vInM, wInM;
    size_t fieldId, antenna1, antenna2;
    double time;
  };
  virtual ~MSProvider() = default;
  virtual std::unique_ptr<MSReader> MakeReader() = 0;
  virtual void NextOutputRow() = 0;
  virtual void ResetWritePosition() = 0;
  // Writes one row of model data; non-finite samples are skipped
  // (msproviders/msprovider.cpp:456-657). add: add instead of overwrite.
  virtual void WriteModel(const std::complex<float>* buffer, bool add) = 0;
  virtual double StartTime() = 0;
  virtual size_t NChannels() = 0;
  // Wavelength (m) of channel ch (BandData::ChannelWavelength).
  virtual double ChannelWavelength(size_t ch) = 0;
};

class MSReader {
 public:
  explicit MSReader(MSProvider* provider) : provider_(provider) {}
  virtual ~MSReader() = default;
  virtual size_t RowId() const = 0;
  virtual bool CurrentRowAvailable() = 0;
  virtual void NextInputRow() = 0;
  virtual void ReadMeta(double& u, double& v, double& w) = 0;
  virtual void ReadMeta(MSProvider::MetaData& metaData) = 0;
  virtual void ReadData(std::complex<float>* buffer) = 0;
  virtual void ReadModel(std::complex<float>* buffer) = 0;
  virtual void ReadWeights(float* buffer) = 0;

 protected:
  MSProvider* provider_;
};

// One correlation of an observation held in memory: rows of uvw (m), data,
// weights and model data per channel.
class MemoryMSProvider final : public MSProvider {
 public:
  MemoryMSProvider(std::vector<double> uvwM, std::vector<double> wavelengths,
                   std::vector<std::complex<float>> data, std::vector<float> weights,
                   std::vector<double> times = {})
      : uvw_(std::move(uvwM)),
        wavelengths_(std::move(wavelengths)),
        data_(std::move(data)),
        weights_(std::move(weights)),
        times_(std::move(times)) {
    nrow_ = uvw_.size() / 3;
    model_.assign(nrow_ * wavelengths_.size(), std::complex<float>(0.0f, 0.0f));
  }

  std::unique_ptr<MSReader> MakeReader() override;
  void NextOutputRow() override { ++writeRow_; }
  void ResetWritePosition() override { writeRow_ = 0; }
  void WriteModel(const std::complex<float>* buffer, bool add) override {
    const size_t nchan = wavelengths_.size();
    std::complex<float>* dst = &model_[writeRow_ * nchan];
    for (size_t ch = 0; ch != nchan; ++ch) {
      if (!std::isfinite(buffer[ch].real()) || !std::isfinite(buffer[ch].imag())) continue;
      dst[ch] = add ? dst[ch] + buffer[ch] : buffer[ch];
    }
  }
  double StartTime() override { return times_.empty() ? 0.0 : times_.front(); }
  size_t NChannels() override { return wavelengths_.size(); }
  double ChannelWavelength(size_t ch) override { return wavelengths_[ch]; }

  size_t NRows() const { return nrow_; }
  const std::vector<double>& Uvw() const { return uvw_; }
  const std::vector<std::complex<float>>& Data() const { return data_; }
  const std::vector<float>& Weights() const { return weights_; }
  const std::vector<double>& Wavelengths() const { return wavelengths_; }
  double Time(size_t row) const { return times_.empty() ? 0.0 : times_[row]; }
  std::vector<std::complex<float>>& Model() { return model_; }

 private:
  friend class MemoryMSReader;
  std::vector<double> uvw_, wavelengths_;
  std::vector<std::complex<float>> data_, model_;
  std::vector<float> weights_;
  std::vector<double> times_;
  size_t nrow_ = 0, writeRow_ = 0;
};

class MemoryMSReader final : public MSReader {
 public:
  explicit MemoryMSReader(MemoryMSProvider* p) : MSReader(p), p_(p) {}
  size_t RowId() const override { return row_; }
  bool CurrentRowAvailable() override { return row_ < p_->nrow_; }
  void NextInputRow() override { ++row_; }
  void ReadMeta(double& u, double& v, double& w) override {
    u = p_->uvw_[row_ * 3];
    v = p_->uvw_[row_ * 3 + 1];
    w = p_->uvw_[row_ * 3 + 2];
  }
  void ReadMeta(MSProvider::MetaData& m) override {
    ReadMeta(m.uInM, m.vInM, m.wInM);
    m.fieldId = m.antenna1 = m.antenna2 = 0;
    m.time = p_->Time(row_);
  }
  void ReadData(std::complex<float>* buffer) override {
    const size_t n = p_->wavelengths_.size();
    for (size_t ch = 0; ch != n; ++ch) buffer[ch] = p_->data_[row_ * n + ch];
  }
  void ReadModel(std::complex<float>* buffer) override {
    const size_t n = p_->wavelengths_.size();
    for (size_t ch = 0; ch != n; ++ch) buffer[ch] = p_->model_[row_ * n + ch];
  }
  void ReadWeights(float* buffer) override {
    const size_t n = p_->wavelengths_.size();
    for (size_t ch = 0; ch != n; ++ch)
      buffer[ch] = p_->weights_.empty() ? 1.0f : p_->weights_[row_ * n + ch];
  }

 private:
  MemoryMSProvider* p_;
  size_t row_ = 0;
};

inline std::unique_ptr<MSReader> MemoryMSProvider::MakeReader() {
  return std::unique_ptr<MSReader>(new MemoryMSReader(this));
}

}  // namespace wsclean_amd

#endif
