// ORACLE / TEST INFRASTRUCTURE ONLY. Nothing in the shipped product path links
// or calls this file: only tests, the smoke check and the CPU-baseline leg of
// the benchmark use the oracle library it is part of.
//
// Project-written complex FFT used by the CPU restatement of the reference
// w-stacking gridder. The reference calls FFTW (un-vendored system library,
// version not pinned; call sites gridding/wstackinggridder.cpp:166-169,
// 207-210, 249-252, 289-292): fftw[f]_plan_dft_2d(H, W, ..., FFTW_BACKWARD or
// FFTW_FORWARD, FFTW_ESTIMATE), i.e. the unnormalised 2-D DFT
//   out[k1,k2] = sum_{n1,n2} in[n1,n2] * exp(sign * 2 pi i (k1 n1/H + k2 n2/W))
// with sign = +1 for BACKWARD and -1 for FORWARD. This file computes that
// transform with a mixed-radix Stockham algorithm (radices 4, 2, 3, 5, 7 and a
// generic odd-radix DFT for other prime factors). Twiddles are evaluated in
// long double and rounded to T.
#ifndef ORACLE_CPUFFT_H
#define ORACLE_CPUFFT_H

#include <cmath>
#include <complex>
#include <cstddef>
#include <vector>

namespace oracle {

template <typename T>
class Fft1D {
 public:
  // sign: +1 = backward (exp(+2 pi i)), -1 = forward (exp(-2 pi i)).
  Fft1D(size_t n, int sign) : n_(n), sign_(sign) {
    size_t m = n;
    while (m % 4 == 0) { radices_.push_back(4); m /= 4; }
    while (m % 2 == 0) { radices_.push_back(2); m /= 2; }
    for (size_t p = 3; m > 1; p += 2) {
      while (m % p == 0) { radices_.push_back(p); m /= p; }
    }
    tw_.resize(n);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (size_t k = 0; k != n; ++k) {
      long double a = sign * 2.0L * pi * (long double)k / (long double)n;
      tw_[k] = std::complex<T>((T)std::cos(a), (T)std::sin(a));
    }
    // per-stage copies of the twiddles a stage uses, in the order its loop
    // reads them: entry jm * (radix - 1) + (r - 1) is tw_[jm * r * n / (ns * radix)]
    size_t ns = 1;
    for (size_t radix : radices_) {
      const size_t twStep = n / (ns * radix);
      std::vector<std::complex<T>> st(ns * (radix - 1));
      for (size_t jm = 0; jm != ns; ++jm)
        for (size_t r = 1; r != radix; ++r)
          st[jm * (radix - 1) + (r - 1)] = tw_[(jm * r * twStep) % n];
      stageTw_.push_back(std::move(st));
      ns *= radix;
    }
  }

  size_t Size() const { return n_; }

  // In-place transform of data[0..n) (stride 1); scratch must hold n values.
  void Transform(std::complex<T>* data, std::complex<T>* scratch) const {
    std::complex<T>* in = data;
    std::complex<T>* out = scratch;
    size_t ns = 1;
    const size_t n = n_;
    std::complex<T> v[64];
    for (size_t s = 0; s != radices_.size(); ++s) {
      const size_t radix = radices_[s];
      const size_t nr = n / radix;
      const std::complex<T>* tw = stageTw_[s].data();
      // j = q * ns + jm runs over [0, nr) as in the Stockham recurrence
      for (size_t q = 0; q != nr / ns; ++q) {
        for (size_t jm = 0; jm != ns; ++jm) {
          const size_t j = q * ns + jm;
          v[0] = in[j];
          if (jm != 0) {
            const std::complex<T>* t = tw + jm * (radix - 1) - 1;
            for (size_t r = 1; r != radix; ++r) v[r] = in[j + r * nr] * t[r];
          } else {
            for (size_t r = 1; r != radix; ++r) v[r] = in[j + r * nr];
          }
          butterfly(v, radix);
          const size_t idxD = q * ns * radix + jm;
          for (size_t r = 0; r != radix; ++r) out[idxD + r * ns] = v[r];
        }
      }
      std::swap(in, out);
      ns *= radix;
    }
    if (in != data)
      for (size_t i = 0; i != n; ++i) data[i] = in[i];
  }

 private:
  void butterfly(std::complex<T>* v, size_t radix) const {
    if (radix == 2) {
      std::complex<T> a = v[0], b = v[1];
      v[0] = a + b;
      v[1] = a - b;
    } else if (radix == 4) {
      std::complex<T> a0 = v[0] + v[2], a1 = v[0] - v[2];
      std::complex<T> b0 = v[1] + v[3], b1 = v[1] - v[3];
      // multiply b1 by sign * i
      std::complex<T> jb1 = sign_ > 0 ? std::complex<T>(-b1.imag(), b1.real())
                                      : std::complex<T>(b1.imag(), -b1.real());
      v[0] = a0 + b0;
      v[2] = a0 - b0;
      v[1] = a1 + jb1;
      v[3] = a1 - jb1;
    } else {
      // generic small DFT: v'[k] = sum_r v[r] w^(r k), w = exp(sign 2 pi i / radix)
      std::complex<T> res[64];
      const size_t step = n_ / radix;
      for (size_t k = 0; k != radix; ++k) {
        std::complex<T> acc = v[0];
        for (size_t r = 1; r != radix; ++r)
          acc += v[r] * tw_[((r * k) % radix) * step];
        res[k] = acc;
      }
      for (size_t k = 0; k != radix; ++k) v[k] = res[k];
    }
  }

  size_t n_;
  int sign_;
  std::vector<size_t> radices_;
  std::vector<std::complex<T>> tw_;
  std::vector<std::vector<std::complex<T>>> stageTw_;
};

// Unnormalised 2-D transform of a height x width row-major array (x fastest),
// rows first, then columns (separable, identical to the joint 2-D DFT).
template <typename T>
class Fft2D {
 public:
  Fft2D(size_t width, size_t height, int sign)
      : w_(width), h_(height), rowFft_(width, sign), colFft_(height, sign) {}

  void Transform(std::complex<T>* data) const {
    std::vector<std::complex<T>> scratch(std::max(w_, h_) * kColBlock);
    for (size_t y = 0; y != h_; ++y)
      rowFft_.Transform(data + y * w_, scratch.data());
    // columns, kColBlock at a time through a contiguous buffer whose column
    // pitch is padded off a power of two (a 2^k pitch puts every element of a
    // gathered row into one cache set)
    const size_t pitch = h_ + 8;
    std::vector<std::complex<T>> col(pitch * kColBlock);
    for (size_t x0 = 0; x0 < w_; x0 += kColBlock) {
      const size_t nb = std::min(kColBlock, w_ - x0);
      for (size_t y = 0; y != h_; ++y)
        for (size_t b = 0; b != nb; ++b) col[b * pitch + y] = data[y * w_ + x0 + b];
      for (size_t b = 0; b != nb; ++b)
        colFft_.Transform(col.data() + b * pitch, scratch.data());
      for (size_t y = 0; y != h_; ++y)
        for (size_t b = 0; b != nb; ++b) data[y * w_ + x0 + b] = col[b * pitch + y];
    }
  }

 private:
  static constexpr size_t kColBlock = 16;
  size_t w_, h_;
  Fft1D<T> rowFft_, colFft_;
};

}  // namespace oracle

#endif
