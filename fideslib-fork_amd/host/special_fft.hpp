// CKKS special FFT (Cheon-Kim-Kim-Song) of the facade's encoder and of the bootstrapping
// transforms: slots at zeta^(5^j), explicit real arithmetic so the rounding sequence is fixed.
// Twiddles come from one sincos call per angle (glibc's sincos and separate sin / cos calls differ
// in the last bit for some angles), as in the device encoder's tables.
#pragma once
#include <cmath>
#include <cstdint>
#include <utility>
#include <vector>

namespace fideslib {
namespace detail {

constexpr double kPi = 3.14159265358979323846;

struct Cx {
  double re, im;
};

inline void bitrev(std::vector<Cx>& v) {
  const int n = (int)v.size();
  for (int i = 1, j = 0; i < n; ++i) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(v[i], v[j]);
  }
}

inline void special_ifft(std::vector<Cx>& v, uint64_t M) {
  const int n = (int)v.size();
  for (int len = n; len >= 1; len >>= 1) {
    const int lenh = len >> 1;
    const uint64_t lenq = (uint64_t)len << 2;
    for (int i = 0; i < n; i += len) {
      uint64_t g = 1;
      for (int j = 0; j < lenh; ++j) {
        const uint64_t idx = (lenq - (g % lenq)) * (M / lenq);
        const double ang = 2.0 * kPi * (double)idx / (double)M;
        double wr, wi;
        ::sincos(ang, &wi, &wr);
        const Cx a = v[i + j], b = v[i + j + lenh];
        const double ur = a.re + b.re, ui = a.im + b.im;
        const double dr = a.re - b.re, di = a.im - b.im;
        v[i + j] = Cx{ur, ui};
        v[i + j + lenh] = Cx{dr * wr - di * wi, dr * wi + di * wr};
        g = (g * 5) % M;
      }
    }
  }
  bitrev(v);
  for (auto& x : v) {
    x.re = x.re / n;
    x.im = x.im / n;
  }
}

inline void special_fft(std::vector<Cx>& v, uint64_t M) {
  const int n = (int)v.size();
  bitrev(v);
  for (int len = 2; len <= n; len <<= 1) {
    const int lenh = len >> 1;
    const uint64_t lenq = (uint64_t)len << 2;
    for (int i = 0; i < n; i += len) {
      uint64_t g = 1;
      for (int j = 0; j < lenh; ++j) {
        const uint64_t idx = (g % lenq) * (M / lenq);
        const double ang = 2.0 * kPi * (double)idx / (double)M;
        double wr, wi;
        ::sincos(ang, &wi, &wr);
        const Cx a = v[i + j], b = v[i + j + lenh];
        const double tr = b.re * wr - b.im * wi, ti = b.re * wi + b.im * wr;
        v[i + j] = Cx{a.re + tr, a.im + ti};
        v[i + j + lenh] = Cx{a.re - tr, a.im - ti};
        g = (g * 5) % M;
      }
    }
  }
}

}  // namespace detail
}  // namespace fideslib
