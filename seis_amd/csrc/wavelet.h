// Ricker source wavelet of the reference (util/wavelet.cpp:7-18):
// w[i] = (1 - 2 a t^2) exp(-a t^2), a = pi^2 fr^2 formed in double with
// pi = 3.1415926 and stored to float, t = (i - num_delay) dt in float.
#pragma once
#include <cmath>

namespace seis {

inline void ricker_host(float dt, float fr, int num_delay, int n, float* w) {
  const float a = 3.1415926 * 3.1415926 * fr * fr;
  for (int i = 0; i < n; i++) {
    const float t = (i - num_delay) * 1.0f * dt;
    const float t2 = t * t;
    w[i] = (1.0f - 2.0f * a * t2) * std::exp(-a * t2);
  }
}

}  // namespace seis
