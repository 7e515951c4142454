// Helpers shared by the renderer and its adjoint: detector mapping and the
// 64-bit fixed-point LDS accumulation.
#pragma once

#include "dsx_internal.h"

namespace dsx {

typedef long long acc_t;
constexpr float kFixedOne = 1073741824.0f;  // 2^30

// Pixel position of a lab-frame spot (Simulation2D._get_transformed_coordinates,
// simulation2d.py:261-285, in pixel units).
template <class Args>
__device__ __forceinline__ bool to_pixel(const float4 &sp, const Args &a, float &X, float &Y) {
  const float xp = sp.x / a.cal, yp = sp.y / a.cal;
  if (a.mirrored) {
    X = xp * a.ca + yp * a.sa + a.cx;
    Y = xp * a.sa - yp * a.ca + a.cy;
  } else {
    X = xp * a.ca - yp * a.sa + a.cx;
    Y = xp * a.sa + yp * a.ca + a.cy;
  }
  return X >= 0.0f && X < static_cast<float>(a.W) && Y >= 0.0f && Y < static_cast<float>(a.H);
}

// Clip box of _subpixel_gaussian (detector_functions.py:314-359) for a spot of
// intensity I at pixel (X, Y): rows [lo, hi] and columns [xlo, xhi] with
// radius sqrt(log(clip / (I pref)) / ep), clipped to the frame.  False when the
// radius is NaN (I pref < clip) or the box is empty.
__device__ __forceinline__ bool subpixel_box(float I, float X, float Y, float pref, float ep,
                                             float clip, int H, int W, int &lo, int &hi, int &xlo,
                                             int &xhi) {
  const double amp = static_cast<double>(I) * static_cast<double>(pref);
  const double rad2 = log(static_cast<double>(clip) / amp) / static_cast<double>(ep);
  if (!(amp > 0.0) || !(rad2 >= 0.0)) return false;
  const double rad = sqrt(fmin(rad2, 1e12));
  lo = max(0, static_cast<int>(ceil(static_cast<double>(Y) - rad)));
  hi = min(H, static_cast<int>(floor(static_cast<double>(Y) + rad + 1.0))) - 1;
  xlo = max(0, static_cast<int>(ceil(static_cast<double>(X) - rad)));
  xhi = min(W, static_cast<int>(floor(static_cast<double>(X) + rad + 1.0))) - 1;
  return hi >= lo && xhi >= xlo;
}

__device__ __forceinline__ void acc_add(acc_t *p, float v, float scale) {
  const long long q = __float2int_rn(v * scale);
  atomicAdd(reinterpret_cast<unsigned long long *>(p), static_cast<unsigned long long>(q));
}

// int64 fixed point -> float (exact to float rounding)
__device__ __forceinline__ float acc_to_float(acc_t v) {
  const int hi = static_cast<int>(v >> 32);
  const unsigned lo = static_cast<unsigned>(v);
  return fmaf(static_cast<float>(hi), 4294967296.0f, static_cast<float>(lo));
}

__device__ __forceinline__ acc_t wave_max_acc(acc_t v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const acc_t o = __shfl_xor(v, off, kWave);
    v = o > v ? o : v;
  }
  return v;
}

}  // namespace dsx
