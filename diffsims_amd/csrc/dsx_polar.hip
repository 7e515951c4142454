// Polar templates on the device: Simulation2D.polar_flatten_simulations
// (diffsims/simulations/simulation2d.py:313-355) with to_flat_polar
// (diffsims/crystallography/_diffracting_vector.py:186-194) and get_closest
// (simulation2d.py:767-781), computed straight from the spot lists the
// per-orientation kernel wrote, so template matching needs no host copy of
// 10^5-10^6 spot lists.
//
// One wave per orientation.  r = |(x, y)| and theta = atan2(y, x) in FP64 from
// the stored lab-frame coordinates.  With radial and azimuthal axes each value
// becomes the index of the nearest axis entry (the left neighbour only when it
// is strictly closer, as searchsorted(side="left") + the distance test do),
// entries on or past the last axis index are dropped and the rest compacted in
// order; rows are zero-padded to `width` like the reference's np.zeros arrays
// (entries past `width` are not written).
#include "dsx_internal.h"

namespace dsx {

struct PolarArgs {
  const float4 *spots;
  const int *count;
  int cap, n, width;
  const double *radial;
  int n_radial;
  const double *azimuthal;
  int n_azimuthal;
  float *r, *theta;
  int *r_idx, *theta_idx;
  float *I;
  int *n_out;
};

// index of the nearest entry of the ascending array a[0..n)
__device__ __forceinline__ int nearest_index(const double *a, int n, double v) {
  int lo = 0, hi = n;  // first index with a[i] >= v
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  const int left = lo > 0 ? lo - 1 : 0;
  const int right = lo < n ? lo : n - 1;
  if (lo == n || fabs(v - a[left]) < fabs(v - a[right])) return lo - 1;
  return lo;
}

constexpr int kPolarThreads = 256;

__global__ __launch_bounds__(kPolarThreads) void polar_kernel(PolarArgs a) {
  const int lane = lane_id();
  const int waves = gridDim.x * (kPolarThreads / kWave);
  const bool axes = a.radial != nullptr && a.azimuthal != nullptr;
  for (int o = blockIdx.x * (kPolarThreads / kWave) + threadIdx.x / kWave; o < a.n; o += waves) {
    const int n = min(a.count[o], a.cap);
    const float4 *sp = a.spots + (size_t)o * a.cap;
    const size_t row = (size_t)o * a.width;
    int base = 0;
    for (int j0 = 0; j0 < n; j0 += kWave) {
      const int j = j0 + lane;
      const bool valid = j < n;
      const float4 s = valid ? sp[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      const double x = s.x, y = s.y;
      const double r = sqrt(x * x + y * y);
      const double t = atan2(y, x);
      int ri = 0, ti = 0;
      bool keep = valid;
      if (axes && valid) {
        ri = nearest_index(a.radial, a.n_radial, r);
        ti = nearest_index(a.azimuthal, a.n_azimuthal, t);
        keep = ri < a.n_radial - 1 && ti < a.n_azimuthal - 1;
      }
      const uint64_t m = __ballot(keep);
      const int pos = base + mask_rank(m);
      if (keep && pos < a.width) {  // a row narrower than the list keeps its head
        const size_t q = row + pos;
        if (axes) {
          a.r_idx[q] = ri;
          a.theta_idx[q] = ti;
        } else {
          a.r[q] = static_cast<float>(r);
          a.theta[q] = static_cast<float>(t);
        }
        a.I[q] = s.w;
      }
      base += __popcll(m);
    }
    for (int q = base + lane; q < a.width; q += kWave) {
      if (axes) {
        a.r_idx[row + q] = 0;
        a.theta_idx[row + q] = 0;
      } else {
        a.r[row + q] = 0.0f;
        a.theta[row + q] = 0.0f;
      }
      a.I[row + q] = 0.0f;
    }
    if (lane == 0 && a.n_out) a.n_out[o] = base;
  }
}

}  // namespace dsx

using namespace dsx;

extern "C" int dsx_polar_templates(const float *spot_d, const int *count_d, int cap, int n,
                                   int width, const double *radial_d, int n_radial,
                                   const double *azimuthal_d, int n_azimuthal, float *r_d,
                                   float *theta_d, int *r_idx_d, int *theta_idx_d, float *I_d,
                                   int *n_out_d, void *stream) {
  clear_error();
  if (n < 0 || cap < 1 || width < 0) {
    set_error("dsx_polar_templates: bad arguments (n=%d cap=%d width=%d)", n, cap, width);
    return DSX_EINVAL;
  }
  if (n == 0) return DSX_OK;
  const bool axes = radial_d != nullptr && azimuthal_d != nullptr;
  if (axes && (n_radial < 1 || n_azimuthal < 1)) {
    set_error("dsx_polar_templates: empty axes");
    return DSX_EINVAL;
  }
  if (!spot_d || !count_d || !I_d || (axes && (!r_idx_d || !theta_idx_d)) ||
      (!axes && (!r_d || !theta_d))) {
    set_error("dsx_polar_templates: null device pointer");
    return DSX_EINVAL;
  }
  PolarArgs a;
  a.spots = reinterpret_cast<const float4 *>(spot_d);
  a.count = count_d;
  a.cap = cap;
  a.n = n;
  a.width = width;
  a.radial = axes ? radial_d : nullptr;
  a.n_radial = n_radial;
  a.azimuthal = axes ? azimuthal_d : nullptr;
  a.n_azimuthal = n_azimuthal;
  a.r = r_d;
  a.theta = theta_d;
  a.r_idx = r_idx_d;
  a.theta_idx = theta_idx_d;
  a.I = I_d;
  a.n_out = n_out_d;
  const int waves_per_block = kPolarThreads / kWave;
  const int blocks = std::min((n + waves_per_block - 1) / waves_per_block, cu_count() * 8);
  hipLaunchKernelGGL(polar_kernel, dim3(blocks), dim3(kPolarThreads), 0, as_stream(stream), a);
  return check_launch("dsx_polar_templates");
}
