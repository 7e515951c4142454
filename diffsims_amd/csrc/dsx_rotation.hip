// Orientation parameterisation for the differentiable path: Bunge Euler
// angles (phi1, Phi, phi2, radians) -> the rows of M^T that the spot kernels
// read (g_lab = M^T g, M = orix Rotation.to_matrix(), the passive Bunge
// matrix), and the adjoint dL/d(angles) = sum_ij dL/dRt_ij dRt_ij/d(angle).
// One thread per orientation, FP64; replaces ~70 small tensor operations of a
// generic autograd graph with two launches.
#include "dsx_internal.h"

namespace dsx {

struct Bunge {
  double M[3][3];
  double d[3][3][3];  // d[k] = dM / d(angle k)
};

__device__ __forceinline__ void bunge(const double *e, Bunge &b, bool grad) {
  double s1, c1, s, c, s2, c2;
  sincos(e[0], &s1, &c1);
  sincos(e[1], &s, &c);
  sincos(e[2], &s2, &c2);
  b.M[0][0] = c1 * c2 - s1 * s2 * c;
  b.M[0][1] = s1 * c2 + c1 * s2 * c;
  b.M[0][2] = s2 * s;
  b.M[1][0] = -c1 * s2 - s1 * c2 * c;
  b.M[1][1] = -s1 * s2 + c1 * c2 * c;
  b.M[1][2] = c2 * s;
  b.M[2][0] = s1 * s;
  b.M[2][1] = -c1 * s;
  b.M[2][2] = c;
  if (!grad) return;
  // phi1
  b.d[0][0][0] = -s1 * c2 - c1 * s2 * c;
  b.d[0][0][1] = c1 * c2 - s1 * s2 * c;
  b.d[0][0][2] = 0.0;
  b.d[0][1][0] = s1 * s2 - c1 * c2 * c;
  b.d[0][1][1] = -c1 * s2 - s1 * c2 * c;
  b.d[0][1][2] = 0.0;
  b.d[0][2][0] = c1 * s;
  b.d[0][2][1] = s1 * s;
  b.d[0][2][2] = 0.0;
  // Phi
  b.d[1][0][0] = s1 * s2 * s;
  b.d[1][0][1] = -c1 * s2 * s;
  b.d[1][0][2] = s2 * c;
  b.d[1][1][0] = s1 * c2 * s;
  b.d[1][1][1] = -c1 * c2 * s;
  b.d[1][1][2] = c2 * c;
  b.d[1][2][0] = s1 * c;
  b.d[1][2][1] = -c1 * c;
  b.d[1][2][2] = -s;
  // phi2
  b.d[2][0][0] = -c1 * s2 - s1 * c2 * c;
  b.d[2][0][1] = -s1 * s2 + c1 * c2 * c;
  b.d[2][0][2] = c2 * s;
  b.d[2][1][0] = -c1 * c2 + s1 * s2 * c;
  b.d[2][1][1] = -s1 * c2 - c1 * s2 * c;
  b.d[2][1][2] = -s2 * s;
  b.d[2][2][0] = 0.0;
  b.d[2][2][1] = 0.0;
  b.d[2][2][2] = 0.0;
}

__global__ void euler_rt_kernel(const double *euler, double *rt, int n) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n) return;
  Bunge b;
  bunge(euler + 3 * (size_t)o, b, false);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) rt[9 * (size_t)o + 3 * i + j] = b.M[j][i];  // Rt = M^T
}

__global__ void euler_rt_backward_kernel(const double *euler, const double *drt, double *de, int n) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n) return;
  Bunge b;
  bunge(euler + 3 * (size_t)o, b, true);
  const double *g = drt + 9 * (size_t)o;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) acc += g[3 * i + j] * b.d[k][j][i];
    de[3 * (size_t)o + k] = acc;
  }
}

}  // namespace dsx

using namespace dsx;

extern "C" int dsx_euler_to_rt(const double *euler_d, int n, double *rt_d, void *stream) {
  clear_error();
  if (n < 0) {
    set_error("dsx_euler_to_rt: bad count %d", n);
    return DSX_EINVAL;
  }
  if (n == 0) return DSX_OK;
  if (!euler_d || !rt_d) {
    set_error("dsx_euler_to_rt: null device pointer");
    return DSX_EINVAL;
  }
  hipLaunchKernelGGL(euler_rt_kernel, dim3((n + 255) / 256), dim3(256), 0, as_stream(stream), euler_d,
                     rt_d, n);
  return check_launch("dsx_euler_to_rt");
}

extern "C" int dsx_euler_to_rt_backward(const double *euler_d, const double *drt_d, int n,
                                        double *deuler_d, void *stream) {
  clear_error();
  if (n < 0) {
    set_error("dsx_euler_to_rt_backward: bad count %d", n);
    return DSX_EINVAL;
  }
  if (n == 0) return DSX_OK;
  if (!euler_d || !drt_d || !deuler_d) {
    set_error("dsx_euler_to_rt_backward: null device pointer");
    return DSX_EINVAL;
  }
  hipLaunchKernelGGL(euler_rt_backward_kernel, dim3((n + 255) / 256), dim3(256), 0, as_stream(stream),
                     euler_d, drt_d, deuler_d, n);
  return check_launch("dsx_euler_to_rt_backward");
}
