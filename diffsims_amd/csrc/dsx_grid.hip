// Beam-direction grids for template libraries: get_beam_directions_grid
// (diffsims/generators/rotation_list_generators.py:176-267) for the
// deterministic meshes of diffsims/generators/sphere_mesh_generators.py --
// the UV sphere (:42-93) and the three cube meshes (:96-197) -- cropped to a
// crystal system's stereographic triangle and converted to Bunge angles with
// phi1 = 0 (beam_directions_grid_to_euler, :486-526), written on the device in
// the reference's row order.
//
// One thread per mesh vertex (candidate): the vertex in FP64 with the
// reference's operation order (contraction off), the triangle test, the
// angles.  Rows stay in candidate order through a count / scan / write
// compaction (block counts, one scan block, block-local ballot ranks).
#include "dsx_internal.h"

namespace dsx {

constexpr int kGridThreads = 256;

struct GridArgs {
  int mesh;
  long long n_cand;
  // UV sphere: psi outer (n_psi steps of dpsi), theta inner (n_theta steps of
  // dtheta, the last exactly pi); pi_ref = deg2rad(180) as numpy forms it
  int n_psi, n_theta;
  double dpsi, dtheta, pi_ref;
  // cube meshes: m = 2 steps grid lines per face, coordinate k = j - steps
  int m;
  double steps, theta, sqrt2;
  // crop: n_planes of {normal (3 ints), sign}; keep s (n . v) >= -1e-13
  int n_planes;
  double plane[3][4];
  double *euler;
  long long capacity;
};

// Per-face grid coordinate j in [0, m) of the cube meshes.
__device__ __forceinline__ double cube_coord(const GridArgs &a, int j) {
  const double k = static_cast<double>(j) - a.steps;
  if (a.mesh == DSX_MESH_NORMALIZED_CUBE) return k / a.steps;
  if (a.mesh == DSX_MESH_SPHERIFIED_EDGE) return tan(k * a.theta);
  return tan(k * a.theta) / a.sqrt2;  // spherified corner
}

// Vertex `c` of the mesh; false when the reference drops it (the UV sphere
// keeps one point per pole).
__device__ __forceinline__ bool mesh_vertex(const GridArgs &a, long long c, double &x, double &y,
                                            double &z) {
#pragma clang fp contract(off)
  if (a.mesh == DSX_MESH_UV_SPHERE) {
    const int ip = static_cast<int>(c / a.n_theta), it = static_cast<int>(c % a.n_theta);
    const double psi = static_cast<double>(ip) * a.dpsi;
    const double theta = it == a.n_theta - 1 ? M_PI : static_cast<double>(it) * a.dtheta;
    if (!(fabs(theta) > 0.0 || psi == 0.0)) return false;
    if (!(fabs(theta) < a.pi_ref || psi == 0.0)) return false;
    const double st = sin(theta);
    x = cos(psi) * st;
    y = sin(psi) * st;
    z = cos(theta);
    return true;
  }
  const long long per_face = static_cast<long long>(a.m) * a.m;
  double vx, vy, vz;
  if (c < 6 * per_face) {
    const int f = static_cast<int>(c / per_face);
    const long long e = c - f * per_face;
    const double gx = cube_coord(a, static_cast<int>(e % a.m));  // meshgrid: x along rows
    const double gy = cube_coord(a, static_cast<int>(e / a.m));
    switch (f) {  // bottom, top, east, west, south, north
      case 0: vx = -gx; vy = -gy; vz = -1.0; break;
      case 1: vx = gx; vy = gy; vz = 1.0; break;
      case 2: vx = 1.0; vy = gx; vz = -gy; break;
      case 3: vx = -1.0; vy = -gx; vz = gy; break;
      case 4: vx = gx; vy = -1.0; vz = gy; break;
      default: vx = -gx; vy = 1.0; vz = -gy; break;
    }
  } else if (c == 6 * per_face) {  // the two corners the faces miss
    vx = -1.0; vy = 1.0; vz = 1.0;
  } else {
    vx = 1.0; vy = -1.0; vz = -1.0;
  }
  const double nrm = sqrt((vx * vx + vy * vy) + vz * vz);
  x = vx / nrm;
  y = vy / nrm;
  z = vz / nrm;
  return true;
}

__device__ __forceinline__ bool grid_keep(const GridArgs &a, long long c, double &x, double &y,
                                          double &z) {
#pragma clang fp contract(off)
  if (c >= a.n_cand || !mesh_vertex(a, c, x, y, z)) return false;
  for (int i = 0; i < a.n_planes; ++i) {
    const double d = (a.plane[i][0] * x + a.plane[i][1] * y) + a.plane[i][2] * z;
    if (!(a.plane[i][3] * d >= -1e-13)) return false;
  }
  return true;
}

template <int kPass>  // 0: count per block; 1: write rows
__global__ __launch_bounds__(kGridThreads) void grid_kernel(GridArgs a, int *work) {
#pragma clang fp contract(off)
  __shared__ int s_wave[kGridThreads / kWave];
  const long long c = static_cast<long long>(blockIdx.x) * kGridThreads + threadIdx.x;
  double x = 0.0, y = 0.0, z = 0.0;
  const bool keep = grid_keep(a, c, x, y, z);
  const uint64_t m = __ballot(keep);
  const int lane = lane_id(), wid = threadIdx.x / kWave;
  if (lane == 0) s_wave[wid] = __popcll(m);
  __syncthreads();
  if (kPass == 0) {
    if (threadIdx.x == 0) {
      int t = 0;
      for (int w = 0; w < kGridThreads / kWave; ++w) t += s_wave[w];
      work[blockIdx.x] = t;
    }
    return;
  }
  if (!keep) return;
  long long row = work[gridDim.x + blockIdx.x] + mask_rank(m);
  for (int w = 0; w < wid; ++w) row += s_wave[w];
  if (row >= a.capacity) return;
  // beam_directions_grid_to_euler: Phi = acos(z/|v|), phi2 = sign * acos(x/|v_xy|)
  const double nrm = sqrt((x * x + y * y) + z * z);
  const double Phi = acos(z / nrm);
  const double nproj = sqrt(x * x + y * y);
  const double sgn = y > 0.0 ? 1.0 : y < 0.0 ? -1.0 : (x > 0.0 ? 1.0 : x < 0.0 ? -1.0 : 0.0);
  double ac = acos(x / nproj);
  if (isnan(ac)) ac = 0.0;  // the pole: nan_to_num
  const double phi2 = sgn * ac;
  const double r2d = 180.0 / M_PI;
  a.euler[3 * row + 0] = 0.0;
  a.euler[3 * row + 1] = Phi * r2d;
  a.euler[3 * row + 2] = (M_PI / 2.0 - phi2) * r2d;
}

// Exclusive scan of the block counts (one block); total to count_out.
__global__ __launch_bounds__(1024) void grid_scan_kernel(int *work, int nblocks, long long *count_out) {
  __shared__ int s[1024];
  __shared__ long long carry;
  const int tid = threadIdx.x;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nblocks; base += 1024) {
    const int i = base + tid;
    const int v = i < nblocks ? work[i] : 0;
    s[tid] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int t = tid >= off ? s[tid - off] : 0;
      __syncthreads();
      s[tid] += t;
      __syncthreads();
    }
    if (i < nblocks) work[nblocks + i] = static_cast<int>(carry) + s[tid] - v;
    __syncthreads();
    if (tid == 1023) carry += s[1023];
    __syncthreads();
  }
  if (tid == 0) *count_out = carry;
}

// Host-side mesh parameters with numpy's arithmetic (FP64).
static bool grid_params(int mesh, double res, GridArgs &a) {
  const double deg = M_PI / 180.0;  // np.deg2rad multiplies by pi/180
  a.mesh = mesh;
  if (mesh == DSX_MESH_UV_SPHERE) {
    a.n_theta = static_cast<int>(ceil(180.0 / res)) + 1;
    a.n_psi = static_cast<int>(ceil(360.0 / res));
    a.dpsi = 2.0 * M_PI / a.n_psi;
    a.dtheta = M_PI / (a.n_theta - 1);
    a.pi_ref = 180.0 * deg;
    a.n_cand = static_cast<long long>(a.n_psi) * a.n_theta;
    return a.n_theta >= 2;
  }
  if (mesh == DSX_MESH_NORMALIZED_CUBE) {
    a.steps = ceil(1.0 / tan(res * deg));
    a.theta = 0.0;
  } else if (mesh == DSX_MESH_SPHERIFIED_EDGE) {
    a.steps = ceil((45.0 * deg) * (180.0 / M_PI) / res);
    a.theta = atan(1.0) / a.steps;
  } else if (mesh == DSX_MESH_SPHERIFIED_CORNER) {
    a.steps = ceil(acos(1.0 / sqrt(3.0)) / (res * deg));
    a.theta = atan(sqrt(2.0)) / a.steps;
  } else {
    return false;
  }
  a.sqrt2 = sqrt(2.0);
  if (!(a.steps >= 1.0) || a.steps > 8192.0) return false;
  a.m = static_cast<int>(2.0 * a.steps);
  a.n_cand = 6ll * a.m * a.m + 2;
  return true;
}

}  // namespace dsx

using namespace dsx;

extern "C" long long dsx_beam_directions_candidates(int mesh, double resolution) {
  GridArgs a;
  if (!(resolution > 0.0) || !grid_params(mesh, resolution, a)) return -1;
  return a.n_cand;
}

extern "C" long long dsx_beam_directions_workspace_ints(int mesh, double resolution) {
  const long long n = dsx_beam_directions_candidates(mesh, resolution);
  return n < 0 ? -1 : 2 * ((n + kGridThreads - 1) / kGridThreads);
}

extern "C" int dsx_beam_directions(int mesh, double resolution, int n_planes, const int *planes,
                                   double *euler_d, long long capacity, long long *count_d,
                                   int *work_d, long long work_ints, void *stream) {
  clear_error();
  GridArgs a;
  if (!(resolution > 0.0) || !grid_params(mesh, resolution, a)) {
    set_error("dsx_beam_directions: unknown mesh %d or bad resolution %g", mesh, resolution);
    return DSX_EINVAL;
  }
  if (n_planes < 0 || n_planes > 3 || (n_planes > 0 && !planes) || capacity < 0) {
    set_error("dsx_beam_directions: bad crop planes (%d) or capacity", n_planes);
    return DSX_EINVAL;
  }
  const long long nb = (a.n_cand + kGridThreads - 1) / kGridThreads;
  if (nb > (1ll << 30)) {
    set_error("dsx_beam_directions: %lld candidates is too many", a.n_cand);
    return DSX_EINVAL;
  }
  if (!euler_d || !count_d || !work_d || work_ints < 2 * nb) {
    set_error("dsx_beam_directions: null pointer or workspace of %lld ints < %lld", work_ints, 2 * nb);
    return DSX_EINVAL;
  }
  a.n_planes = n_planes;
  for (int i = 0; i < n_planes; ++i)
    for (int j = 0; j < 4; ++j) a.plane[i][j] = static_cast<double>(planes[4 * i + j]);
  a.euler = euler_d;
  a.capacity = capacity;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(grid_kernel<0>, dim3(static_cast<unsigned>(nb)), dim3(kGridThreads), 0, s, a, work_d);
  hipLaunchKernelGGL(grid_scan_kernel, dim3(1), dim3(1024), 0, s, work_d, static_cast<int>(nb), count_d);
  hipLaunchKernelGGL(grid_kernel<1>, dim3(static_cast<unsigned>(nb)), dim3(kGridThreads), 0, s, a, work_d);
  return check_launch("dsx_beam_directions");
}
