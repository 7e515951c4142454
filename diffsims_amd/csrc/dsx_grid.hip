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
#include <algorithm>

#include "dsx_internal.h"

namespace dsx {

constexpr int kGridThreads = 256;

// Icosahedron of get_icosahedral_mesh_vertices (sphere_mesh_generators.py:394-431)
// with the caller's edge order (the order the reference's edge set iterates in).
struct IcosaTopo {
  double corner[12][3];
  int face[20][3];
  int edge[30][2];              // sorted corner pairs, in the caller's order
  signed char edge_id[12][12];  // edge holding a corner pair, -1 if none
  signed char edge_face[30][2]; // the two faces of each edge
  signed char corner_face[12][5];
};

struct GridArgs {
  int mesh;  // DSX_MESH_*, or kMeshIcosahedral / kMeshVertices
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
  int xyz_out;  // write the (kept) vertices instead of Bunge angles
  // icosahedral mesh at refinement level `level` (10 level^2 + 2 vertices)
  int level;
  IcosaTopo ico;
  // explicit vertices (double [n_cand][3])
  const double *verts;
};

constexpr int kMeshIcosahedral = 100;
constexpr int kMeshVertices = 101;

// ---- icosahedral mesh ------------------------------------------------------
//
// Vertex ids at level n (_compose_from_faces, sphere_mesh_generators.py:200-325):
// the 12 corners; per edge e (caller order) n-1 points (1-t) c0 + t c1 with
// t = m step + 1/n, step = (1 - 1/n) / (n-1) (numpy's linspace); per face
// (f0, f1, f2) the interior points of rows i = 1..n-2 (outer) and
// j = 1..n-i-1 (inner) with weights (1 - i/n - j/n, j/n, i/n).  In face-local
// coordinates (a, b), a + b <= n, the point is f0 (1 - a/n - b/n) + f1 b/n +
// f2 a/n, so interior (i, j) is (a, b) = (i, j).

__device__ __forceinline__ long long icosa_count(long long n) { return 10 * n * n + 2; }

__host__ __device__ __forceinline__ void icosa_vertex(const IcosaTopo &t, int n, long long g, double &x,
                                             double &y, double &z) {
#pragma clang fp contract(off)
  if (g < 12) {
    x = t.corner[g][0];
    y = t.corner[g][1];
    z = t.corner[g][2];
    return;
  }
  const long long ne = 30ll * (n - 1);
  if (g < 12 + ne) {
    const int e = static_cast<int>((g - 12) / (n - 1));
    const int m = static_cast<int>((g - 12) - static_cast<long long>(e) * (n - 1));
    const double start = 1.0 / n;
    const double step = (1.0 - start) / (n - 1);
    const double tt = static_cast<double>(m) * step + start;
    const double u = 1.0 - tt;
    const int i0 = t.edge[e][0], i1 = t.edge[e][1];
    x = u * t.corner[i0][0] + tt * t.corner[i1][0];
    y = u * t.corner[i0][1] + tt * t.corner[i1][1];
    z = u * t.corner[i0][2] + tt * t.corner[i1][2];
    return;
  }
  const long long per = static_cast<long long>(n - 1) * (n - 2) / 2;
  const long long r = g - 12 - ne;
  const int f = static_cast<int>(r / per);
  const long long q = r - f * per;
  // row u = i - 1 starts at u (2L - u + 1) / 2, L = n - 2
  const double L2 = 2.0 * (n - 2) + 1.0;
  long long u = static_cast<long long>((L2 - sqrt(L2 * L2 - 8.0 * static_cast<double>(q))) * 0.5);
  auto off = [&](long long uu) { return uu * (2ll * (n - 2) - uu + 1) / 2; };
  while (u > 0 && off(u) > q) --u;
  while (off(u + 1) <= q) ++u;
  const int i = static_cast<int>(u) + 1;
  const int j = static_cast<int>(q - off(u)) + 1;
  const double b0 = static_cast<double>(i) / n, b1 = static_cast<double>(j) / n;
  const double w0 = (1.0 - b0) - b1;
  const int f0 = t.face[f][0], f1 = t.face[f][1], f2 = t.face[f][2];
  x = (t.corner[f0][0] * w0 + t.corner[f1][0] * b1) + t.corner[f2][0] * b0;
  y = (t.corner[f0][1] * w0 + t.corner[f1][1] * b1) + t.corner[f2][1] * b0;
  z = (t.corner[f0][2] * w0 + t.corner[f1][2] * b1) + t.corner[f2][2] * b0;
}

// Vertex id of face-local point (a, b) of face f at level n.
__host__ __device__ __forceinline__ long long icosa_local(const IcosaTopo &t, int n, int f, int a, int b) {
  const int f0 = t.face[f][0], f1 = t.face[f][1], f2 = t.face[f][2];
  if (a == 0 && b == 0) return f0;
  if (a == 0 && b == n) return f1;
  if (a == n && b == 0) return f2;
  int p, q, d;  // edge p-q, distance d/n from p
  if (a == 0) {
    p = f0; q = f1; d = b;
  } else if (b == 0) {
    p = f0; q = f2; d = a;
  } else if (a + b == n) {
    p = f1; q = f2; d = a;
  } else {
    const long long per = static_cast<long long>(n - 1) * (n - 2) / 2;
    const long long u = a - 1;
    return 12 + 30ll * (n - 1) + f * per + u * (2ll * (n - 2) - u + 1) / 2 + (b - 1);
  }
  const int e = t.edge_id[p][q];
  const int m = p < q ? d - 1 : n - d - 1;
  return 12 + static_cast<long long>(e) * (n - 1) + m;
}

__host__ __device__ __forceinline__ void unit(double &x, double &y, double &z) {
#pragma clang fp contract(off)
  const double nrm = sqrt((x * x + y * y) + z * z);
  x = x / nrm;
  y = y / nrm;
  z = z / nrm;
}

// Largest dot product between unit vertex g and its neighbours in face f,
// where g sits at face-local (a, b).  The nearest neighbour of a vertex of
// this mesh is one of its triangulation neighbours: radial projection
// distorts a face's triangles by at most a factor 1/cos(37.4 deg) = 1.26,
// while a vertex two steps away is at least sqrt(3) steps from it.
__host__ __device__ __forceinline__ double icosa_face_best(const IcosaTopo &t, int n, int f, int a, int b,
                                                  double ux, double uy, double uz, double best) {
#pragma clang fp contract(off)
  const int da[6] = {1, -1, 0, 0, 1, -1};
  const int db[6] = {0, 0, 1, -1, -1, 1};
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int a2 = a + da[k], b2 = b + db[k];
    if (a2 < 0 || b2 < 0 || a2 + b2 > n) continue;
    double x, y, z;
    icosa_vertex(t, n, icosa_local(t, n, f, a2, b2), x, y, z);
    unit(x, y, z);
    const double d = (ux * x + uy * y) + uz * z;
    best = d > best ? d : best;
  }
  return best;
}

// Dot product of vertex g (level n) with its nearest neighbour.
__host__ __device__ double icosa_nn_dot(const IcosaTopo &t, int n, long long g) {
  double ux, uy, uz;
  icosa_vertex(t, n, g, ux, uy, uz);
  unit(ux, uy, uz);
  double best = -2.0;
  if (g < 12) {
    const int c = static_cast<int>(g);
    for (int k = 0; k < 5; ++k) {
      const int f = t.corner_face[c][k];
      const int a = t.face[f][2] == c ? n : 0;
      const int b = t.face[f][1] == c ? n : 0;
      best = icosa_face_best(t, n, f, a, b, ux, uy, uz, best);
    }
    return best;
  }
  const long long ne = 30ll * (n - 1);
  if (g < 12 + ne) {
    const int e = static_cast<int>((g - 12) / (n - 1));
    const int m = static_cast<int>((g - 12) - static_cast<long long>(e) * (n - 1));
    const int i0 = t.edge[e][0], i1 = t.edge[e][1];
    for (int k = 0; k < 2; ++k) {
      const int f = t.edge_face[e][k];
      const int f0 = t.face[f][0], f1 = t.face[f][1], f2 = t.face[f][2];
      int p, q, a = 0, b = 0;  // the face side holding the edge, from p to q
      if ((f0 == i0 && f1 == i1) || (f0 == i1 && f1 == i0)) {
        p = f0; q = f1;
      } else if ((f0 == i0 && f2 == i1) || (f0 == i1 && f2 == i0)) {
        p = f0; q = f2;
      } else {
        p = f1; q = f2;
      }
      const int d = p < q ? m + 1 : n - m - 1;  // distance from p in steps
      if (p == f0 && q == f1) {
        b = d;
      } else if (p == f0 && q == f2) {
        a = d;
      } else {
        a = d;
        b = n - d;
      }
      best = icosa_face_best(t, n, f, a, b, ux, uy, uz, best);
    }
    return best;
  }
  const long long per = static_cast<long long>(n - 1) * (n - 2) / 2;
  const long long r = g - 12 - ne;
  const int f = static_cast<int>(r / per);
  const long long q = r - f * per;
  const double L2 = 2.0 * (n - 2) + 1.0;
  long long u = static_cast<long long>((L2 - sqrt(L2 * L2 - 8.0 * static_cast<double>(q))) * 0.5);
  auto off = [&](long long uu) { return uu * (2ll * (n - 2) - uu + 1) / 2; };
  while (u > 0 && off(u) > q) --u;
  while (off(u + 1) <= q) ++u;
  return icosa_face_best(t, n, f, static_cast<int>(u) + 1, static_cast<int>(q - off(u)) + 1, ux,
                         uy, uz, best);
}

// Vertices of levels first..: S(L) = sum_{n=1..L} (10 n^2 + 2).
__host__ __device__ __forceinline__ long long icosa_cum(long long L) {
  return 10 * L * (L + 1) * (2 * L + 1) / 6 + 2 * L;
}

constexpr int kIcoThreads = 256;

__global__ __launch_bounds__(kIcoThreads) void icosa_init_kernel(unsigned long long *bits, int n) {
  const int i = blockIdx.x * kIcoThreads + threadIdx.x;
  if (i < n) bits[i] = 0x7FF0000000000000ull;  // +inf
}

// Smallest nearest-neighbour dot product of each level in [first, last]
// (stored as the bits of a positive double, so unsigned order is value order).
__global__ __launch_bounds__(kIcoThreads) void icosa_nn_kernel(IcosaTopo t, int first, int last,
                                                               unsigned long long *bits) {
  const long long base = icosa_cum(first - 1);
  const long long total = icosa_cum(last) - base;
  for (long long w0 = static_cast<long long>(blockIdx.x) * kIcoThreads; w0 < total;
       w0 += static_cast<long long>(gridDim.x) * kIcoThreads) {
    const long long flat = w0 + threadIdx.x;
    double d = 2.0;
    int lv = 0;
    if (flat < total) {
      const long long gf = base + flat;
      // level: largest L with icosa_cum(L - 1) <= gf
      long long L = static_cast<long long>(cbrt(static_cast<double>(gf) * 0.3));
      if (L < 1) L = 1;
      while (L > 1 && icosa_cum(L - 1) > gf) --L;
      while (icosa_cum(L) <= gf) ++L;
      lv = static_cast<int>(L);
      d = icosa_nn_dot(t, lv, gf - icosa_cum(L - 1));
    }
    // one atomic per level present in the wave
    const int lo = __shfl(lv, 0, kWave);
    for (int L = lo;; ++L) {
      double m = lv == L ? d : 2.0;
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) m = fmin(m, __shfl_xor(m, o, kWave));
      if (lane_id() == 0 && m < 2.0)
        atomicMin(bits + (L - first), static_cast<unsigned long long>(__double_as_longlong(m)));
      const uint64_t more = __ballot(lv > L);
      if (!more) break;
    }
  }
}

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
  if (a.mesh == kMeshVertices) {
    x = a.verts[3 * c];
    y = a.verts[3 * c + 1];
    z = a.verts[3 * c + 2];
    return true;
  }
  if (a.mesh == kMeshIcosahedral) {  // rows divided by sqrt(einsum('ij,ij->i'))
    icosa_vertex(a.ico, a.level, c, x, y, z);
    unit(x, y, z);
    return true;
  }
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
  if (a.xyz_out) {
    a.euler[3 * row + 0] = x;
    a.euler[3 * row + 1] = y;
    a.euler[3 * row + 2] = z;
    return;
  }
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

// Count / scan / write launch shared by the grid entry points.
static int launch_grid(GridArgs &a, int n_planes, const int *planes, double *out_d,
                       long long capacity, long long *count_d, int *work_d, long long work_ints,
                       void *stream, const char *name) {
  if (n_planes < 0 || n_planes > 3 || (n_planes > 0 && !planes) || capacity < 0) {
    set_error("%s: bad crop planes (%d) or capacity", name, n_planes);
    return DSX_EINVAL;
  }
  const long long nb = (a.n_cand + kGridThreads - 1) / kGridThreads;
  if (nb > (1ll << 30)) {
    set_error("%s: %lld candidates is too many", name, a.n_cand);
    return DSX_EINVAL;
  }
  if (!out_d || !count_d || !work_d || work_ints < 2 * nb) {
    set_error("%s: null pointer or workspace of %lld ints < %lld", name, work_ints, 2 * nb);
    return DSX_EINVAL;
  }
  a.n_planes = n_planes;
  for (int i = 0; i < n_planes; ++i)
    for (int j = 0; j < 4; ++j) a.plane[i][j] = static_cast<double>(planes[4 * i + j]);
  a.euler = out_d;
  a.capacity = capacity;
  hipStream_t s = as_stream(stream);
  if (nb == 0) {
    hipLaunchKernelGGL(grid_scan_kernel, dim3(1), dim3(1024), 0, s, work_d, 0, count_d);
    return check_launch(name);
  }
  hipLaunchKernelGGL(grid_kernel<0>, dim3(static_cast<unsigned>(nb)), dim3(kGridThreads), 0, s, a, work_d);
  hipLaunchKernelGGL(grid_scan_kernel, dim3(1), dim3(1024), 0, s, work_d, static_cast<int>(nb), count_d);
  hipLaunchKernelGGL(grid_kernel<1>, dim3(static_cast<unsigned>(nb)), dim3(kGridThreads), 0, s, a, work_d);
  return check_launch(name);
}

// Icosahedron topology from the caller's edge order (30 sorted corner pairs);
// false unless they are exactly the 30 edges of the faces.
bool icosa_topology(const int *edge_order, IcosaTopo &t) {
  static const int faces[20][3] = {
      {0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11}, {1, 5, 9},  {5, 11, 4},
      {11, 10, 2}, {10, 7, 6}, {7, 1, 8},  {3, 9, 4},  {3, 4, 2},   {3, 2, 6},  {3, 6, 8},
      {3, 8, 9},  {4, 9, 5},  {2, 4, 11}, {6, 2, 10}, {8, 6, 7},   {9, 8, 1}};
  const double g = (1.0 + sqrt(5.0)) / 2.0;
  const double c[12][3] = {{-1, g, 0}, {1, g, 0},  {-1, -g, 0}, {1, -g, 0},
                           {0, -1, g}, {0, 1, g},  {0, -1, -g}, {0, 1, -g},
                           {g, 0, -1}, {g, 0, 1},  {-g, 0, -1}, {-g, 0, 1}};
  for (int i = 0; i < 12; ++i)
    for (int k = 0; k < 3; ++k) t.corner[i][k] = c[i][k];
  for (int f = 0; f < 20; ++f)
    for (int k = 0; k < 3; ++k) t.face[f][k] = faces[f][k];
  for (int i = 0; i < 12; ++i)
    for (int j = 0; j < 12; ++j) t.edge_id[i][j] = -1;
  if (!edge_order) return false;
  for (int e = 0; e < 30; ++e) {
    const int p = edge_order[2 * e], q = edge_order[2 * e + 1];
    if (p < 0 || q > 11 || p >= q || t.edge_id[p][q] >= 0) return false;
    t.edge[e][0] = p;
    t.edge[e][1] = q;
    t.edge_id[p][q] = t.edge_id[q][p] = static_cast<signed char>(e);
    t.edge_face[e][0] = t.edge_face[e][1] = -1;
  }
  int n_corner_faces[12] = {0};
  for (int f = 0; f < 20; ++f) {
    for (int k = 0; k < 3; ++k) {
      const int p = faces[f][k], q = faces[f][(k + 1) % 3];
      const int e = t.edge_id[p][q];
      if (e < 0) return false;
      signed char *ef = t.edge_face[e];
      if (ef[0] < 0) ef[0] = static_cast<signed char>(f);
      else if (ef[1] < 0) ef[1] = static_cast<signed char>(f);
      else return false;
      if (n_corner_faces[p] >= 5) return false;
      t.corner_face[p][n_corner_faces[p]++] = static_cast<signed char>(f);
    }
  }
  for (int e = 0; e < 30; ++e)
    if (t.edge_face[e][1] < 0) return false;
  return true;
}

constexpr int kIcoMaxLevel = 1 << 14;

// ---- SO(3) cubochoric grids ------------------------------------------------
//
// get_fundamental_zone_grid / get_local_grid (rotation_list_generators.py:
// 85-134) sample SO(3) on the cubochoric grid (Rosca, Morawiec & De Graef,
// Modelling Simul. Mater. Sci. Eng. 22 (2014) 075013): points i a/n,
// i = -n+1 .. n, of the cube of semi-edge a = pi^(2/3)/2 (the -a faces are
// antipodes of the +a faces), mapped volume-preservingly onto the homochoric
// ball and from there to rotations.  One thread per grid point, FP64 with
// contraction off; rows leave in grid order (x outermost, z innermost) through
// the count / scan / write compaction of the sphere meshes.  Bunge angles
// follow Rowenhorst et al., Modelling Simul. Mater. Sci. Eng. 23 (2015)
// 083501, eq. A.14 with P = +1.

constexpr int kSo3MaxSteps = 1024;
constexpr int kSo3NewtonSteps = 6;
constexpr double kFzEps = 1e-12;

struct So3Args {
  int n;             // semi-edge steps
  long long side;    // 2 n points per axis
  long long n_cand;  // side^3
  double step;       // a / n
  // map constants, formed on the host: (pi/6)^(1/6); (3pi/4)^(1/3) 2^(1/4) /
  // (pi^(5/6) / (2 6^(1/6))); sqrt(6/pi); sqrt(2); pi/12; sqrt(pi); sqrt(24)
  double sc, prek, pref, sqrt2, pi12, sqrtpi, sqrt24;
  int mode;          // DSX_SO3_FUNDAMENTAL_ZONE or DSX_SO3_LOCAL
  int n_ops;         // group rotations other than the identity
  double ops[24][4];
  double half;       // local: keep acos(w) < half
  int has_center;
  double center[4];
  double *euler;
  long long capacity;
};

// Cube -> homochoric ball.  The dominant axis (z, then x, then y on ties)
// is rotated onto z; the square at height Z goes area-preservingly onto the
// Lambert-azimuthal preimage of the curved square, then onto the sphere of
// radius sqrt(6/pi) Z.
__device__ __forceinline__ void cubo_to_homo(const So3Args &a, double x, double y, double z,
                                             double h[3]) {
#pragma clang fp contract(off)
  const double ax = fabs(x), ay = fabs(y), az = fabs(z);
  int pyr;  // 0: z, 1: x, 2: y
  double s0, s1, s2;
  if (fmax(ax, ay) <= az) {
    pyr = 0; s0 = x; s1 = y; s2 = z;
  } else if (fmax(az, ay) <= ax) {
    pyr = 1; s0 = y; s1 = z; s2 = x;
  } else {
    pyr = 2; s0 = z; s1 = x; s2 = y;
  }
  const double X = a.sc * s0, Y = a.sc * s1, Z = a.sc * s2;
  double L0 = 0.0, L1 = 0.0, L2;
  if (X == 0.0 && Y == 0.0) {
    L2 = a.pref * Z;
  } else {
    const bool ylex = fabs(Y) <= fabs(X);
    const double num = ylex ? Y : X, den = ylex ? X : Y;
    const double t = a.pi12 * num / den;
    double sn, c;
    sincos(t, &sn, &c);
    const double q = a.prek * den / sqrt(a.sqrt2 - c);
    const double u = (a.sqrt2 * c - 1.0) * q;
    const double v = a.sqrt2 * sn * q;
    const double T1 = ylex ? u : v, T2 = ylex ? v : u;
    const double rho2 = T1 * T1 + T2 * T2;
    const double lift = sqrt(1.0 - M_PI * rho2 / (24.0 * Z * Z));
    L0 = T1 * lift;
    L1 = T2 * lift;
    L2 = a.pref * Z - a.sqrtpi * rho2 / a.sqrt24 / Z;
  }
  if (pyr == 0) {
    h[0] = L0; h[1] = L1; h[2] = L2;
  } else if (pyr == 1) {
    h[0] = L2; h[1] = L0; h[2] = L1;
  } else {
    h[0] = L1; h[1] = L2; h[2] = L0;
  }
}

// w - sin(w), by its Taylor series below 0.5 (no cancellation).
__device__ __forceinline__ double omega_minus_sin(double w) {
#pragma clang fp contract(off)
  if (w < 0.5) {
    const double w2 = w * w;
    return w * w2 / 6.0 *
           (1.0 - w2 / 20.0 * (1.0 - w2 / 42.0 * (1.0 - w2 / 72.0 * (1.0 - w2 / 110.0 *
                                                                    (1.0 - w2 / 156.0)))));
  }
  return w - sin(w);
}

// Homochoric vector -> unit quaternion (w >= 0): |h|^3 = (3/4)(w - sin w)
// by a fixed number of Newton steps from w0 = 2|h|, which lies below the root
// of this convex function, so the steps converge from above; four reach the
// last bit anywhere on [0, pi].
__device__ __forceinline__ void homo_to_quat(const double h[3], double q[4]) {
#pragma clang fp contract(off)
  const double hn = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
  const double t = 4.0 / 3.0 * (hn * hn * hn);
  double w = fmin(2.0 * hn, M_PI);
  for (int it = 0; it < kSo3NewtonSteps; ++it) {
    const double f = omega_minus_sin(w) - t;
    const double sh = sin(0.5 * w);
    const double d = 2.0 * (sh * sh);
    const double st = d > 0.0 ? f / d : 0.0;
    w = fmin(fmax(w - st, 0.0), M_PI);
  }
  double sw, cw;
  sincos(0.5 * w, &sw, &cw);
  q[0] = cw;
  if (hn > 0.0) {
    q[1] = sw * (h[0] / hn);
    q[2] = sw * (h[1] / hn);
    q[3] = sw * (h[2] / hn);
  } else {
    q[1] = 0.0;
    q[2] = 0.0;
    q[3] = sw;
  }
}

__device__ __forceinline__ void quat_mul(const double p[4], const double q[4], double r[4]) {
#pragma clang fp contract(off)
  r[0] = p[0] * q[0] - p[1] * q[1] - p[2] * q[2] - p[3] * q[3];
  r[1] = p[0] * q[1] + p[1] * q[0] + p[2] * q[3] - p[3] * q[2];
  r[2] = p[0] * q[2] - p[1] * q[3] + p[2] * q[0] + p[3] * q[1];
  r[3] = p[0] * q[3] + p[1] * q[2] - p[2] * q[1] + p[3] * q[0];
}

// Bunge angles (radians) of a quaternion: Rowenhorst et al. (2015) eq. A.14,
// P = +1, with the Phi = 0 and Phi = pi branches; negative angles moved into
// [0, 2 pi) (phi1, phi2) and [0, pi] (Phi).
__device__ __forceinline__ void quat_to_bunge(const double qin[4], double e[3]) {
#pragma clang fp contract(off)
  const double nrm = sqrt(qin[0] * qin[0] + qin[1] * qin[1] + qin[2] * qin[2] + qin[3] * qin[3]);
  const double w = qin[0] / nrm, x = qin[1] / nrm, y = qin[2] / nrm, z = qin[3] / nrm;
  const double q03 = w * w + z * z, q12 = x * x + y * y;
  const double chi = sqrt(q03 * q12);
  if (chi == 0.0 && q12 == 0.0) {
    e[0] = atan2(-2.0 * w * z, w * w - z * z);
    e[1] = 0.0;
    e[2] = 0.0;
  } else if (chi == 0.0 && q03 == 0.0) {
    e[0] = atan2(2.0 * x * y, x * x - y * y);
    e[1] = M_PI;
    e[2] = 0.0;
  } else {
    e[0] = atan2((x * z - w * y) / chi, (-w * x - y * z) / chi);
    e[1] = atan2(2.0 * chi, q03 - q12);
    e[2] = atan2((w * y + x * z) / chi, (y * z - w * x) / chi);
  }
  if (e[0] < 0.0) e[0] = fmod(e[0] + 100.0 * M_PI, 2.0 * M_PI);
  if (e[1] < 0.0) e[1] = fmod(e[1] + 100.0 * M_PI, M_PI);
  if (e[2] < 0.0) e[2] = fmod(e[2] + 100.0 * M_PI, 2.0 * M_PI);
}

__device__ __forceinline__ bool so3_keep(const So3Args &a, long long c, double q[4]) {
#pragma clang fp contract(off)
  if (c >= a.n_cand) return false;
  const long long i = c / (a.side * a.side);
  const long long j = (c / a.side) % a.side;
  const long long k = c % a.side;
  const double x = static_cast<double>(i - a.n + 1) * a.step;
  const double y = static_cast<double>(j - a.n + 1) * a.step;
  const double z = static_cast<double>(k - a.n + 1) * a.step;
  double h[3];
  cubo_to_homo(a, x, y, z, h);
  homo_to_quat(h, q);
  if (a.mode == DSX_SO3_FUNDAMENTAL_ZONE) {
    double worst = 0.0;
    for (int s = 0; s < a.n_ops; ++s) {
      const double d = fabs(q[0] * a.ops[s][0] + q[1] * a.ops[s][1] + q[2] * a.ops[s][2] +
                            q[3] * a.ops[s][3]);
      worst = fmax(worst, d);
    }
    return a.n_ops == 0 || fabs(q[0]) - worst >= -kFzEps;
  }
  if (!(acos(q[0]) < a.half)) return false;
  if (a.has_center) {
    double r[4];
    quat_mul(a.center, q, r);
    q[0] = r[0]; q[1] = r[1]; q[2] = r[2]; q[3] = r[3];
  }
  return true;
}

template <int kPass>  // 0: count per block; 1: write rows
__global__ __launch_bounds__(kGridThreads) void so3_grid_kernel(So3Args a, int *work) {
  __shared__ int s_wave[kGridThreads / kWave];
  const long long c = static_cast<long long>(blockIdx.x) * kGridThreads + threadIdx.x;
  double q[4];
  const bool keep = so3_keep(a, c, q);
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
  const long long nb = gridDim.x;
  const long long *offs = reinterpret_cast<const long long *>(work + ((nb + 1) & ~1ll));
  long long row = offs[blockIdx.x] + mask_rank(m);
  for (int w = 0; w < wid; ++w) row += s_wave[w];
  if (row >= a.capacity) return;
  double e[3];
  quat_to_bunge(q, e);
  a.euler[3 * row + 0] = e[0];
  a.euler[3 * row + 1] = e[1];
  a.euler[3 * row + 2] = e[2];
}

// 64-bit exclusive scan of the block counts (the SO(3) grids can pass 2^31
// rows): block offsets as long long in the workspace after the counts.
__global__ __launch_bounds__(1024) void so3_scan_kernel(int *work, long long nblocks,
                                                        long long *count_out) {
  __shared__ long long s[1024];
  __shared__ long long carry;
  const int tid = threadIdx.x;
  long long *offs = reinterpret_cast<long long *>(work + ((nblocks + 1) & ~1ll));
  if (tid == 0) carry = 0;
  __syncthreads();
  for (long long base = 0; base < nblocks; base += 1024) {
    const long long i = base + tid;
    const long long v = i < nblocks ? work[i] : 0;
    s[tid] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const long long t = tid >= off ? s[tid - off] : 0;
      __syncthreads();
      s[tid] += t;
      __syncthreads();
    }
    if (i < nblocks) offs[i] = carry + s[tid] - v;
    __syncthreads();
    if (tid == 1023) carry += s[1023];
    __syncthreads();
  }
  if (tid == 0) *count_out = carry;
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
  a.xyz_out = 0;
  if (!(resolution > 0.0) || !grid_params(mesh, resolution, a)) {
    set_error("dsx_beam_directions: unknown mesh %d or bad resolution %g", mesh, resolution);
    return DSX_EINVAL;
  }
  return launch_grid(a, n_planes, planes, euler_d, capacity, count_d, work_d, work_ints, stream,
                     "dsx_beam_directions");
}

extern "C" long long dsx_grid_workspace_ints(long long n_vertices) {
  return n_vertices < 0 ? -1 : 2 * ((n_vertices + kGridThreads - 1) / kGridThreads);
}

extern "C" long long dsx_icosahedral_vertices(int level) {
  return level < 1 || level > kIcoMaxLevel ? -1 : 10ll * level * level + 2;
}

extern "C" int dsx_icosahedral_nn_dots(int level_first, int level_last, const int *edge_order,
                                       double *min_dot_d, void *stream) {
  clear_error();
  IcosaTopo t;
  if (level_first < 1 || level_last < level_first || level_last > kIcoMaxLevel) {
    set_error("dsx_icosahedral_nn_dots: bad level range [%d, %d] (1..%d)", level_first,
              level_last, kIcoMaxLevel);
    return DSX_EINVAL;
  }
  if (!min_dot_d || !icosa_topology(edge_order, t)) {
    set_error("dsx_icosahedral_nn_dots: null output or the edge order is not the 30 edges of "
              "the icosahedron as sorted corner pairs");
    return DSX_EINVAL;
  }
  hipStream_t s = as_stream(stream);
  const int nl = level_last - level_first + 1;
  auto *bits = reinterpret_cast<unsigned long long *>(min_dot_d);
  hipLaunchKernelGGL(icosa_init_kernel, dim3((nl + kIcoThreads - 1) / kIcoThreads),
                     dim3(kIcoThreads), 0, s, bits, nl);
  const long long L1 = level_first - 1, L2 = level_last;
  const long long total = (10 * L2 * (L2 + 1) * (2 * L2 + 1) / 6 + 2 * L2) -
                          (10 * L1 * (L1 + 1) * (2 * L1 + 1) / 6 + 2 * L1);
  const long long want = (total + kIcoThreads - 1) / kIcoThreads;
  const int blocks = static_cast<int>(std::min<long long>(want, 64ll * cu_count()));
  hipLaunchKernelGGL(icosa_nn_kernel, dim3(blocks), dim3(kIcoThreads), 0, s, t, level_first,
                     level_last, bits);
  return check_launch("dsx_icosahedral_nn_dots");
}

extern "C" int dsx_beam_directions_icosahedral(int level, const int *edge_order, int n_planes,
                                               const int *planes, int xyz_out, double *out_d,
                                               long long capacity, long long *count_d,
                                               int *work_d, long long work_ints, void *stream) {
  clear_error();
  GridArgs a;
  if (level < 1 || level > kIcoMaxLevel || !icosa_topology(edge_order, a.ico)) {
    set_error("dsx_beam_directions_icosahedral: bad level %d or edge order", level);
    return DSX_EINVAL;
  }
  a.mesh = kMeshIcosahedral;
  a.level = level;
  a.n_cand = 10ll * level * level + 2;
  a.xyz_out = xyz_out ? 1 : 0;
  return launch_grid(a, n_planes, planes, out_d, capacity, count_d, work_d, work_ints, stream,
                     "dsx_beam_directions_icosahedral");
}

extern "C" int dsx_beam_directions_from_vertices(const double *xyz_d, long long n, int n_planes,
                                                 const int *planes, double *euler_d,
                                                 long long capacity, long long *count_d,
                                                 int *work_d, long long work_ints, void *stream) {
  clear_error();
  GridArgs a;
  if (n < 0 || (n > 0 && !xyz_d)) {
    set_error("dsx_beam_directions_from_vertices: bad vertex array (n = %lld)", n);
    return DSX_EINVAL;
  }
  a.mesh = kMeshVertices;
  a.verts = xyz_d;
  a.n_cand = n;
  a.xyz_out = 0;
  return launch_grid(a, n_planes, planes, euler_d, capacity, count_d, work_d, work_ints, stream,
                     "dsx_beam_directions_from_vertices");
}

extern "C" long long dsx_so3_grid_candidates(int semi_edge_steps) {
  if (semi_edge_steps < 1 || semi_edge_steps > kSo3MaxSteps) return -1;
  const long long side = 2ll * semi_edge_steps;
  return side * side * side;
}

extern "C" long long dsx_so3_grid_workspace_ints(int semi_edge_steps) {
  const long long n = dsx_so3_grid_candidates(semi_edge_steps);
  if (n < 0) return -1;
  const long long nb = (n + kGridThreads - 1) / kGridThreads;
  return ((nb + 1) & ~1ll) + 2 * nb;
}

extern "C" int dsx_so3_grid(int semi_edge_steps, int mode, int n_ops, const double *ops,
                            double half_angle, const double *center, double *euler_d,
                            long long capacity, long long *count_d, int *work_d,
                            long long work_ints, void *stream) {
  clear_error();
  const long long n_cand = dsx_so3_grid_candidates(semi_edge_steps);
  if (n_cand < 0) {
    set_error("dsx_so3_grid: semi-edge steps %d outside 1..%d", semi_edge_steps, kSo3MaxSteps);
    return DSX_EINVAL;
  }
  if ((mode != DSX_SO3_FUNDAMENTAL_ZONE && mode != DSX_SO3_LOCAL) || n_ops < 0 || n_ops > 24 ||
      (n_ops > 0 && !ops) || capacity < 0 || (euler_d == nullptr && capacity > 0)) {
    set_error("dsx_so3_grid: bad mode %d, group of %d rotations or capacity %lld", mode, n_ops,
              capacity);
    return DSX_EINVAL;
  }
  const long long nb = (n_cand + kGridThreads - 1) / kGridThreads;
  if (!count_d || !work_d || work_ints < dsx_so3_grid_workspace_ints(semi_edge_steps)) {
    set_error("dsx_so3_grid: null pointer or workspace of %lld ints < %lld", work_ints,
              dsx_so3_grid_workspace_ints(semi_edge_steps));
    return DSX_EINVAL;
  }
  So3Args a;
  a.n = semi_edge_steps;
  a.side = 2ll * semi_edge_steps;
  a.n_cand = n_cand;
  const double semi = 0.5 * pow(M_PI, 2.0 / 3.0);
  a.step = semi / semi_edge_steps;
  a.sc = pow(M_PI / 6.0, 1.0 / 6.0);
  const double beta = 0.5 * pow(M_PI, 5.0 / 6.0) / pow(6.0, 1.0 / 6.0);
  a.prek = pow(0.75 * M_PI, 1.0 / 3.0) * pow(2.0, 0.25) / beta;
  a.pref = sqrt(6.0 / M_PI);
  a.sqrt2 = sqrt(2.0);
  a.pi12 = M_PI / 12.0;
  a.sqrtpi = sqrt(M_PI);
  a.sqrt24 = sqrt(24.0);
  a.mode = mode;
  a.n_ops = n_ops;
  for (int s = 0; s < n_ops; ++s)
    for (int k = 0; k < 4; ++k) a.ops[s][k] = ops[4 * s + k];
  a.half = half_angle;
  a.has_center = center != nullptr;
  for (int k = 0; k < 4; ++k) a.center[k] = center ? center[k] : 0.0;
  a.euler = euler_d;
  a.capacity = euler_d ? capacity : 0;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(so3_grid_kernel<0>, dim3(static_cast<unsigned>(nb)), dim3(kGridThreads), 0, s,
                     a, work_d);
  hipLaunchKernelGGL(so3_scan_kernel, dim3(1), dim3(1024), 0, s, work_d, nb, count_d);
  if (euler_d)
    hipLaunchKernelGGL(so3_grid_kernel<1>, dim3(static_cast<unsigned>(nb)), dim3(kGridThreads), 0,
                       s, a, work_d);
  return check_launch("dsx_so3_grid");
}
