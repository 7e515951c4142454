// Per-phase tables: (a) integer hkl enumeration with the reference's box rule,
// and the kinematical structure factor |F(hkl)|^2 with g-vectors.
//
// hkl enumeration follows ReciprocalLatticeVector.from_min_dspacing
// (diffsims/crystallography/reciprocal_lattice_vector.py:1077-1142): a box of
// integer indices, a d-spacing filter in FP64, (000) dropped, and the order
// orix's Vector3d.unique leaves (descending lexicographic; see hkl_index_to_hkl).
// The structure factor follows _get_kinematical_structure_factor
// (diffsims/utils/sim_utils.py:256-304), computed once per reflection instead of
// once per orientation.
#include "dsx_internal.h"
#include "dsx_physics.h"

namespace dsx {

struct Metric {
  double g00, g11, g22, g01, g02, g12;
};

struct Box {
  int H, K, L;
  int nk, nl;      // 2K+1, 2L+1
  long long total; // (2H+1)(2K+1)(2L+1)
};

constexpr int kEnumThreads = 256;
constexpr int kEnumItems = 8;  // consecutive box entries per thread
constexpr int kEnumPerBlock = kEnumThreads * kEnumItems;

// Box entry i (row-major over descending h, k, l; ascending with kAscend,
// the order of itertools.product over ascending aranges).
template <bool kAscend = false>
__device__ __forceinline__ void hkl_index_to_hkl(long long i, const Box &b, int &h, int &k,
                                                 int &l) {
  const long long plane = static_cast<long long>(b.nk) * b.nl;
  const int ih = static_cast<int>(i / plane);
  const int rem = static_cast<int>(i - static_cast<long long>(ih) * plane);
  const int ik = rem / b.nl;
  const int il = rem - ik * b.nl;
  h = kAscend ? ih - b.H : b.H - ih;
  k = kAscend ? ik - b.K : b.K - ik;
  l = kAscend ? il - b.L : b.L - il;
}

// d-spacing filter in IEEE FP64 with no contraction, in the same operation
// order as the host-side reference metric evaluation, so both agree bit for bit.
__device__ __forceinline__ bool keep_hkl(int h, int k, int l, const Metric &m, double dmin) {
#pragma clang fp contract(off)
  if ((h | k | l) == 0) return false;
  const double H = h, K = k, L = l;
  double q = H * H * m.g00;
  q = q + K * K * m.g11;
  q = q + L * L * m.g22;
  double cross = H * K * m.g01;
  cross = cross + H * L * m.g02;
  cross = cross + K * L * m.g12;
  q = q + 2.0 * cross;
  const double d = 1.0 / sqrt(q);
  return d >= dmin;
}

// Sphere test of the legacy generator's get_points_in_sphere
// (diffsims/utils/sim_utils.py:436-474): |g| < radius from the same FP64
// quadratic form, (000) kept.
__device__ __forceinline__ bool keep_sphere(int h, int k, int l, const Metric &m, double radius) {
#pragma clang fp contract(off)
  const double H = h, K = k, L = l;
  double q = H * H * m.g00;
  q = q + K * K * m.g11;
  q = q + L * L * m.g22;
  double cross = H * K * m.g01;
  cross = cross + H * L * m.g02;
  cross = cross + K * L * m.g12;
  q = q + 2.0 * cross;
  return sqrt(q) < radius;
}

// Pass 0 counts survivors per block; pass 1 writes them, order preserved.
// kSphere: ascending box order and keep_sphere (dmin is then the radius).
template <int kPass, bool kSphere = false>
__global__ __launch_bounds__(kEnumThreads) void hkl_enum_kernel(Box box, Metric m, double dmin,
                                                                int *__restrict__ work,
                                                                int *__restrict__ hkl_out,
                                                                int capacity) {
  __shared__ int s_scan[kEnumThreads];
  const int tid = threadIdx.x;
  const long long first = static_cast<long long>(blockIdx.x) * kEnumPerBlock + tid * kEnumItems;
  unsigned keep_bits = 0;
  int n = 0;
#pragma unroll
  for (int j = 0; j < kEnumItems; ++j) {
    const long long i = first + j;
    if (i < box.total) {
      int h, k, l;
      hkl_index_to_hkl<kSphere>(i, box, h, k, l);
      if (kSphere ? keep_sphere(h, k, l, m, dmin) : keep_hkl(h, k, l, m, dmin)) {
        keep_bits |= 1u << j;
        ++n;
      }
    }
  }
  if (kPass == 0) {
    // block total
    s_scan[tid] = n;
    __syncthreads();
    for (int off = kEnumThreads / 2; off > 0; off >>= 1) {
      if (tid < off) s_scan[tid] += s_scan[tid + off];
      __syncthreads();
    }
    if (tid == 0) work[blockIdx.x] = s_scan[0];
    return;
  }
  // exclusive scan of n across the block (Hillis-Steele in LDS)
  s_scan[tid] = n;
  __syncthreads();
  for (int off = 1; off < kEnumThreads; off <<= 1) {
    int t = tid >= off ? s_scan[tid - off] : 0;
    __syncthreads();
    s_scan[tid] += t;
    __syncthreads();
  }
  int pos = work[gridDim.x + blockIdx.x] + s_scan[tid] - n;
#pragma unroll
  for (int j = 0; j < kEnumItems; ++j) {
    if (keep_bits & (1u << j)) {
      if (pos < capacity) {
        int h, k, l;
        hkl_index_to_hkl<kSphere>(first + j, box, h, k, l);
        reinterpret_cast<int4 *>(hkl_out)[pos] = make_int4(h, k, l, 0);
      }
      ++pos;
    }
  }
}

// Exclusive scan of per-block counts (one block), total to count_out.
__global__ __launch_bounds__(1024) void hkl_scan_kernel(int *__restrict__ work, int nblocks,
                                                        int *__restrict__ count_out) {
  __shared__ int s[1024];
  __shared__ int carry;
  const int tid = threadIdx.x;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nblocks; base += 1024) {
    const int i = base + tid;
    const int v = i < nblocks ? work[i] : 0;
    s[tid] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      int t = tid >= off ? s[tid - off] : 0;
      __syncthreads();
      s[tid] += t;
      __syncthreads();
    }
    if (i < nblocks) work[nblocks + i] = carry + s[tid] - v;
    __syncthreads();
    if (tid == 1023) carry += s[1023];
    __syncthreads();
  }
  if (tid == 0) *count_out = carry;
}

// ---------------------------------------------------------------------------
// Line index of a reflection table
// ---------------------------------------------------------------------------
//
// A line is a maximal run of consecutive rows with equal (h, k) and l falling
// by one per row -- the rows of one (h, k) column of the hkl box, which the
// enumeration's descending order keeps contiguous.  Per line: first row, row
// count, h, k, l of the first row and the smallest / largest |g|^2 (g4.w) of
// its rows, which bound where along the beam its rows can pass the simulate
// pretest.

constexpr int kLineThreads = 256;

__device__ __forceinline__ bool line_continues(const int4 *hkl, int i) {
  const int4 a = hkl[i - 1], b = hkl[i];
  return a.x == b.x && a.y == b.y && a.z == b.z + 1;
}

template <int kPass>  // 0: count line starts per block; 1: write lines
__global__ __launch_bounds__(kLineThreads) void hkl_lines_kernel(const int4 *hkl, const float4 *g4,
                                                                 int n, int *lines, int capacity,
                                                                 int *work) {
  __shared__ int s_wave[kLineThreads / kWave];
  const int i = blockIdx.x * kLineThreads + threadIdx.x;
  const bool start = i < n && (i == 0 || !line_continues(hkl, i));
  const uint64_t m = __ballot(start);
  const int lane = lane_id(), wid = threadIdx.x / kWave;
  if (lane == 0) s_wave[wid] = __popcll(m);
  __syncthreads();
  if (kPass == 0) {
    if (threadIdx.x == 0) {
      int t = 0;
      for (int w = 0; w < kLineThreads / kWave; ++w) t += s_wave[w];
      work[blockIdx.x] = t;
    }
    return;
  }
  if (!start) return;
  int q = work[gridDim.x + blockIdx.x] + mask_rank(m);
  for (int w = 0; w < wid; ++w) q += s_wave[w];
  if (q >= capacity) return;
  float wmin = g4[i].w, wmax = wmin;
  int e = i + 1;
  while (e < n && line_continues(hkl, e)) {
    const float w = g4[e].w;
    wmin = fminf(wmin, w);
    wmax = fmaxf(wmax, w);
    ++e;
  }
  const int4 h = hkl[i];
  int4 *out = reinterpret_cast<int4 *>(lines) + 2 * static_cast<size_t>(q);
  out[0] = make_int4(i, e - i, h.x, h.y);
  out[1] = make_int4(h.z, __float_as_int(wmin), __float_as_int(wmax), 0);
}

// ---------------------------------------------------------------------------
// Structure factor
// ---------------------------------------------------------------------------

constexpr int kSfThreads = 256;

struct SfArgs {
  const int4 *hkl;
  int n;
  double rec[9];
  const uint4 *xyz_fx;
  const int *elem;
  const float *occ;
  const float *B;
  int n_atoms;
  const double *coeff;  // [n_elem][10]
  int n_elem;
  int scatter;
  float4 *g4;
  float *F2;
  float2 *F;
};

// One thread per reflection; atoms are staged through LDS in chunks.  The
// atomic scattering factor is recomputed only when the element changes along
// the atom list, so a host-side sort by element makes it once per element.
__global__ __launch_bounds__(kSfThreads) void structure_factor_kernel(SfArgs a) {
  __shared__ uint4 s_xyz[kSfThreads];
  __shared__ float s_occ[kSfThreads];
  __shared__ float s_B[kSfThreads];
  __shared__ int s_el[kSfThreads];

  const int tid = threadIdx.x;
  const int i = blockIdx.x * kSfThreads + tid;
  const bool valid = i < a.n;
  int4 hkl = valid ? a.hkl[i] : make_int4(0, 0, 0, 0);

  double gx = a.rec[0] * hkl.x + a.rec[1] * hkl.y + a.rec[2] * hkl.z;
  double gy = a.rec[3] * hkl.x + a.rec[4] * hkl.y + a.rec[5] * hkl.z;
  double gz = a.rec[6] * hkl.x + a.rec[7] * hkl.y + a.rec[8] * hkl.z;
  const double g2d = gx * gx + gy * gy + gz * gz;
  const float g2 = static_cast<float>(g2d);

  const uint32_t uh = static_cast<uint32_t>(hkl.x), uk = static_cast<uint32_t>(hkl.y),
                 ul = static_cast<uint32_t>(hkl.z);
  const float mq = -0.25f * g2;
  double Fr = 0.0, Fi = 0.0;  // FP64 accumulation: full rate on CDNA4
  int cur_el = -1;
  double f = 0.0;
  for (int base = 0; base < a.n_atoms; base += kSfThreads) {
    __syncthreads();
    const int j = base + tid;
    if (j < a.n_atoms) {
      s_xyz[tid] = a.xyz_fx[j];
      s_occ[tid] = a.occ[j];
      s_B[tid] = a.B[j];
      s_el[tid] = a.elem[j];
    }
    __syncthreads();
    const int m = min(kSfThreads, a.n_atoms - base);
    for (int jj = 0; jj < m; ++jj) {
      const int el = min(max(s_el[jj], 0), a.n_elem - 1);
      if (el != cur_el) {  // uniform across the block
        f = atomic_f(a.coeff + 10 * el, g2d, a.scatter);
        cur_el = el;
      }
      const uint4 x = s_xyz[jj];
      // exact phase in turns mod 1: 32-bit wrap-around of h x + k y + l z
      const uint32_t t = uh * x.x + uk * x.y + ul * x.z;
      const float turns = static_cast<float>(static_cast<int32_t>(t)) * 2.3283064365386963e-10f;
      float sn, cs;
      sincospif(2.0f * turns, &sn, &cs);
      const double amp = f * static_cast<double>(s_occ[jj] * expf(mq * s_B[jj]));
      Fr = fma(amp, static_cast<double>(cs), Fr);
      Fi = fma(amp, static_cast<double>(sn), Fi);
    }
  }
  if (valid) {
    a.g4[i] = make_float4(static_cast<float>(gx), static_cast<float>(gy), static_cast<float>(gz), g2);
    a.F2[i] = static_cast<float>(Fr * Fr + Fi * Fi);
    if (a.F) a.F[i] = make_float2(static_cast<float>(Fr), static_cast<float>(Fi));
  }
}

}  // namespace dsx

using namespace dsx;

extern "C" size_t dsx_hkl_workspace_ints(const int highest[3]) {
  const long long total = static_cast<long long>(2 * highest[0] + 1) * (2 * highest[1] + 1) *
                          (2 * highest[2] + 1);
  const long long nb = (total + kEnumPerBlock - 1) / kEnumPerBlock;
  return static_cast<size_t>(2 * nb);
}

static int hkl_box_launch(bool sphere, const int highest[3], const double metric[6], double bound,
                          int *hkl_d, int capacity, int *count_d, int *work_d, void *stream,
                          const char *name) {
  clear_error();
  if (!highest || !metric || !count_d || !work_d || (capacity > 0 && !hkl_d)) {
    set_error("%s: null argument", name);
    return DSX_EINVAL;
  }
  if (highest[0] < 0 || highest[1] < 0 || highest[2] < 0 || !(bound > 0.0)) {
    set_error("%s: bad box {%d,%d,%d} or bound %g", name, highest[0], highest[1], highest[2],
              bound);
    return DSX_EINVAL;
  }
  Box b;
  b.H = highest[0];
  b.K = highest[1];
  b.L = highest[2];
  b.nk = 2 * b.K + 1;
  b.nl = 2 * b.L + 1;
  b.total = static_cast<long long>(2 * b.H + 1) * b.nk * b.nl;
  if (b.total > (1ll << 31)) {
    set_error("%s: box of %lld entries exceeds 2^31", name, b.total);
    return DSX_EINVAL;
  }
  Metric m{metric[0], metric[1], metric[2], metric[3], metric[4], metric[5]};
  const int nb = static_cast<int>((b.total + kEnumPerBlock - 1) / kEnumPerBlock);
  hipStream_t s = as_stream(stream);
  if (sphere) {
    hipLaunchKernelGGL((hkl_enum_kernel<0, true>), dim3(nb), dim3(kEnumThreads), 0, s, b, m, bound,
                       work_d, hkl_d, capacity);
    hipLaunchKernelGGL(hkl_scan_kernel, dim3(1), dim3(1024), 0, s, work_d, nb, count_d);
    hipLaunchKernelGGL((hkl_enum_kernel<1, true>), dim3(nb), dim3(kEnumThreads), 0, s, b, m, bound,
                       work_d, hkl_d, capacity);
  } else {
    hipLaunchKernelGGL((hkl_enum_kernel<0, false>), dim3(nb), dim3(kEnumThreads), 0, s, b, m,
                       bound, work_d, hkl_d, capacity);
    hipLaunchKernelGGL(hkl_scan_kernel, dim3(1), dim3(1024), 0, s, work_d, nb, count_d);
    hipLaunchKernelGGL((hkl_enum_kernel<1, false>), dim3(nb), dim3(kEnumThreads), 0, s, b, m,
                       bound, work_d, hkl_d, capacity);
  }
  return check_launch(name);
}

extern "C" int dsx_hkl_enumerate(const int highest[3], const double metric[6], double d_min,
                                 int *hkl_d, int capacity, int *count_d, int *work_d,
                                 void *stream) {
  return hkl_box_launch(false, highest, metric, d_min, hkl_d, capacity, count_d, work_d, stream,
                        "dsx_hkl_enumerate");
}

extern "C" int dsx_hkl_points_in_sphere(const int highest[3], const double metric[6],
                                        double radius, int *hkl_d, int capacity, int *count_d,
                                        int *work_d, void *stream) {
  return hkl_box_launch(true, highest, metric, radius, hkl_d, capacity, count_d, work_d, stream,
                        "dsx_hkl_points_in_sphere");
}

// Atom staging for dsx_structure_factor from device-resident parameters:
// row i takes atom order[i]; fractional position -> uint32 turns
// round((f - floor(f)) * 2^32) mod 2^32 (rint: half to even), B -> float.
__global__ void stage_atoms_kernel(const double *xyz, const double *B, const long long *order, int n,
                                   uint4 *fx, float *Bs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long j = order[i];
  uint32_t u[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double f = xyz[3 * j + c];
    const double r = rint((f - floor(f)) * 4294967296.0);
    u[c] = static_cast<uint32_t>(static_cast<unsigned long long>(r) & 0xFFFFFFFFull);
  }
  fx[i] = make_uint4(u[0], u[1], u[2], 0u);
  Bs[i] = static_cast<float>(B[j]);
}

extern "C" int dsx_stage_atoms(const double *xyz_d, const double *B_d, const long long *order_d,
                               int n_atoms, uint32_t *xyz_fx_d, float *B_out_d, void *stream) {
  clear_error();
  if (n_atoms < 0) {
    set_error("dsx_stage_atoms: bad atom count %d", n_atoms);
    return DSX_EINVAL;
  }
  if (n_atoms == 0) return DSX_OK;
  if (!xyz_d || !B_d || !order_d || !xyz_fx_d || !B_out_d) {
    set_error("dsx_stage_atoms: null device pointer");
    return DSX_EINVAL;
  }
  hipLaunchKernelGGL(stage_atoms_kernel, dim3((n_atoms + 255) / 256), dim3(256), 0, as_stream(stream),
                     xyz_d, B_d, order_d, n_atoms, reinterpret_cast<uint4 *>(xyz_fx_d), B_out_d);
  return check_launch("dsx_stage_atoms");
}

extern "C" int dsx_structure_factor(const int *hkl_d, int n, const double rec[9],
                                    const uint32_t *xyz_fx_d, const int *elem_d,
                                    const float *occ_d, const float *B_d, int n_atoms,
                                    const double *coeff_d, int n_elem, int scatter, float *g4_d,
                                    float *F2_d, float *F_d, void *stream) {
  clear_error();
  if (n < 0 || n_atoms < 0 || n_elem < 0 || !rec) {
    set_error("dsx_structure_factor: bad sizes n=%d n_atoms=%d n_elem=%d", n, n_atoms, n_elem);
    return DSX_EINVAL;
  }
  if (scatter != DSX_SCATTER_NONE && scatter != DSX_SCATTER_LOBATO && scatter != DSX_SCATTER_XTABLES) {
    set_error("dsx_structure_factor: unknown scattering parameterisation %d", scatter);
    return DSX_EINVAL;
  }
  if (n == 0) return DSX_OK;
  if (n_atoms > 0 && n_elem == 0) {
    set_error("dsx_structure_factor: atoms given but no element coefficients");
    return DSX_EINVAL;
  }
  if (!hkl_d || !g4_d || !F2_d || (n_atoms > 0 && (!xyz_fx_d || !elem_d || !occ_d || !B_d)) ||
      (n_elem > 0 && !coeff_d)) {
    set_error("dsx_structure_factor: null device pointer");
    return DSX_EINVAL;
  }
  SfArgs a;
  a.hkl = reinterpret_cast<const int4 *>(hkl_d);
  a.n = n;
  for (int i = 0; i < 9; ++i) a.rec[i] = rec[i];
  a.xyz_fx = reinterpret_cast<const uint4 *>(xyz_fx_d);
  a.elem = elem_d;
  a.occ = occ_d;
  a.B = B_d;
  a.n_atoms = n_atoms;
  a.coeff = coeff_d;
  a.n_elem = n_elem;
  a.scatter = scatter;
  a.g4 = reinterpret_cast<float4 *>(g4_d);
  a.F2 = F2_d;
  a.F = reinterpret_cast<float2 *>(F_d);
  hipLaunchKernelGGL(structure_factor_kernel, dim3((n + kSfThreads - 1) / kSfThreads),
                     dim3(kSfThreads), 0, as_stream(stream), a);
  return check_launch("dsx_structure_factor");
}

extern "C" long long dsx_hkl_lines_workspace_ints(int n_refl) {
  if (n_refl < 0) return -1;
  return 2ll * ((n_refl + kLineThreads - 1) / kLineThreads);
}

extern "C" int dsx_hkl_lines(const int *hkl_d, const float *g4_d, int n_refl, int *lines_d,
                             int capacity, int *count_d, int *work_d, long long work_ints,
                             void *stream) {
  clear_error();
  if (n_refl < 0 || capacity < 0 || !count_d || (capacity > 0 && !lines_d) ||
      (n_refl > 0 && (!hkl_d || !g4_d || !work_d))) {
    set_error("dsx_hkl_lines: bad arguments (n_refl=%d capacity=%d) or null pointer", n_refl,
              capacity);
    return DSX_EINVAL;
  }
  const int nb = (n_refl + kLineThreads - 1) / kLineThreads;
  if (n_refl > 0 && work_ints < 2ll * nb) {
    set_error("dsx_hkl_lines: workspace of %lld ints < %lld", work_ints, 2ll * nb);
    return DSX_EINVAL;
  }
  hipStream_t s = as_stream(stream);
  if (n_refl == 0) {  // the scan of no blocks writes a zero count
    hipLaunchKernelGGL(hkl_scan_kernel, dim3(1), dim3(1024), 0, s, work_d, 0, count_d);
    return check_launch("dsx_hkl_lines");
  }
  const int4 *hkl = reinterpret_cast<const int4 *>(hkl_d);
  const float4 *g4 = reinterpret_cast<const float4 *>(g4_d);
  hipLaunchKernelGGL(hkl_lines_kernel<0>, dim3(nb), dim3(kLineThreads), 0, s, hkl, g4, n_refl,
                     lines_d, capacity, work_d);
  hipLaunchKernelGGL(hkl_scan_kernel, dim3(1), dim3(1024), 0, s, work_d, nb, count_d);
  hipLaunchKernelGGL(hkl_lines_kernel<1>, dim3(nb), dim3(kLineThreads), 0, s, hkl, g4, n_refl,
                     lines_d, capacity, work_d);
  return check_launch("dsx_hkl_lines");
}
