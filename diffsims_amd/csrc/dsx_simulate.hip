// (b) Persistent per-orientation kernel: rotate g, excitation error against the
// Ewald sphere, shape factor, intensity, relative cut, ordered compaction.
//
// Follows calculate_diffraction2d's orientation loop and
// get_intersecting_reflections (diffsims/generators/simulation_generator.py:
// 211-241, :319-412) and the models of diffsims/utils/shape_factor_models.py.
//
// Mapping: one wave64 owns a tile of kOT orientations; its lanes sweep the
// reflection table 64 rows at a time.  A cheap rotation-invariant FP32 pretest
// (|g|^2 - 2 k z, one dot product) rejects almost all rows; rows that pass are
// re-evaluated in FP64 from their integer hkl (g = rec hkl, g_lab = R g) with
// the reference's arithmetic made cancellation-free, and survivors are
// appended in reflection order with a wave ballot + mbcnt prefix.  The relative intensity cut needs the maximum over the whole
// intersecting set, so survivors are written first and, only if some fall
// below the cut, re-compacted in place by the same wave.
#include <algorithm>

#include "dsx_internal.h"
#include "dsx_physics.h"

namespace dsx {

constexpr int kOT = 4;            // orientations per wave
constexpr int kSimThreads = 256;  // 4 waves per block

struct SimArgs {
  const int4 *hkl;
  const float4 *g4;  // {gx, gy, gz, |g|^2}
  const float *F2;
  int n_refl;
  const double *Rt;  // [n_orient][9], g_lab = Rt g
  int n_orient;
  double rec[9];
  double k, k2;
  double me;         // max_excitation_error
  double width;      // shape width
  double sigma;      // pi / width
  double fac;        // pi * minima_number / width (|width| for atanc)
  int model;
  int precession;
  double phi, Pz, Pt;  // precession angle (rad), k cos(phi), k sin(phi)
  float min_I;
  float k_f;
  float pre_thr;
  int cap;
  int *idx;
  float4 *spots;
  int *count;
  int *ncand;
  float *maxI;
};

// Candidate queue per wave: (reflection row << 2) | orientation slot, in the
// order the sweep produced them (row-major, so per orientation ascending rows).
constexpr int kQueue = 64 + kOT * kWave;  // one batch plus one worst-case chunk
constexpr int kWavesPerBlock = kSimThreads / kWave;

__global__ __launch_bounds__(kSimThreads) void simulate_kernel(SimArgs a) {
  __shared__ int s_queue[kWavesPerBlock][kQueue];
  __shared__ double s_R[kWavesPerBlock][kOT][9];
  const int lane = lane_id();
  const int wib = threadIdx.x / kWave;
  int *queue = s_queue[wib];
  const int wave = blockIdx.x * kWavesPerBlock + wib;
  const int n_waves = gridDim.x * kWavesPerBlock;
  const int n_tiles = (a.n_orient + kOT - 1) / kOT;
  const float two_k = 2.0f * a.k_f;

  for (int tile = wave; tile < n_tiles; tile += n_waves) {
    const int o0 = tile * kOT;
    const int n_here = min(kOT, a.n_orient - o0);
    float R[kOT][3];  // third rows (beam direction) for the FP32 pretest
    int cnt[kOT];
    float lmax[kOT], lmin[kOT];
    if (lane < kOT * 9) {
      const int o = lane / 9, j = lane - o * 9;
      s_R[wib][o][j] = o < n_here ? a.Rt[(size_t)(o0 + o) * 9 + j] : 0.0;
    }
#pragma unroll
    for (int o = 0; o < kOT; ++o) {
      const bool ok = o < n_here;
#pragma unroll
      for (int j = 0; j < 3; ++j) R[o][j] = ok ? static_cast<float>(a.Rt[(size_t)(o0 + o) * 9 + 6 + j]) : 0.0f;
      cnt[o] = 0;
      lmax[o] = 0.0f;
      lmin[o] = __builtin_huge_valf();
    }
    int head = 0, qn = 0;  // ring buffer state (uniform)

    // Evaluate one batch of up to 64 queued candidates in FP64, all lanes busy.
    auto drain = [&](int nb) {
      bool in = false;
      int row = 0, slot = 0;
      double x = 0.0, y = 0.0, z = 0.0;
      float I = 0.0f;
      if (lane < nb) {
        const int e = queue[(head + lane) % kQueue];
        row = e >> 2;
        slot = e & 3;
        const int4 h = a.hkl[row];
        const double hx = h.x, hy = h.y, hz = h.z;
        const double gx = a.rec[0] * hx + a.rec[1] * hy + a.rec[2] * hz;
        const double gy = a.rec[3] * hx + a.rec[4] * hy + a.rec[5] * hz;
        const double gz = a.rec[6] * hx + a.rec[7] * hy + a.rec[8] * hz;
        const double *Rd = s_R[wib][slot];
        x = Rd[0] * gx + Rd[1] * gy + Rd[2] * gz;
        y = Rd[3] * gx + Rd[4] * gy + Rd[5] * gz;
        z = Rd[6] * gx + Rd[7] * gy + Rd[8] * gz;
        double s, r;
        in = intersects(x, y, z, a, s, r);
        if (in) I = static_cast<float>(shape_factor(s, r, a) * static_cast<double>(a.F2[row]));
      }
#pragma unroll
      for (int o = 0; o < kOT; ++o) {
        const bool mine = in && slot == o;
        const uint64_t m = __ballot(mine);
        if (m == 0) continue;
        if (mine) {
          const int pos = cnt[o] + mask_rank(m);
          lmax[o] = fmaxf(lmax[o], I);
          lmin[o] = fminf(lmin[o], I);
          if (pos < a.cap) {
            const size_t off = (size_t)(o0 + o) * a.cap + pos;
            a.idx[off] = row;
            a.spots[off] = make_float4(static_cast<float>(x), static_cast<float>(y),
                                       static_cast<float>(z), I);
          }
        }
        cnt[o] += __popcll(m);
      }
      head = (head + nb) % kQueue;
      qn -= nb;
    };

    for (int base = 0;; base += kWave) {
      const bool more = base < a.n_refl;
      if (more) {
        const int i = base + lane;
        const bool v = i < a.n_refl;
        const float4 g = v ? a.g4[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int o = 0; o < kOT; ++o) {
          if (o >= n_here) break;  // uniform
          const float zf = fmaf(R[o][0], g.x, fmaf(R[o][1], g.y, R[o][2] * g.z));
          const bool pre = v && fabsf(fmaf(-two_k, zf, g.w)) < a.pre_thr;
          const uint64_t pm = __ballot(pre);
          if (pm == 0) continue;
          if (pre) queue[(head + qn + mask_rank(pm)) % kQueue] = (i << 2) | o;
          qn += __popcll(pm);
        }
      }
      // drain full batches during the sweep, everything at its end
      const int need = more ? kWave : 1;
      if (qn >= need) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
        while (qn >= need) drain(min(qn, kWave));
      }
      if (!more) break;
    }

#pragma unroll
    for (int o = 0; o < kOT; ++o) {
      const int orient = o0 + o;
      if (o >= n_here) break;
      const float mx = wave_max(lmax[o]);
      const float mn = wave_min(lmin[o]);
      const int n = cnt[o];
      const int nw = min(n, a.cap);
      int final_count = nw;
      const float thr = mx * a.min_I;
      // mx == 0 with n > 0: every intersecting reflection is kinematically
      // forbidden (|F| exactly 0); the reference keeps them all (its FP64
      // round-off passes the relative cut), so they are kept with I = 0.
      if (n > 0 && mx > 0.0f && !(mn > thr)) {
        __threadfence_block();
        const size_t off = (size_t)orient * a.cap;
        int w = 0;
        for (int b = 0; b < nw; b += kWave) {
          const int j = b + lane;
          float4 sp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
          int id = 0;
          bool keep = false;
          if (j < nw) {
            sp = a.spots[off + j];
            id = a.idx[off + j];
            keep = sp.w > thr;
          }
          const uint64_t km = __ballot(keep);
          if (keep) {
            const int pos = w + mask_rank(km);
            a.spots[off + pos] = sp;
            a.idx[off + pos] = id;
          }
          w += __popcll(km);
        }
        final_count = w;
      }
      if (lane == 0) {
        a.count[orient] = final_count;
        a.ncand[orient] = n;
        a.maxI[orient] = mx;
      }
    }
    // the next tile rewrites s_R and the queue
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
  }
}

}  // namespace dsx

using namespace dsx;

extern "C" int dsx_simulate(const int *hkl_d, const float *g4_d, const float *F2_d, int n_refl,
                            const double *Rt_d, int n_orient, const dsx_sim_params *p, int cap,
                            int *idx_d, float *spot_d, int *count_d, int *ncand_d, float *maxI_d,
                            void *stream) {
  clear_error();
  if (!p || n_refl < 0 || n_orient < 0 || cap < 1) {
    set_error("dsx_simulate: bad arguments (n_refl=%d n_orient=%d cap=%d)", n_refl, n_orient, cap);
    return DSX_EINVAL;
  }
  if (n_orient == 0) return DSX_OK;
  if (!Rt_d || !idx_d || !spot_d || !count_d || !ncand_d || !maxI_d ||
      (n_refl > 0 && (!hkl_d || !g4_d || !F2_d))) {
    set_error("dsx_simulate: null device pointer");
    return DSX_EINVAL;
  }
  if (p->shape_model < DSX_SF_LINEAR || p->shape_model > DSX_SF_LORENTZIAN_PRECESSION) {
    set_error("dsx_simulate: unknown shape model %d", p->shape_model);
    return DSX_EINVAL;
  }
  if (!(p->k > 0.0) || !(p->max_excitation_error > 0.0) || p->shape_width == 0.0 ||
      !(p->precession_angle >= 0.0) || !(p->g_max >= 0.0)) {
    set_error("dsx_simulate: bad physical parameters");
    return DSX_EINVAL;
  }
  if (p->shape_model == DSX_SF_LORENTZIAN_PRECESSION && !(p->precession_angle > 0.0)) {
    set_error("dsx_simulate: the precession shape factor needs a positive precession angle");
    return DSX_EINVAL;
  }
  if ((size_t)n_orient * (size_t)cap > (size_t)1 << 40) {
    set_error("dsx_simulate: n_orient*cap too large");
    return DSX_EINVAL;
  }
  SimArgs a;
  a.hkl = reinterpret_cast<const int4 *>(hkl_d);
  a.g4 = reinterpret_cast<const float4 *>(g4_d);
  a.F2 = F2_d;
  a.n_refl = n_refl;
  a.Rt = Rt_d;
  a.n_orient = n_orient;
  for (int i = 0; i < 9; ++i) a.rec[i] = p->rec[i];
  fill_physics(a, p);
  const double k = p->k;
  a.k_f = static_cast<float>(k);
  const double phi = p->precession_angle;
  a.min_I = static_cast<float>(p->minimum_intensity);
  // Pretest bound: e = |g|^2 - 2 k z = s (2k - z_sphere - z), and |s| <= me
  // (+ r sin(phi) with precession); the slack covers FP32 rounding of e.
  const double gmax = p->g_max;
  const double me_eff = p->max_excitation_error + (phi > 0.0 ? 1.1 * gmax * sin(phi) : 0.0);
  a.pre_thr = static_cast<float>(1.05 * me_eff * (2.0 * k + gmax) + 1e-3 +
                                 1e-6 * (2.0 * k * gmax + gmax * gmax));
  a.cap = cap;
  a.idx = idx_d;
  a.spots = reinterpret_cast<float4 *>(spot_d);
  a.count = count_d;
  a.ncand = ncand_d;
  a.maxI = maxI_d;
  const int n_tiles = (n_orient + kOT - 1) / kOT;
  const int waves_per_block = kSimThreads / kWave;
  int blocks = (n_tiles + waves_per_block - 1) / waves_per_block;
  blocks = std::min(blocks, cu_count() * 8);
  hipLaunchKernelGGL(simulate_kernel, dim3(blocks), dim3(kSimThreads), 0, as_stream(stream), a);
  return check_launch("dsx_simulate");
}
