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
//
// With a line index (dsx_hkl_lines) the sweep visits only rows that can pass
// the pretest: a row passes only if its beam coordinate z lies in a slab
// [(wmin - thr) / 2k, (wmax + thr) / 2k] (w = |g|^2 over the row's line), and
// along a line z is linear in l, so each line contributes one interval of
// rows per orientation.  The intervals of 64 lines are flattened across the
// wave (prefix sum, per-lane binary search) in row order, so the pretest sees
// exactly the rows it would pass in a full sweep, in the same order, and the
// outputs are identical.
#include <algorithm>

#include "dsx_internal.h"
#include "dsx_physics.h"

namespace dsx {

// Orientations per wave: one 16-byte table load feeds kOT dot products
// (measured on the 40-atom library: 2.25 ms with 4, 2.82 ms with 2, 5.88 ms
// with 8, which spills).
constexpr int kOT = 4;
constexpr int kSlotBits = 2;      // bits of an orientation slot in a queue entry
static_assert((1 << kSlotBits) == kOT, "orientation slot bits");
// one wave per block: the waves share nothing, and single-wave blocks let the
// dispatcher refill a finished wave's slot at once (the library's spot kernel
// 1.46 -> 1.43 ms, the 64-phase batch's 18.5 -> 18.2 ms against four waves)
constexpr int kSimThreads = 64;
// The line sweep pays once lines average this many rows.  Measured with the
// intervals of a group formed for all orientations at once: the 40-atom
// library table (14.6 rows per line) 2.19 ms full sweep against 1.55 ms line
// sweep for 200,000 orientations; thresholds of 6 or 3 gain nothing more on
// the 64-phase ragged batch.
constexpr int kMinLineRows = DSX_LINE_SWEEP_MIN_ROWS;

__host__ __device__ __forceinline__ bool lines_pay(int n_refl, int n_lines) {
  return n_lines > 0 && static_cast<long long>(n_refl) >= static_cast<long long>(kMinLineRows) * n_lines;
}

struct SimArgs {
  double k, k2;
  double me;         // max_excitation_error
  double width;      // shape width
  double sigma;      // pi / width
  double fac;        // pi * minima_number / width (|width| for atanc)
  int model;
  int precession;
  double phi, Pz, Pt;  // precession angle (rad), k cos(phi), k sin(phi)
  double sin_phi;
  float min_I;
  float k_f;
  int n_items;                  // tiles of kOT orientations
  dsx_phase_desc single;        // plain launch: the one phase
  const dsx_phase_desc *phases; // ragged launch: device array of phases
  const int2 *items;            // ragged launch: {phase, first orientation} per tile
};

// Pretest bound: e = |g|^2 - 2 k z = s (2k - z_sphere - z), and |s| <= me
// (+ r sin(phi) with precession); the slack covers FP32 rounding of e.
__host__ __device__ __forceinline__ float pretest_threshold(double me, double k, double gmax,
                                                            double sin_phi) {
  const double me_eff = me + (sin_phi > 0.0 ? 1.1 * gmax * sin_phi : 0.0);
  return static_cast<float>(1.05 * me_eff * (2.0 * k + gmax) + 1e-3 +
                            1e-6 * (2.0 * k * gmax + gmax * gmax));
}

// Candidate queue per wave: (reflection row << kSlotBits) | orientation slot, in the
// order the sweep produced them (row-major, so per orientation ascending rows).
// Ring of at least one batch plus one worst-case step (64 + kOT * 64
// entries), a power of two so that positions wrap with a mask: candidates
// are spread thinly over the rows, so nearly every step queues a few for
// every orientation, and a modulo by a non-power of two there cost more
// than the pretest itself.
constexpr int kQueue = 512;
static_assert(kQueue >= 64 + kOT * kWave && (kQueue & (kQueue - 1)) == 0, "queue ring size");
__device__ __forceinline__ int qpos(int i) { return i & (kQueue - 1); }
constexpr int kWavesPerBlock = kSimThreads / kWave;

// Register budgets (waves per SIMD), measured: the full sweep at 5 (96 VGPRs)
// and the line sweep at 4 (128 VGPRs) are 5-6 % faster than the compiler's own
// choice (111 and 151-165 VGPRs) despite a few spilled registers; the
// averaged-precession models keep theirs.
constexpr int sim_min_waves(bool avg, bool lines) { return avg ? 1 : (lines ? 4 : 5); }

template <bool kRagged, bool kAvg, bool kLines>
__global__ __launch_bounds__(kSimThreads, sim_min_waves(kAvg, kLines)) void simulate_kernel(SimArgs a) {
  __shared__ int s_queue[kWavesPerBlock][kQueue];
  __shared__ double s_R[kWavesPerBlock][kOT][9];
  __shared__ double s_abc[kWavesPerBlock][kOT][4];  // beam . (a*, b*, c*), 1 / (beam . c*)
  // FP32 beam rows of the full-sweep pretest: in the line-sweep kernels, which
  // run the full sweep only for short-line phases, they wait in LDS instead
  // of holding 12 VGPRs through the line machinery
  __shared__ float s_Rf[kWavesPerBlock][kOT][3];
  // line sweep: the current group's intervals per orientation and lane
  constexpr int kLO = kLines ? kOT : 1, kLW = kLines ? kWave : 1;
  __shared__ int s_P[kWavesPerBlock][kLO][kLW];
  __shared__ int s_rb[kWavesPerBlock][kLO][kLW];
  __shared__ float s_eA[kWavesPerBlock][kLO][kLW];
  __shared__ float s_eB[kWavesPerBlock][kLO][kLW];
  __shared__ int s_tot[kWavesPerBlock][kLO];
  const int lane = lane_id();
  const int wib = threadIdx.x / kWave;
  int *queue = s_queue[wib];
  const int wave = blockIdx.x * kWavesPerBlock + wib;
  const int n_waves = gridDim.x * kWavesPerBlock;
  const float two_k = 2.0f * a.k_f;

  for (int tile = wave; tile < a.n_items; tile += n_waves) {
    const dsx_phase_desc *d;
    int o0;
    if (kRagged) {
      const int2 it = a.items[tile];
      d = a.phases + it.x;
      o0 = it.y;
    } else {
      d = &a.single;
      o0 = tile * kOT;
    }
    const int n_here = min(kOT, d->n_orient - o0);
    const int4 *hkl = reinterpret_cast<const int4 *>(d->hkl);
    const float4 *g4 = reinterpret_cast<const float4 *>(d->g4);
    const float *F2 = d->F2;
    const double *Rt = d->Rt;
    const double *rec = d->rec;
    const int n_refl = d->n_refl;
    const int cap = d->cap;
    int *idx_out = d->idx;
    float4 *spot_out = reinterpret_cast<float4 *>(d->spots);
    const float pre_thr = pretest_threshold(a.me, a.k, d->g_max, a.sin_phi);
    float R[kLines ? 1 : kOT][3];  // third rows (beam direction) for the FP32 pretest
    int cnt[kOT];
    float lmax[kOT], lmin[kOT];
    for (int e = lane; e < kOT * 9; e += kWave) {
      const int o = e / 9, j = e - o * 9;
      s_R[wib][o][j] = o < n_here ? Rt[(size_t)(o0 + o) * 9 + j] : 0.0;
    }
#pragma unroll
    for (int o = 0; o < kOT; ++o) {
      const bool ok = o < n_here;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const float v = ok ? static_cast<float>(Rt[(size_t)(o0 + o) * 9 + 6 + j]) : 0.0f;
        if (kLines)
          s_Rf[wib][o][j] = v;
        else
          R[o][j] = v;
      }
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
        const int e = queue[qpos(head + lane)];
        row = e >> kSlotBits;
        slot = e & (kOT - 1);
        const int4 h = hkl[row];
        const double hx = h.x, hy = h.y, hz = h.z;
        const double gx = rec[0] * hx + rec[1] * hy + rec[2] * hz;
        const double gy = rec[3] * hx + rec[4] * hy + rec[5] * hz;
        const double gz = rec[6] * hx + rec[7] * hy + rec[8] * hz;
        const double *Rd = s_R[wib][slot];
        x = Rd[0] * gx + Rd[1] * gy + Rd[2] * gz;
        y = Rd[3] * gx + Rd[4] * gy + Rd[5] * gz;
        z = Rd[6] * gx + Rd[7] * gy + Rd[8] * gz;
        double s, r;
        in = intersects(x, y, z, a, s, r);
        if (in) I = static_cast<float>(shape_factor<kAvg>(s, r, a) * static_cast<double>(F2[row]));
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
          if (pos < cap) {
            const size_t off = (size_t)(o0 + o) * cap + pos;
            idx_out[off] = row;
            spot_out[off] = make_float4(static_cast<float>(x), static_cast<float>(y),
                                       static_cast<float>(z), I);
          }
        }
        cnt[o] += __popcll(m);
      }
      head = qpos(head + nb);
      qn -= nb;
    };

    // Queue the rows of this lane that pass the pretest of orientation o
    // (beam direction r0, r1, r2 in FP32).
    auto pretest = [&](bool v, int row, const float4 &g, float r0, float r1, float r2, int o) {
      const float zf = fmaf(r0, g.x, fmaf(r1, g.y, r2 * g.z));
      const bool pre = v && fabsf(fmaf(-two_k, zf, g.w)) < pre_thr;
      const uint64_t pm = __ballot(pre);
      if (pm == 0) return;
      if (pre) queue[qpos(head + qn + mask_rank(pm))] = (row << kSlotBits) | o;
      qn += __popcll(pm);
    };

    // Line sweep state: the current group of 64 lines (lb), orientation (lo)
    // and window (win).  When a group loads, the row interval of every line
    // (one per lane) is formed for all kOT orientations at once -- four
    // independent chains of FP64 arithmetic and prefix sums instead of one
    // chain per orientation -- and parked in LDS (per orientation and lane:
    // first candidate P, first row rb, pretest coefficients eA, eB; the
    // orientation's candidate total in s_tot), so the line data is dead while
    // the windows and drains run.
    const bool use_lines = kLines && d->lines != nullptr && lines_pay(n_refl, d->n_lines);
    const int4 *lines = reinterpret_cast<const int4 *>(d->lines);
    const int n_lines = use_lines ? d->n_lines : 0;
    int lb = -kWave, lo = kOT - 1, win = 0, total = 0;
    // along a line (hkl = (h, k, l0 - t)) the pretest value is a quadratic in
    // t: e(t) = |g0 - t c*|^2 - 2k (zt - t ga) = A + B t + c2 t^2, so a window
    // forms it from its line's coefficients (eA, eB at the interval's first
    // row) instead of loading the row's g4; w0 = |g0|^2, cd = g0 . c* (FP32:
    // their rounding, ~1e-6 of e, is far inside the pretest's slack)
    const double c2 = rec[2] * rec[2] + rec[5] * rec[5] + rec[8] * rec[8];
    const float c2f = static_cast<float>(c2);
    // pretest pass => |w - 2k zf| < thr (1 + 2^-24); zf (FP32 R and g, two
    // fmaf) is within 1e-6 |g| of the FP64 beam coordinate, ten times inside
    // the margin zm = 1e-5 (1 + max |g| of the line) below
    const double two_kd = static_cast<double>(two_k);
    const double thr = static_cast<double>(pre_thr) * (1.0 + 1e-6);
    if (use_lines) {
      // beam direction (third row of R) against the reciprocal axes, FP64
      if (lane < kOT * 4) {
        const int o = lane / 4, c = lane - 4 * o, cc = c < 3 ? c : 2;
        double v = 0.0;
        if (o < n_here) {
          const double *b = Rt + (size_t)(o0 + o) * 9 + 6;
          v = b[0] * rec[cc] + b[1] * rec[3 + cc] + b[2] * rec[6 + cc];
          // 1 / (beam . c*); 0 marks a line direction perpendicular to the beam
          if (c == 3) v = fabs(v) < 1e-200 ? 0.0 : 1.0 / v;
        }
        s_abc[wib][o][c] = v;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
    }
    // Next group of lines: false once the lines are exhausted.
    auto load_group = [&]() -> bool {
      lb += kWave;
      if (lb >= n_lines) return false;
      const int j = lb + lane;
      int first = 0, nrow = 0, lh = 0, lk = 0, l0 = 0;
      double zlo = 0.0, zhi = -1.0;
      float w0 = 0.0f, cd = 0.0f;
      if (j < n_lines) {
        const int4 e0 = lines[2 * (size_t)j], e1 = lines[2 * (size_t)j + 1];
        first = e0.x;
        nrow = e0.y;
        lh = e0.z;
        lk = e0.w;
        l0 = e1.x;
        const double wmin = __int_as_float(e1.y), wmax = __int_as_float(e1.z);
        const double zm = 1e-5 * (1.0 + sqrt(wmax));
        zlo = (wmin - thr) / two_kd - zm;
        zhi = (wmax + thr) / two_kd + zm;
        const double hx = lh, hy = lk, hz = l0;
        const double gx = rec[0] * hx + rec[1] * hy + rec[2] * hz;
        const double gy = rec[3] * hx + rec[4] * hy + rec[5] * hz;
        const double gz = rec[6] * hx + rec[7] * hy + rec[8] * hz;
        w0 = static_cast<float>(gx * gx + gy * gy + gz * gz);
        cd = static_cast<float>(gx * rec[2] + gy * rec[5] + gz * rec[8]);
      }
      int t0[kOT], len[kOT], incl[kOT];
#pragma unroll
      for (int o = 0; o < kOT; ++o) {
        // z of row first + t is zt - t ga (l = l0 - t): the rows with z in [zlo, zhi]
        const double al = s_abc[wib][o][0], be = s_abc[wib][o][1], ga = s_abc[wib][o][2];
        const double inv = s_abc[wib][o][3];
        const double zt = static_cast<double>(lh) * al + static_cast<double>(lk) * be +
                          static_cast<double>(l0) * ga;
        t0[o] = 0;
        len[o] = 0;
        if (o >= n_here || nrow <= 0) {
          // no rows (an orientation past the tile's end, or no line)
        } else if (inv == 0.0) {  // line perpendicular to the beam: all rows or none
          if (zt >= zlo && zt <= zhi) len[o] = nrow;
        } else {
          const double q1 = (zt - zhi) * inv, q2 = (zt - zlo) * inv;
          const double tlo = fmax(ceil(fmin(q1, q2)), 0.0);
          const double thi = fmin(floor(fmax(q1, q2)), static_cast<double>(nrow - 1));
          if (thi >= tlo) {
            t0[o] = static_cast<int>(tlo);
            len[o] = static_cast<int>(thi - tlo) + 1;
          }
        }
        const double tt = static_cast<double>(t0[o]);
        const double eb = two_kd * ga - 2.0 * static_cast<double>(cd);
        s_eA[wib][o][lane] =
            static_cast<float>((static_cast<double>(w0) - two_kd * zt) + tt * (eb + c2 * tt));
        s_eB[wib][o][lane] = static_cast<float>(eb + 2.0 * c2 * tt);
        s_rb[wib][o][lane] = first + t0[o];
        incl[o] = len[o];  // inclusive prefix over the lanes (lines in row order)
      }
#pragma unroll
      for (int off = 1; off < kWave; off <<= 1) {
#pragma unroll
        for (int o = 0; o < kOT; ++o) {
          const int v = __shfl_up(incl[o], off);
          if (lane >= off) incl[o] += v;
        }
      }
#pragma unroll
      for (int o = 0; o < kOT; ++o) {
        s_P[wib][o][lane] = incl[o] - len[o];  // first candidate of this lane's line
        if (lane == kWave - 1) s_tot[wib][o] = incl[o];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
      return true;
    };

    // One step per iteration: a window of 64 line candidates of one
    // orientation, or 64 table rows for every orientation; full batches are
    // drained as they fill, everything after the last step.
    for (int base = 0;; base += kWave) {
      bool done;
      if (use_lines) {
        done = false;
        win += kWave;
        while (win >= total) {
          if (++lo >= n_here) {  // next group of lines
            lo = 0;
            if (!load_group()) {
              done = true;
              break;
            }
          }
          total = s_tot[wib][lo];
          win = 0;
        }
        if (!done) {
          const int c = win + lane;
          // the last lane whose line starts at or before candidate c holds it
          const int *Pw = s_P[wib][lo];
          int src = 0;
#pragma unroll
          for (int step = kWave / 2; step > 0; step >>= 1)
            if (Pw[src + step] <= c) src += step;
          const int u = c - Pw[src];
          const int row = s_rb[wib][lo][src] + u;
          const bool v = c < total && static_cast<unsigned>(row) < static_cast<unsigned>(n_refl);
          const float uf = static_cast<float>(u);
          const float e = fmaf(fmaf(c2f, uf, s_eB[wib][lo][src]), uf, s_eA[wib][lo][src]);
          const bool pre = v && fabsf(e) < pre_thr;
          const uint64_t pm = __ballot(pre);
          if (pm != 0) {
            if (pre) queue[qpos(head + qn + mask_rank(pm))] = (row << kSlotBits) | lo;
            qn += __popcll(pm);
          }
        }
      } else {
        done = base >= n_refl;
        if (!done) {
          const int i = base + lane;
          const bool v = i < n_refl;
          const float4 g = v ? g4[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
          for (int o = 0; o < kOT; ++o) {
            if (o >= n_here) break;  // uniform
            if (kLines)
              pretest(v, i, g, s_Rf[wib][o][0], s_Rf[wib][o][1], s_Rf[wib][o][2], o);
            else
              pretest(v, i, g, R[o][0], R[o][1], R[o][2], o);
          }
        }
      }
      const int need = done ? 1 : kWave;
      if (qn >= need) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
        while (qn >= need) drain(min(qn, kWave));
      }
      if (done) break;
    }

#pragma unroll
    for (int o = 0; o < kOT; ++o) {
      const int orient = o0 + o;
      if (o >= n_here) break;
      const float mx = wave_max(lmax[o]);
      const float mn = wave_min(lmin[o]);
      const int n = cnt[o];
      const int nw = min(n, cap);
      int final_count = nw;
      const float thr = mx * a.min_I;
      // mx == 0 with n > 0: every intersecting reflection is kinematically
      // forbidden (|F| exactly 0); the reference keeps them all (its FP64
      // round-off passes the relative cut), so they are kept with I = 0.
      if (n > 0 && mx > 0.0f && !(mn > thr)) {
        __threadfence_block();
        const size_t off = (size_t)orient * cap;
        int w = 0;
        for (int b = 0; b < nw; b += kWave) {
          const int j = b + lane;
          float4 sp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
          int id = 0;
          bool keep = false;
          if (j < nw) {
            sp = spot_out[off + j];
            id = idx_out[off + j];
            keep = sp.w > thr;
          }
          const uint64_t km = __ballot(keep);
          if (keep) {
            const int pos = w + mask_rank(km);
            spot_out[off + pos] = sp;
            idx_out[off + pos] = id;
          }
          w += __popcll(km);
        }
        final_count = w;
      }
      if (lane == 0) {
        d->count[orient] = final_count;
        d->ncand[orient] = n;
        d->maxI[orient] = mx;
      }
    }
    // the next tile rewrites s_R and the queue
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
  }
}

static int check_physics(const dsx_sim_params *p, const char *name) {
  const bool averaged = p->shape_model >= DSX_SF_PRECESSION_AVG + DSX_SF_LINEAR &&
                        p->shape_model <= DSX_SF_PRECESSION_AVG + DSX_SF_BINARY;
  if (!averaged && (p->shape_model < DSX_SF_LINEAR || p->shape_model > DSX_SF_LORENTZIAN_PRECESSION)) {
    set_error("%s: unknown shape model %d", name, p->shape_model);
    return DSX_EINVAL;
  }
  if (averaged && !(p->precession_angle > 0.0)) {
    set_error("%s: the full precession average needs a positive precession angle", name);
    return DSX_EINVAL;
  }
  if (!(p->k > 0.0) || !(p->max_excitation_error > 0.0) || p->shape_width == 0.0 ||
      !(p->precession_angle >= 0.0) || !(p->g_max >= 0.0)) {
    set_error("%s: bad physical parameters", name);
    return DSX_EINVAL;
  }
  if (p->shape_model == DSX_SF_LORENTZIAN_PRECESSION && !(p->precession_angle > 0.0)) {
    set_error("%s: the precession shape factor needs a positive precession angle", name);
    return DSX_EINVAL;
  }
  return DSX_OK;
}

static void fill_args(SimArgs &a, const dsx_sim_params *p) {
  fill_physics(a, p);
  a.k_f = static_cast<float>(p->k);
  a.sin_phi = p->precession_angle > 0.0 ? sin(p->precession_angle) : 0.0;
  a.min_I = static_cast<float>(p->minimum_intensity);
  a.phases = nullptr;
  a.items = nullptr;
}

// Blocks of the spot kernel's launch: each wave takes every n_waves-th tile.
// Far more waves than stay resident (default 128 one-wave blocks per CU, one
// or two tiles per wave on the library) let the dispatcher hand the slots of
// waves that finish early to the next ones, which balances tiles of unequal
// cost; a launch of only the resident waves leaves each a fixed share.  The
// library's 50,000 tiles, in waves per CU (four-wave blocks): 1.60 ms with 16,
// 1.55 with 32, 1.52 with 48, 1.50 with 64, 1.45-1.46 with 128 or 256.
static int launch_blocks(int n_items, int blocks_per_cu) {
  const int waves_per_block = kSimThreads / kWave;
  const int blocks = (n_items + waves_per_block - 1) / waves_per_block;
  return std::max(1, std::min(blocks, cu_count() * (blocks_per_cu > 0 ? blocks_per_cu : 128)));
}

}  // namespace dsx

using namespace dsx;

extern "C" int dsx_simulate(const int *hkl_d, const float *g4_d, const float *F2_d, int n_refl,
                            const int *lines_d, int n_lines, const double *Rt_d, int n_orient,
                            const dsx_sim_params *p, int cap, int *idx_d, float *spot_d,
                            int *count_d, int *ncand_d, float *maxI_d, void *stream) {
  clear_error();
  if (!p || n_refl < 0 || n_orient < 0 || cap < 1 || n_lines < 0 || (n_lines > 0 && !lines_d)) {
    set_error("dsx_simulate: bad arguments (n_refl=%d n_orient=%d cap=%d)", n_refl, n_orient, cap);
    return DSX_EINVAL;
  }
  if (n_orient == 0) return DSX_OK;
  if (!Rt_d || !idx_d || !spot_d || !count_d || !ncand_d || !maxI_d ||
      (n_refl > 0 && (!hkl_d || !g4_d || !F2_d))) {
    set_error("dsx_simulate: null device pointer");
    return DSX_EINVAL;
  }
  const int st = check_physics(p, "dsx_simulate");
  if (st != DSX_OK) return st;
  if ((size_t)n_orient * (size_t)cap > (size_t)1 << 40) {
    set_error("dsx_simulate: n_orient*cap too large");
    return DSX_EINVAL;
  }
  SimArgs a;
  fill_args(a, p);
  dsx_phase_desc &d = a.single;
  d.hkl = hkl_d;
  d.g4 = g4_d;
  d.F2 = F2_d;
  d.n_refl = n_refl;
  d.cap = cap;
  for (int i = 0; i < 9; ++i) d.rec[i] = p->rec[i];
  d.g_max = p->g_max;
  d.Rt = Rt_d;
  d.n_orient = n_orient;
  d.n_lines = lines_d ? n_lines : 0;
  d.idx = idx_d;
  d.spots = spot_d;
  d.count = count_d;
  d.ncand = ncand_d;
  d.maxI = maxI_d;
  d.lines = lines_d;
  a.n_items = (n_orient + kOT - 1) / kOT;
  const dim3 grid(launch_blocks(a.n_items, 0));
  hipStream_t s = as_stream(stream);
  // the averaged-precession model keeps the full sweep: its FP64 evaluation
  // dominates, and the line state would cost it occupancy
  const bool avg = p->shape_model >= DSX_SF_PRECESSION_AVG;
  if (lines_d && !avg && lines_pay(n_refl, n_lines))
    hipLaunchKernelGGL((simulate_kernel<false, false, true>), grid, dim3(kSimThreads), 0, s, a);
  else if (avg)
    hipLaunchKernelGGL((simulate_kernel<false, true, false>), grid, dim3(kSimThreads), 0, s, a);
  else
    hipLaunchKernelGGL((simulate_kernel<false, false, false>), grid, dim3(kSimThreads), 0, s, a);
  return check_launch("dsx_simulate");
}

extern "C" int dsx_simulate_tile_size(void) { return kOT; }

extern "C" int dsx_simulate_ragged(const dsx_phase_desc *phases_d, int n_phases, const int *items_d,
                                   int n_items, const dsx_sim_params *p, int blocks_per_cu,
                                   void *stream) {
  clear_error();
  if (!p || n_phases < 0 || n_items < 0 || blocks_per_cu < 0) {
    set_error("dsx_simulate_ragged: bad arguments (n_phases=%d n_items=%d)", n_phases, n_items);
    return DSX_EINVAL;
  }
  if (n_items == 0) return DSX_OK;
  if (!phases_d || !items_d || n_phases == 0) {
    set_error("dsx_simulate_ragged: null device pointer or no phases");
    return DSX_EINVAL;
  }
  const int st = check_physics(p, "dsx_simulate_ragged");
  if (st != DSX_OK) return st;
  SimArgs a;
  fill_args(a, p);
  a.phases = phases_d;
  a.items = reinterpret_cast<const int2 *>(items_d);
  a.n_items = n_items;
  // each phase uses its line index when it has one (lines != NULL) and its
  // lines are long enough, except under the averaged-precession model
  if (p->shape_model >= DSX_SF_PRECESSION_AVG)
    hipLaunchKernelGGL((simulate_kernel<true, true, false>), dim3(launch_blocks(n_items, blocks_per_cu)),
                       dim3(kSimThreads), 0, as_stream(stream), a);
  else
    hipLaunchKernelGGL((simulate_kernel<true, false, true>), dim3(launch_blocks(n_items, blocks_per_cu)),
                       dim3(kSimThreads), 0, as_stream(stream), a);
  return check_launch("dsx_simulate_ragged");
}
