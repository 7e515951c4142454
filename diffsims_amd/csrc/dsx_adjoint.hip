// Adjoint of the pattern path: gradients of a loss on the rendered sub-pixel
// patterns with respect to the orientation matrices, the reciprocal basis
// (hence the lattice parameters), the fractional atom positions and the
// Debye-Waller factors.  diffsims has no gradients; this is the reverse mode
// of dsx_simulate + dsx_render_subpixel + dsx_structure_factor.
//
// render_adjoint_kernel (one block per pattern, persistent over patterns)
//   recomputes the pattern as render_kernel<subpixel> does (same spots and
//   clip boxes; 32-bit fixed-point LDS strips), so it needs no stored pattern:
//     pass A  pattern maximum m and its first pixel p* (normalize): exact
//             evaluation of the few 16x16 tiles whose bound can reach the
//             value at the brightest spot (a full pass when too many remain);
//     pass B  per strip: P = Q/m, then the upstream gradient G = dL/dP
//             (2 (P - T) for the L2 loss, or given), kept in LDS as G/m;
//             each spot gathers A = sum G/m K, B = sum G/m K (x - X) over its
//             box (the transpose of the splat: no atomics on pixels);
//     then    the max term of the normalisation, -(sum_q G P)/m at p*, and
//             the per-spot sums go to a workspace.
// spot_chain_kernel (one wave per pattern)
//   per spot the chain through pixel position, intensity I = |F|^2 f(s, r)
//   and the excitation error (recomputed in FP64 from the integer hkl, the
//   orientation and the reciprocal basis): dL/dRt of the pattern, dL/d|F|^2
//   per reflection (atomics) and dL/drec through g (block partials, summed by
//   sum_partials_kernel in a fixed order).
// render_adjoint_large_kernel: the same for spot lists longer than the LDS
//   spot table: their bins live in a per-block workspace and the spots pass
//   through LDS in chunks; each spot's gathered sums are added into its
//   position of the pattern's gsum row (float atomics over the strips it
//   spans), which spot_chain_kernel then reads like the small kernel's.
// structure_factor_adjoint_kernel (one thread per reflection)
//   recomputes F(hkl) and reduces dL/d|F|^2 into dL/dxyz and dL/dB per atom
//   and into dL/d(rec) through |g|.
#include <algorithm>
#include <mutex>

#include "dsx_internal.h"
#include "dsx_physics.h"
#include "dsx_splat.h"

namespace dsx {

constexpr int kAdjThreads = 512;
constexpr int kAdjWaves = kAdjThreads / kWave;
constexpr int kAdjMaxSpots = 1024;   // longest spot list of the two-blocks-per-CU kernel
constexpr int kAdjMaxStrips = 256;
// The large-pattern kernel splits its strip bins by clip-box width (<= 4,
// <= 8, wider columns), so the (spot, column) items of a class run with that
// class's column count instead of the pattern's widest box.  (Small patterns
// hold few spots per strip; one pass per strip fills the waves better.)
constexpr int kWidthClasses = 3;
constexpr int kAdjBinCap = 4096;
// large-pattern kernel: one 64 KB strip buffer (the splat's fixed-point words
// in pass 1, G in pass 2) and as many LDS spot slots as the rest holds
constexpr int kAdjStripBytes = 64 * 1024;
constexpr int kAdjLargeMaxSlots = 4096;
constexpr int kPixBits = 22;  // argmax key: value << 22 | (2^22 - 1 - pixel)
constexpr int kNRed = 20;     // block sums: loss, sum G P, dRt[9], dL/drec[9]
constexpr int kAdjTile = 16;      // tile edge of the fixed-point bound
constexpr int kAdjMaxTiles = 1024;
constexpr int kAdjMaxCand = 16;    // candidate tiles of the maximum search
constexpr int kAdjCandSpots = 64;  // spots touching one candidate tile
// float4 target values each thread of the two-blocks-per-CU kernel prefetches
// for one strip of at most kAdjSmallStripBytes / 8 pixels (the large-pattern
// kernel, whose strips take long to splat, loads them in the read-out)
constexpr int kAdjPre = 4;
constexpr int kAdjSmallStripBytes = kAdjPre * 4 * kAdjThreads * 8;
constexpr float kAdjRange = 2040109465.6f;  // 1.9 * 2^30 units per bound
constexpr float kAdjBoundUnits = 1048576.0f;  // 2^20 bound units per largest spot

// 32-bit fixed-point strip words: the largest per-tile sum U of the spots'
// peak values bounds every pixel, so one unit = U / (1.9 * 2^30) cannot
// overflow (the bound is an integer sum, independent of the atomics' order).
typedef int aacc_t;

struct AdjArgs {
  const int *idx;
  const float4 *spots;
  const int *count;
  int cap;
  int n;
  const int4 *hkl;
  const float *F2;
  int n_refl;
  const double *Rt;
  double rec[9];
  // physics (fill_physics)
  double k, k2, me, width, sigma, fac;
  int model, precession;
  double phi, Pz, Pt;
  // detector
  int H, W;
  float cal, cx, cy, ca, sa;
  int mirrored, normalize;
  float pref, ep, inv_s2, clip;
  float e2ep;  // exp(2 ep): ratio of consecutive row ratios of the Gaussian
  int strip_rows, n_strips, acc_words;
  int slots, bin_cap;       // LDS spot slots and strip-bin entries
  int mode;
  int want_grad;
  long long target_stride;  // elements between targets (0: one target for every pattern)
  int *work;                // large patterns: per-block records, bins and boxes (may be NULL)
  long long work_per_block; // ints
  int work_entries;         // bin entries per block
  float4 *gsum;             // per pattern and slot: A, BX, BY, list position (chain kernel input)
  int *gcount;              // per pattern: slots used
  const float *target;
  float *pattern;
  float *loss;
  double *dRt;
  float *dF2;    // [n_refl] dL/d|F|^2, accumulated (may be NULL)
  double *drec;  // [9] sum dL/dg (x) hkl, accumulated (may be NULL)
  double *dpart; // chain kernel: per-block partial sums of drec
  int *status;
};

template <int kCls>  // width classes per strip
struct AdjShared {
  int bin_start[kCls * kAdjMaxStrips + 1];  // bin b = kCls * strip + width class
  int bin_fill[kCls * kAdjMaxStrips];
  int ns, cols_max;
  float wmaxf[kAdjWaves];
  long long wkey[kAdjWaves];
  int wu[kAdjWaves];
  int qstar, n_cand, fail;
  long long lkey;
  int cand_tile[kAdjMaxCand], cand_n[kAdjMaxCand];
  double red[kAdjWaves][kNRed];
};

// The spot table lives in dynamic LDS after the two strip buffers (acc_words
// fixed-point words, acc_words gradient floats), sized by the host: `slots`
// entries of each field, then `bin_cap` strip-bin entries.
struct SpotTab {
  float *I, *X, *Y, *A, *BX, *BY;
  int *xb;    // (xlo << 16) | xhi
  int *jpos;  // position in the pattern's spot list
  short *ylo, *yhi, *bin;
};

constexpr int kSpotTabBytes = 6 * 4 + 2 * 4 + 2 * 2;  // per slot

__device__ __forceinline__ SpotTab spot_tab(aacc_t *s_acc, size_t strip_words, int slots) {
  float *f = reinterpret_cast<float *>(s_acc + strip_words);
  SpotTab t;
  t.I = f;
  t.X = f + slots;
  t.Y = f + 2 * slots;
  t.A = f + 3 * slots;
  t.BX = f + 4 * slots;
  t.BY = f + 5 * slots;
  int *iv = reinterpret_cast<int *>(f + 6 * slots);
  t.xb = iv;
  t.jpos = iv + slots;
  short *sv = reinterpret_cast<short *>(iv + 2 * slots);
  t.ylo = sv;
  t.yhi = sv + slots;
  t.bin = sv + 2 * slots;
  return t;
}

// Reflection-side constants of the chain rule, staged in LDS once per block so
// the kernels do not hold them in scalar registers throughout.
struct ChainConsts {
  double rec[9];
  double k, k2, me, width, sigma, fac;
  int model, precession;
  double phi, Pz, Pt;
};

__device__ __forceinline__ void stage_chain_consts(const AdjArgs &a, ChainConsts &c) {
  for (int i = 0; i < 9; ++i) c.rec[i] = a.rec[i];
  c.k = a.k;
  c.k2 = a.k2;
  c.me = a.me;
  c.width = a.width;
  c.sigma = a.sigma;
  c.fac = a.fac;
  c.model = a.model;
  c.precession = a.precession;
  c.phi = a.phi;
  c.Pz = a.Pz;
  c.Pt = a.Pt;
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Fixed-point value of one Gaussian term amp exp(ep r^2) at offset (dx, dy),
// written once so the splat and the maximum search round it identically.
__device__ __forceinline__ int gauss_term(float amp, float dx, float dy, float ep, float scale) {
  return __float2int_rn(amp * fast_exp(ep * fmaf(dy, dy, dx * dx)) * scale);
}

__device__ __forceinline__ int width_class(int w) { return w <= 4 ? 0 : (w <= 8 ? 1 : 2); }
__device__ __forceinline__ int class_cols(int c, int cols_max) {
  return c == 0 ? min(4, cols_max) : (c == 1 ? min(8, cols_max) : cols_max);
}
__device__ __forceinline__ int box_width(int xb) { return (xb & 0xffff) - (xb >> 16) + 1; }
// it / ncols, a shift for powers of two (ncols is uniform)
__device__ __forceinline__ int item_spot(int it, int ncols) {
  return (ncols & (ncols - 1)) == 0 ? it >> __builtin_ctz(ncols) : it / ncols;
}
// Column count of a gather pass: the next power of two up to a wave, so a
// spot's run of items never straddles two waves and its first lane ends up
// with the whole sum (no LDS float atomics, which run far slower than
// integer ones)
__device__ __forceinline__ int gather_cols(int w) {
  return w <= 1 ? 1 : (w <= kWave ? 1 << (32 - __builtin_clz(w - 1)) : w);
}

// Splat (kGather = false) or gather (true) of every binned spot footprint
// inside strip rows [y0, y1).  Work item = (spot, footprint column) so a wave
// touches consecutive LDS columns.
template <bool kGather>
__device__ __forceinline__ void strip_pass(const AdjArgs &a, const SpotTab &t, int cols_max,
                                           aacc_t *acc, const float *G, int b0, int nb,
                                           bool binned, int y0, int y1, float scale, int tid,
                                           float4 *hand = nullptr) {
  // unbinned: LDS slots b0 .. b0 + nb - 1; hand: the gather adds each spot's
  // sums into row jpos of the gradient hand-over instead of its LDS slot
  const int W = a.W;
  const int ncols = cols_max;
  const int n_items = nb * ncols;
  if (!kGather) {
    for (int it = tid; it < n_items; it += kAdjThreads) {
      const int bi = item_spot(it, ncols);
      const int c = it - bi * ncols;
      const int q = binned ? t.bin[b0 + bi] : b0 + bi;
      const int ya = max(static_cast<int>(t.ylo[q]), y0);
      const int yb = min(static_cast<int>(t.yhi[q]), y1 - 1);
      if (ya > yb) continue;
      const int xlo = t.xb[q] >> 16, xhi = t.xb[q] & 0xffff;
      const int x = xlo + c;
      if (x > xhi) continue;
      const float dx = static_cast<float>(x) - t.X[q];
      const float cy = t.Y[q];
      const float amp = t.I[q] * a.pref;
      aacc_t *col = acc + x;
      lds_reads_done();
      for (int y = ya; y <= yb; ++y) {
        const float dy = static_cast<float>(y) - cy;
        atomicAdd(col + (y - y0) * W, gauss_term(amp, dx, dy, a.ep, scale));
      }
    }
    return;
  }
  // Gather: the ncols consecutive items of a spot sit in consecutive lanes,
  // so their column sums are combined across lanes first (segmented shuffle
  // reduction) and only the first lane of each spot's run adds them to LDS:
  // LDS float atomics to one address serialise, and a spot's columns would
  // otherwise all hit the same three words.
  const int lane = lane_id();
  for (int base = tid - lane; base < n_items; base += kAdjThreads) {
    const int it = base + lane;
    const int bi = item_spot(it, ncols);
    const int c = it - bi * ncols;
    int q = -1;
    float sa = 0.0f, sx = 0.0f, sy = 0.0f;
    if (it < n_items) {
      q = binned ? t.bin[b0 + bi] : b0 + bi;
      const int ya = max(static_cast<int>(t.ylo[q]), y0);
      const int yb = min(static_cast<int>(t.yhi[q]), y1 - 1);
      const int xlo = t.xb[q] >> 16, xhi = t.xb[q] & 0xffff;
      const int x = xlo + c;
      if (ya <= yb && x <= xhi) {
        // down the column the Gaussian follows K(y + 1) = K(y) r(y) with
        // r(y + 1) = r(y) exp(2 ep): two products per row instead of an exp
        // (a few ulps over a box; the splat keeps its exact per-pixel terms)
        const float dx = static_cast<float>(x) - t.X[q];
        float dy = static_cast<float>(ya) - t.Y[q];
        float k = a.pref * fast_exp(a.ep * (dy * dy + dx * dx));
        float r = fast_exp(a.ep * (2.0f * dy + 1.0f));
        const float *g = G + (ya - y0) * W + x;
        auto row = [&](float gv) {
          const float kg = k * gv;
          sa += kg;
          sy = fmaf(kg, dy, sy);
          k *= r;
          r *= a.e2ep;
          dy += 1.0f;
        };
        // four rows' G words are read together, so one LDS round trip
        // serves four rows
        int y = ya;
        for (; y + 3 <= yb; y += 4, g += 4 * W) {
          const float g0 = g[0], g1 = g[W], g2 = g[2 * W], g3 = g[3 * W];
          row(g0);
          row(g1);
          row(g2);
          row(g3);
        }
        for (; y <= yb; ++y, g += W) row(*g);
        sx = sa * dx;
      }
    }
    // lane L takes lane L + d while both lie in the same spot's run; after
    // the steps the first lane of each run (in this wave) holds its sums
    for (int d = 1; d < ncols && d < kWave; d <<= 1) {
      const float oa = __shfl_down(sa, d, kWave);
      const float ox = __shfl_down(sx, d, kWave);
      const float oy = __shfl_down(sy, d, kWave);
      if (c + d < ncols && lane + d < kWave) {
        sa += oa;
        sx += ox;
        sy += oy;
      }
    }
    if (q >= 0 && (c == 0 || lane == 0) && (sa != 0.0f || sx != 0.0f || sy != 0.0f)) {
      if ((ncols & (ncols - 1)) == 0 && ncols <= kWave) {
        // the run lies in this wave and this lane (c == 0) holds its sum;
        // a spot has one run per pass
        if (hand) {
          float *h = reinterpret_cast<float *>(hand + t.jpos[q]);
          unsafeAtomicAdd(h, sa);
          unsafeAtomicAdd(h + 1, sx);
          unsafeAtomicAdd(h + 2, sy);
        } else {
          t.A[q] += sa;
          t.BX[q] += sx;
          t.BY[q] += sy;
        }
      } else {
        atomicAdd(&t.A[q], sa);
        atomicAdd(&t.BX[q], sx);
        atomicAdd(&t.BY[q], sy);
      }
    }
  }
}

// Load entries [base, base + m) of a strip's global bin into LDS slots 0..m-1
// (large patterns: more in-frame spots than the LDS holds).  The clip boxes
// come from the per-spot cache the first scan wrote (boxes[j] = {(xlo << 16) |
// xhi, (lo << 16) | hi}); only the pixel position is recomputed.  jpos holds
// the spot's row of the gradient hand-over (row_of[j]) when there is one.
__device__ __forceinline__ void load_chunk(const AdjArgs &a, const SpotTab &t, const float4 *spots,
                                           const int *bins, const int2 *boxes, const int *row_of,
                                           int base, int m, int tid) {
  for (int i = tid; i < m; i += kAdjThreads) {
    const int j = bins[base + i];
    const float4 sp = spots[j];
    const int2 bx = boxes[j];
    float X, Y;
    to_pixel(sp, a, X, Y);  // in frame: the scan that binned j tested it
    t.I[i] = sp.w;
    t.X[i] = X;
    t.Y[i] = Y;
    t.A[i] = 0.0f;
    t.BX[i] = 0.0f;
    t.BY[i] = 0.0f;
    t.xb[i] = bx.x;
    t.jpos[i] = row_of ? row_of[j] : j;
    t.ylo[i] = static_cast<short>(bx.y >> 16);
    t.yhi[i] = static_cast<short>(bx.y & 0xffff);
  }
}

// Chain rule of one spot: (dL/dI, dL/dX, dL/dY) -> dL/dRt of the pattern (dR,
// accumulated), dL/dg (x) hkl of its reflection (dG, accumulated when a.drec)
// and dL/d|F|^2 of its reflection (a global atomic when a.dF2).
__device__ __forceinline__ void spot_chain(const AdjArgs &a, const ChainConsts &c, const double *R,
                                           int row, double I, double dI, double BX, double BY,
                                           double *dR, double *dG) {
  const double dX = I * BX * a.inv_s2;
  const double dY = I * BY * a.inv_s2;
  if (dI == 0.0 && dX == 0.0 && dY == 0.0) return;
  const double ca = a.ca, sa = a.sa, ic = 1.0 / static_cast<double>(a.cal);
  double dx, dy;  // d/d(lab x, lab y)
  if (a.mirrored) {
    dx = (dX * ca + dY * sa) * ic;
    dy = (dX * sa - dY * ca) * ic;
  } else {
    dx = (dX * ca + dY * sa) * ic;
    dy = (-dX * sa + dY * ca) * ic;
  }
  const int4 h = a.hkl[row];
  const double hx = h.x, hy = h.y, hz = h.z;
  const double g0 = c.rec[0] * hx + c.rec[1] * hy + c.rec[2] * hz;
  const double g1 = c.rec[3] * hx + c.rec[4] * hy + c.rec[5] * hz;
  const double g2 = c.rec[6] * hx + c.rec[7] * hy + c.rec[8] * hz;
  const double x = R[0] * g0 + R[1] * g1 + R[2] * g2;
  const double y = R[3] * g0 + R[4] * g1 + R[5] * g2;
  const double z = R[6] * g0 + R[7] * g1 + R[8] * g2;
  const double r2 = x * x + y * y;
  const double r = sqrt(r2);
  const double qk = sqrt(fmax(c.k2 - r2, 0.0));
  const double sv = r2 / (c.k + qk) - z;
  double fs, fr;
  const double f = shape_factor_grad(sv, r, c, fs, fr);
  const double F2 = a.F2[row];
  const double dF2 = dI * f;
  const double ds = dI * F2 * fs;
  const double dr = dI * F2 * fr;
  double gx = dx, gy = dy, gz = -ds;
  if (qk > 0.0) {
    gx += ds * x / qk;
    gy += ds * y / qk;
  }
  if (r > 0.0 && dr != 0.0) {
    gx += dr * x / r;
    gy += dr * y / r;
  }
  // g_lab = Rt g: dL/dRt[i][c] = dL/dg_lab[i] g[c]; dL/dg = Rt^T dL/dg_lab
  dR[0] += gx * g0; dR[1] += gx * g1; dR[2] += gx * g2;
  dR[3] += gy * g0; dR[4] += gy * g1; dR[5] += gy * g2;
  dR[6] += gz * g0; dR[7] += gz * g1; dR[8] += gz * g2;
  if (a.drec) {
    // dL/dg (crystal frame) = Rt^T dL/dg_lab; g = rec hkl: dL/drec += dL/dg hkl^T
    const double d0 = R[0] * gx + R[3] * gy + R[6] * gz;
    const double d1 = R[1] * gx + R[4] * gy + R[7] * gz;
    const double d2 = R[2] * gx + R[5] * gy + R[8] * gz;
    dG[0] += d0 * hx; dG[1] += d0 * hy; dG[2] += d0 * hz;
    dG[3] += d1 * hx; dG[4] += d1 * hy; dG[5] += d1 * hz;
    dG[6] += d2 * hx; dG[7] += d2 * hy; dG[8] += d2 * hz;
  }
  if (a.dF2) unsafeAtomicAdd(a.dF2 + row, static_cast<float>(dF2));
}

// The max term of the normalisation (-(sum G P)/m at the maximum's pixel p*)
// for a spot whose box holds p*, added to its gathered sums.
__device__ __forceinline__ void max_term(const AdjArgs &a, float corr, int px, int py, float X,
                                         float Y, int xlo, int xhi, int lo, int hi, float &A,
                                         float &BX, float &BY) {
  if (px < xlo || px > xhi || py < lo || py > hi) return;
  const float dx = static_cast<float>(px) - X, dy = static_cast<float>(py) - Y;
  const float kc = corr * a.pref * expf(a.ep * (dx * dx + dy * dy));
  A -= kc;
  BX -= kc * dx;
  BY -= kc * dy;
}

// Fixed-point value at pixel (x, y) of LDS spot q (0 outside its clip box).
__device__ __forceinline__ int spot_at(const AdjArgs &a, const SpotTab &t, int q, int x, int y,
                                       float scale) {
  if (y < t.ylo[q] || y > t.yhi[q] || x < (t.xb[q] >> 16) || x > (t.xb[q] & 0xffff)) return 0;
  return gauss_term(t.I[q] * a.pref, static_cast<float>(x) - t.X[q],
                    static_cast<float>(y) - t.Y[q], a.ep, scale);
}

// Pattern maximum and its first pixel, as the key value << kPixBits |
// (2^kPixBits - 1 - pixel), without rendering the whole pattern: a tile whose
// bound (the per-tile sums tile_u of the spots' peak values, in units of
// `unit`) stays below the exact value at the brightest spot's nearest pixel
// cannot hold the maximum, and the few remaining tiles are evaluated pixel by
// pixel with the splat's own fixed-point terms (integer sums: the result equals
// a full pass bit for bit).  Every pixel holding the maximum lies in a
// remaining tile, so the first of them is found.  Returns false (uniformly)
// when too many tiles or spots remain; the caller then runs the full pass.
// tile_u is overwritten (candidate slots); cand_q borrows the gradient buffer.
template <class Shared>
__device__ bool bounded_max_key(const AdjArgs &a, Shared &sh, const SpotTab &t, int *tile_u,
                                short (*cand_q)[kAdjCandSpots], int ntx, int n_tiles, int ns,
                                float scale, float unit, int tid, int lane, int wid,
                                long long &key_out) {
  const int H = a.H, W = a.W;
  if (tid == 0) {
    sh.n_cand = 0;
    sh.fail = 0;
  }
  if (wid == 0) {
    const int qs = sh.qstar;
    const int xs = min(max(static_cast<int>(rintf(t.X[qs])), t.xb[qs] >> 16), t.xb[qs] & 0xffff);
    const int ys = min(max(static_cast<int>(rintf(t.Y[qs])), static_cast<int>(t.ylo[qs])),
                       static_cast<int>(t.yhi[qs]));
    long long acc = 0;
    for (int q = lane; q < ns; q += kWave) acc += spot_at(a, t, q, xs, ys, scale);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, kWave);
    if (lane == 0)
      sh.lkey = (max(acc, 0ll) << kPixBits) |
                static_cast<long long>((1 << kPixBits) - 1 - (ys * W + xs));
  }
  lds_sync();
  const long long lkey = sh.lkey;
  const float lfix = static_cast<float>(lkey >> kPixBits);
  // per term: the float product is within 1e-5 of the bound's share and the
  // rounding adds at most one half
  const float margin = 0.5f * static_cast<float>(ns) + 1.0f;
  int *slot_of = tile_u;
  for (int ti = tid; ti < n_tiles; ti += kAdjThreads) {
    const float bound = static_cast<float>(tile_u[ti]) * unit * scale * 1.00001f + margin;
    int slot = -1;
    if (bound >= lfix) {
      slot = atomicAdd(&sh.n_cand, 1);
      if (slot < kAdjMaxCand) {
        sh.cand_tile[slot] = ti;
        sh.cand_n[slot] = 0;
      } else {
        slot = -1;
        sh.fail = 1;
      }
    }
    slot_of[ti] = slot;
  }
  lds_sync();
  if (sh.fail) return false;
  const int n_cand = sh.n_cand;
  for (int q = tid; q < ns; q += kAdjThreads) {
    const int tx0 = (t.xb[q] >> 16) / kAdjTile, tx1 = (t.xb[q] & 0xffff) / kAdjTile;
    const int ty0 = t.ylo[q] / kAdjTile, ty1 = t.yhi[q] / kAdjTile;
    for (int ty = ty0; ty <= ty1; ++ty)
      for (int tx = tx0; tx <= tx1; ++tx) {
        const int sl = slot_of[ty * ntx + tx];
        if (sl < 0) continue;
        const int pos = atomicAdd(&sh.cand_n[sl], 1);
        if (pos < kAdjCandSpots)
          cand_q[sl][pos] = static_cast<short>(q);
        else
          sh.fail = 1;
      }
  }
  lds_sync();
  if (sh.fail) return false;
  long long key = lkey;
  for (int i = tid; i < n_cand * kAdjTile * kAdjTile; i += kAdjThreads) {
    const int sl = i / (kAdjTile * kAdjTile);
    const int ti = sh.cand_tile[sl];
    const int x = (ti % ntx) * kAdjTile + (i % kAdjTile);
    const int y = (ti / ntx) * kAdjTile + (i / kAdjTile) % kAdjTile;
    if (x >= W || y >= H) continue;
    long long acc = 0;
    const int nq = sh.cand_n[sl];
    for (int k = 0; k < nq; ++k) acc += spot_at(a, t, cand_q[sl][k], x, y, scale);
    const long long kv =
        (max(acc, 0ll) << kPixBits) | static_cast<long long>((1 << kPixBits) - 1 - (y * W + x));
    key = kv > key ? kv : key;
  }
  key = wave_max_acc(key);
  if (lane == 0) sh.wkey[wid] = key;
  lds_sync();
  key = 0;
  for (int w = 0; w < kAdjWaves; ++w) key = sh.wkey[w] > key ? sh.wkey[w] : key;
  key_out = key;
  lds_sync();  // wkey and the candidate state are reused
  return true;
}

// One pattern per block iteration.  kLarge = false: every in-frame spot fits
// the LDS slots (the host guarantees slots >= cap) and the per-spot sums go to
// gsum for spot_chain_kernel, which keeps the FP64 chain rule (and its
// registers) out of this kernel.  kLarge = true: patterns may hold more
// in-frame spots than the slots (their bins and sums live in the per-block
// workspace) and the chain rule runs here, per pattern.
template <bool kLarge>
__device__ __forceinline__ void adjoint_body(const AdjArgs &a, aacc_t *s_acc,
                                             AdjShared<kLarge ? kWidthClasses : 1> &sh) {
  constexpr int kCls = kLarge ? kWidthClasses : 1;
  // The large-pattern kernel has one strip buffer: pass 1 splats into it, and
  // pass 2 keeps G there while the fixed-point pixels wait in q.  The small
  // one splats and reads out into G in the same step, so it has two.
  float *s_G = kLarge ? reinterpret_cast<float *>(s_acc)
                      : reinterpret_cast<float *>(s_acc + a.acc_words);
  const SpotTab t = spot_tab(s_acc, (kLarge ? 1 : 2) * static_cast<size_t>(a.acc_words), a.slots);
  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int wid = tid / kWave;
  const int H = a.H, W = a.W;
  const size_t HW = static_cast<size_t>(H) * W;
  // large-pattern workspace of this block: int bins[entries] | int2
  // boxes[cap] (clip boxes cached by the first scan) | int rows[cap] (each
  // in-frame spot's row of the gradient hand-over) | int q[H W] (the
  // fixed-point pattern between the splat and the read-out when there is no
  // pattern buffer to hold it), each part 16-byte aligned
  int *wblock = kLarge && a.work ? a.work + (size_t)blockIdx.x * a.work_per_block : nullptr;
  int *wbins = wblock;
  int2 *wbox = wblock ? reinterpret_cast<int2 *>(wbins + ((static_cast<size_t>(a.work_entries) + 3) &
                                                          ~static_cast<size_t>(3)))
                      : nullptr;
  int *wrow = wblock ? reinterpret_cast<int *>(wbox) + ((2 * static_cast<size_t>(a.cap) + 3) &
                                                         ~static_cast<size_t>(3))
                     : nullptr;
  int *wq = wblock ? wrow + ((static_cast<size_t>(a.cap) + 3) & ~static_cast<size_t>(3)) : nullptr;
  for (int i = tid; i < a.acc_words; i += kAdjThreads) s_acc[i] = 0;
  lds_sync();

  for (int p = blockIdx.x; p < a.n; p += gridDim.x) {
    const int n = min(a.count[p], a.cap);
    const float4 *spots = a.spots + (size_t)p * a.cap;
    if (tid == 0) {
      sh.ns = 0;
      sh.cols_max = 1;
      sh.qstar = 0;
    }
    for (int b = tid; b <= kCls * kAdjMaxStrips; b += kAdjThreads) sh.bin_start[b] = 0;
    lds_sync();

    // ---- 1. spots in frame with their clip boxes --------------------------
    // every such spot counts in the strip bins and the maximum; the first
    // `slots` also go to LDS (all of them unless the pattern is large)
    float tmax = 0.0f;
    float4 *gsp = kLarge && a.gsum ? a.gsum + (size_t)p * a.cap : nullptr;
    for (int j = tid; j < n; j += kAdjThreads) {
      const float4 sp = spots[j];
      float X = 0.0f, Y = 0.0f;
      int lo = 1, hi = 0, xlo = 0, xhi = 0;
      const bool in = to_pixel(sp, a, X, Y) &&
                      subpixel_box(sp.w, X, Y, a.pref, a.ep, a.clip, H, W, lo, hi, xlo, xhi);
      // the large-pattern scans below read the cached box ({0, 1 << 16}: none)
      if (kLarge && wbox) wbox[j] = in ? make_int2((xlo << 16) | xhi, (lo << 16) | hi) : make_int2(0, 1 << 16);
      if (!in) continue;
      tmax = fmaxf(tmax, fabsf(sp.w) * a.pref);
      atomicMax(&sh.cols_max, xhi - xlo + 1);
      const int wc = kCls == 1 ? 0 : width_class(xhi - xlo + 1);
      for (int s = lo / a.strip_rows; s <= hi / a.strip_rows; ++s)
        atomicAdd(&sh.bin_start[kCls * s + wc + 1], 1);
      const int q = atomicAdd(&sh.ns, 1);
      if (kLarge) {
        // gradients: in-frame spot q's gathered sums start at zero in row q
        // of the hand-over (which names its spot j)
        if (gsp) {
          gsp[q] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(j));
          wrow[j] = q;
        }
        continue;  // its chunks are loaded per strip
      }
      t.I[q] = sp.w;
      t.X[q] = X;
      t.Y[q] = Y;
      t.A[q] = 0.0f;
      t.BX[q] = 0.0f;
      t.BY[q] = 0.0f;
      t.xb[q] = (xlo << 16) | xhi;
      t.jpos[q] = j;
      t.ylo[q] = static_cast<short>(lo);
      t.yhi[q] = static_cast<short>(hi);
    }
    tmax = wave_max(tmax);
    if (lane == 0) sh.wmaxf[wid] = tmax;
    if (kLarge)
      __syncthreads();  // the cached boxes in global memory, block-visible
    else
      lds_sync();
    const int ns_all = sh.ns;
    float pmax = 0.0f;
    for (int w = 0; w < kAdjWaves; ++w) pmax = fmaxf(pmax, sh.wmaxf[w]);
    float scale = 1.0f;  // set from the tile bounds below

    // ---- 2. strip bins ----------------------------------------------------
    const int n_bins = kCls * a.n_strips;
    if (tid == 0) {
      int acc = 0;
      for (int b = 0; b < n_bins; ++b) {
        acc += sh.bin_start[b + 1];
        sh.bin_start[b + 1] = acc;
      }
    }
    lds_sync();
    const int entries = sh.bin_start[n_bins];
    // large-pattern kernel: the bins live in the caller's workspace and the
    // spots pass through LDS in chunks
    const bool large = kLarge;
    const bool fits = !large || (wbins != nullptr && entries <= a.work_entries);
    if (tid == 0 && a.status) a.status[p] = fits ? 0 : 1;
    const int ns = large ? 0 : ns_all;  // spots held in LDS for the whole pattern
    const bool binned = !large && entries <= a.bin_cap;
    for (int b = tid; b < n_bins; b += kAdjThreads) sh.bin_fill[b] = sh.bin_start[b];
    lds_sync();
    if (kLarge && !fits) {
      // no room: the pattern is reported in status and left as zeros
      if (a.pattern)
        for (size_t i = tid; i < HW; i += kAdjThreads) a.pattern[(size_t)p * HW + i] = 0.0f;
      if (tid < 9 && a.want_grad) a.dRt[(size_t)p * 9 + tid] = 0.0;
      if (tid == 0 && a.loss) a.loss[p] = 0.0f;
      lds_sync();
      continue;
    }
    if (large) {
      for (int j = tid; j < n; j += kAdjThreads) {
        const int2 bx = wbox[j];
        const int lo = bx.y >> 16, hi = bx.y & 0xffff;
        if (lo > hi) continue;  // not in frame
        const int wc = kCls == 1 ? 0 : width_class(box_width(bx.x));
        for (int s = lo / a.strip_rows; s <= hi / a.strip_rows; ++s)
          wbins[atomicAdd(&sh.bin_fill[kCls * s + wc], 1)] = j;
      }
      __syncthreads();  // bins in global memory, block-visible
    } else if (binned) {
      for (int q = tid; q < ns; q += kAdjThreads) {
        const int wc = kCls == 1 ? 0 : width_class(box_width(t.xb[q]));
        for (int s = t.ylo[q] / a.strip_rows; s <= t.yhi[q] / a.strip_rows; ++s) {
          const int pos = atomicAdd(&sh.bin_fill[kCls * s + wc], 1);
          t.bin[pos] = static_cast<short>(q);
        }
      }
    }
    lds_sync();
    const int cols_max = sh.cols_max;

    // ---- 2b. fixed-point scale: per-tile sums of the spots' peak values ------
    const int ntx = (W + kAdjTile - 1) / kAdjTile;
    const int n_tiles = ntx * ((H + kAdjTile - 1) / kAdjTile);
    const bool tiles = n_tiles <= kAdjMaxTiles && n_tiles <= a.acc_words;
    int *tile_u = s_acc;  // the strip buffer is zero here and zeroed again below
    const float unit = pmax / kAdjBoundUnits;  // value of one bound unit
    {
      const float inv_p = pmax > 0.0f ? kAdjBoundUnits / pmax : 0.0f;
      int usum = 0;
      auto add_bound = [&](float I, int lo, int hi, int xlo, int xhi) {
        const int u = min(static_cast<int>(ceilf(fabsf(I) * a.pref * inv_p)) + 1,
                          static_cast<int>(kAdjBoundUnits) + 1);
        usum += u;
        if (!tiles) return;
        for (int ty = lo / kAdjTile; ty <= hi / kAdjTile; ++ty)
          for (int tx = xlo / kAdjTile; tx <= xhi / kAdjTile; ++tx) atomicAdd(&tile_u[ty * ntx + tx], u);
      };
      if (!large) {
        for (int q = tid; q < ns; q += kAdjThreads) {
          if (fabsf(t.I[q]) * a.pref == pmax) sh.qstar = q;  // the brightest spot (any of a tie)
          add_bound(t.I[q], t.ylo[q], t.yhi[q], t.xb[q] >> 16, t.xb[q] & 0xffff);
        }
      } else {
        for (int j = tid; j < n; j += kAdjThreads) {
          const int2 bx = wbox[j];
          const int lo = bx.y >> 16, hi = bx.y & 0xffff;
          if (lo > hi) continue;  // not in frame
          add_bound(spots[j].w, lo, hi, bx.x >> 16, bx.x & 0xffff);
        }
      }
      int lu = 0;
      if (tiles) {
        lds_sync();
        for (int ti = tid; ti < n_tiles; ti += kAdjThreads) lu = max(lu, tile_u[ti]);
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) lu = max(lu, __shfl_xor(lu, off, kWave));
      } else {
        lu = wave_sum(usum);
      }
      if (lane == 0) sh.wu[wid] = lu;
      lds_sync();
      int Ui = 0;
      for (int w = 0; w < kAdjWaves; ++w) Ui = tiles ? max(Ui, sh.wu[w]) : Ui + sh.wu[w];
      const float U = static_cast<float>(Ui) * unit;
      scale = U > 0.0f ? kAdjRange / (U * 1.00001f) : 1.0f;
    }

    // splat (or gather) strip s: every LDS-resident spot of its bin, or, for a
    // large pattern, its global bin chunk by chunk.  A strip whose bin fits
    // one chunk is gathered from the slots its splat just loaded.  Large
    // patterns store each entry's gathered sums in the workspace record of
    // the entry (no read-modify-write; every entry belongs to one strip).
    auto pass_range = [&](bool gather, int ncols, int b0, int nb, bool bin, int y0, int y1) {
      if (nb <= 0) return;
      if (gather)
        strip_pass<true>(a, t, gather_cols(ncols), s_acc, s_G, b0, nb, bin, y0, y1, scale, tid,
                         large ? gsp : nullptr);
      else
        strip_pass<false>(a, t, ncols, s_acc, nullptr, b0, nb, bin, y0, y1, scale, tid);
    };
    auto strip_spots = [&](int s, bool gather, int y0, int y1) {
      const int bs = kCls * s;
      if (!large) {
        // few spots per strip: one pass over the strip's classes together
        // fills the waves better than a pass per class
        if (binned)
          pass_range(gather, cols_max, sh.bin_start[bs],
                     sh.bin_start[bs + kCls] - sh.bin_start[bs], true, y0, y1);
        else
          pass_range(gather, cols_max, 0, ns, false, y0, y1);
        return;
      }
      const int e0 = sh.bin_start[bs], e1 = sh.bin_start[bs + kCls];
      for (int c0 = e0; c0 < e1; c0 += a.slots) {
        const int m = min(a.slots, e1 - c0);
        load_chunk(a, t, spots, wbins, wbox, gsp ? wrow : nullptr, c0, m, tid);
        lds_sync();
        // the chunk's entries in slots 0..m-1, ordered by width class
        auto classes = [&](bool g) {
          for (int c = 0; c < kCls; ++c) {
            const int lo = max(c0, sh.bin_start[bs + c]);
            const int hi = min(c0 + m, sh.bin_start[bs + c + 1]);
            pass_range(g, kCls == 1 ? cols_max : class_cols(c, cols_max), lo - c0, hi - lo, false,
                       y0, y1);
          }
        };
        if (gather) {
          // with boxes up to a wave wide every run lies in one wave and the
          // gather itself adds the sums into the hand-over rows; wider boxes
          // collect theirs in the LDS slots, added here (a spot's strips add
          // in turn; the first scan zeroed its row behind a barrier)
          classes(true);
          lds_sync();
          if (cols_max > kWave) {
            for (int i = tid; i < m; i += kAdjThreads) {
              const float A = t.A[i], BX = t.BX[i], BY = t.BY[i];
              if (A != 0.0f || BX != 0.0f || BY != 0.0f) {
                float *g = reinterpret_cast<float *>(gsp + t.jpos[i]);
                unsafeAtomicAdd(g, A);
                unsafeAtomicAdd(g + 1, BX);
                unsafeAtomicAdd(g + 2, BY);
              }
            }
            lds_sync();  // the next chunk overwrites the slots
          }
        } else {
          classes(false);
          lds_sync();  // the next chunk overwrites the slots
        }
      }
    };

    // ---- 3. pass A: maximum and its first pixel -----------------------------
    long long key = 0;
    bool have_key = false;
    if (a.normalize && ns_all > 0 && tiles && !large &&
        a.acc_words * sizeof(float) >= sizeof(short) * kAdjMaxCand * kAdjCandSpots) {
      have_key = bounded_max_key(a, sh, t, tile_u, reinterpret_cast<short(*)[kAdjCandSpots]>(s_G),
                                 ntx, n_tiles, ns, scale, unit, tid, lane, wid, key);
    }
    if (tiles) {  // hand the borrowed words back zeroed
      for (int ti = tid; ti < n_tiles; ti += kAdjThreads) tile_u[ti] = 0;
      lds_sync();
    }
    // A large pattern is splatted once, into the strip buffer (zero again
    // from here): each strip's fixed-point values go to q (the pattern buffer
    // when there is one, else the block's workspace) while the maximum key is
    // formed, and pass B reads them back.  (Small patterns find the maximum
    // from the tile bounds and splat in pass B.)
    int *q = large ? (a.pattern ? reinterpret_cast<int *>(a.pattern + (size_t)p * HW) : wq)
                   : nullptr;
    if (large) {
      for (int s = 0; s < a.n_strips; ++s) {
        const int y0 = s * a.strip_rows;
        const int y1 = min(H, y0 + a.strip_rows);
        if (sh.bin_start[kCls * (s + 1)] == sh.bin_start[kCls * s]) continue;
        strip_spots(s, false, y0, y1);
        lds_sync();
        const int cnt = (y1 - y0) * W;
        const int pix0 = y0 * W;
        auto take = [&](int v, int pix) {
          const long long kv = (static_cast<long long>(max(v, 0)) << kPixBits) |
                               static_cast<long long>((1 << kPixBits) - 1 - pix);
          key = kv > key ? kv : key;
        };
        if ((W & 3) == 0) {
          for (int i = tid; i < cnt / 4; i += kAdjThreads) {
            int4 *pv = reinterpret_cast<int4 *>(s_acc) + i;
            const int4 u = *pv;
            *pv = make_int4(0, 0, 0, 0);
            reinterpret_cast<int4 *>(q + pix0)[i] = u;
            if (a.normalize) {
              take(u.x, pix0 + 4 * i);
              take(u.y, pix0 + 4 * i + 1);
              take(u.z, pix0 + 4 * i + 2);
              take(u.w, pix0 + 4 * i + 3);
            }
          }
        } else {
          for (int i = tid; i < cnt; i += kAdjThreads) {
            const aacc_t v = s_acc[i];
            s_acc[i] = 0;
            q[pix0 + i] = v;
            if (a.normalize) take(v, pix0 + i);
          }
        }
        lds_sync();
      }
      if (a.normalize) {
        key = wave_max_acc(key);
        if (lane == 0) sh.wkey[wid] = key;
        lds_sync();
        key = 0;
        for (int w = 0; w < kAdjWaves; ++w) key = sh.wkey[w] > key ? sh.wkey[w] : key;
      }
      have_key = true;
    }
    if (a.normalize && ns_all > 0 && !have_key && !large) {
      for (int s = 0; s < a.n_strips; ++s) {
        const int y0 = s * a.strip_rows;
        const int y1 = min(H, y0 + a.strip_rows);
        if (sh.bin_start[kCls * (s + 1)] == sh.bin_start[kCls * s]) continue;
        strip_spots(s, false, y0, y1);
        lds_sync();
        const int cnt = (y1 - y0) * W;
        const int pix0 = y0 * W;
        for (int i = tid; i < cnt; i += kAdjThreads) {
          const aacc_t v = s_acc[i];
          s_acc[i] = 0;
          const long long kv = (static_cast<long long>(max(v, 0)) << kPixBits) |
                               static_cast<long long>((1 << kPixBits) - 1 - (pix0 + i));
          key = kv > key ? kv : key;
        }
        lds_sync();
      }
      key = wave_max_acc(key);
      if (lane == 0) sh.wkey[wid] = key;
      lds_sync();
      key = 0;
      for (int w = 0; w < kAdjWaves; ++w) key = sh.wkey[w] > key ? sh.wkey[w] : key;
    }
    const long long m_fixed = key >> kPixBits;
    const int pstar = (1 << kPixBits) - 1 - static_cast<int>(key & ((1 << kPixBits) - 1));
    // a zero maximum leaves the pattern unnormalised (the reference returns the
    // zeros as they are)
    const bool norm = a.normalize && m_fixed > 0;
    const float out_scale = norm ? 1.0f / static_cast<float>(m_fixed) : 1.0f / scale;
    const float gscale = scale * out_scale;  // dL/dQ = G / m  (1 without normalisation)

    // ---- 4. pass B: pattern, loss, upstream gradient, gather ----------------
    double lsum = 0.0, csum = 0.0;
    const float *tgt = a.target ? a.target + (size_t)p * a.target_stride : nullptr;
    float *pat = a.pattern ? a.pattern + (size_t)p * HW : nullptr;
    const bool l2 = a.mode == DSX_ADJ_L2;
    const bool fwd_only = a.mode == DSX_ADJ_FORWARD;
    // the target of strip s + 1 is loaded into registers while strip s is
    // gathered and strip s + 1 splatted, so its HBM latency is hidden
    constexpr int kPre = kAdjPre;
    native_float4 tp[kPre];
    auto prefetch = [&](int s) {
      if (kLarge || fwd_only || (W & 3) != 0 || s >= a.n_strips) return;
      const int y0 = s * a.strip_rows;
      const int cnt4 = (min(H, y0 + a.strip_rows) - y0) * W / 4;
      const native_float4 *src = reinterpret_cast<const native_float4 *>(tgt + (size_t)y0 * W);
#pragma unroll
      for (int k = 0; k < kPre; ++k) {
        const int i = tid + k * kAdjThreads;
        if (i < cnt4) tp[k] = __builtin_nontemporal_load(src + i);
      }
    };
    if (large) {
      __syncthreads();  // q complete (each thread reads back what it wrote)
      for (int s = 0; s < a.n_strips; ++s) {
        const int y0 = s * a.strip_rows;
        const int y1 = min(H, y0 + a.strip_rows);
        const int cnt = (y1 - y0) * W;
        const bool any = sh.bin_start[kCls * (s + 1)] > sh.bin_start[kCls * s];
        const size_t off = (size_t)y0 * W;
        // one pixel: pattern, loss and G / m (q may be the pattern buffer: it
        // is read before the same thread overwrites it)
        auto pixel = [&](float P, float tv, float &g) {
          if (l2) {
            const float d = P - tv;
            lsum += static_cast<double>(d * d);
            g = 2.0f * d;
          } else {
            g = tv;
          }
          csum += static_cast<double>(g * P);
        };
        if ((W & 3) == 0) {
          for (int i = tid; i < cnt / 4; i += kAdjThreads) {
            const int4 u = any ? reinterpret_cast<const int4 *>(q + off)[i] : make_int4(0, 0, 0, 0);
            const float4 P = make_float4(static_cast<float>(u.x) * out_scale, static_cast<float>(u.y) * out_scale,
                                         static_cast<float>(u.z) * out_scale, static_cast<float>(u.w) * out_scale);
            if (pat) store_nt(P, reinterpret_cast<float4 *>(pat + off) + i);
            if (fwd_only) continue;
            const native_float4 tv =
                __builtin_nontemporal_load(reinterpret_cast<const native_float4 *>(tgt + off) + i);
            float4 g;
            pixel(P.x, tv.x, g.x);
            pixel(P.y, tv.y, g.y);
            pixel(P.z, tv.z, g.z);
            pixel(P.w, tv.w, g.w);
            reinterpret_cast<float4 *>(s_G)[i] =
                make_float4(g.x * gscale, g.y * gscale, g.z * gscale, g.w * gscale);
          }
        } else {
          for (int i = tid; i < cnt; i += kAdjThreads) {
            const float P = static_cast<float>(any ? q[off + i] : 0) * out_scale;
            if (pat) __builtin_nontemporal_store(P, pat + off + i);
            if (fwd_only) continue;
            float g;
            pixel(P, tgt[off + i], g);
            s_G[i] = g * gscale;
          }
        }
        if (fwd_only) continue;
        lds_sync();
        if (a.want_grad && any) strip_spots(s, true, y0, y1);
        lds_sync();
      }
      // G leaves the shared strip buffer; the next pattern's tile bounds and
      // splat start from zero (ordered by the barrier that ends the pattern)
      if (!fwd_only)
        for (int i = tid; i < a.acc_words; i += kAdjThreads) s_acc[i] = 0;
    }
    prefetch(0);
    for (int s = 0; s < a.n_strips && !large; ++s) {
      const int y0 = s * a.strip_rows;
      const int y1 = min(H, y0 + a.strip_rows);
      const int cnt = (y1 - y0) * W;
      const bool any = sh.bin_start[kCls * (s + 1)] > sh.bin_start[kCls * s];
      if (any) {
        strip_spots(s, false, y0, y1);
        lds_sync();
      }
      const size_t off = (size_t)y0 * W;
      if ((W & 3) == 0) {
        // four pixels i: pattern, loss and G / m from the target values tv
        auto readout4 = [&](int i, const native_float4 &tv) {
          int4 *pv = reinterpret_cast<int4 *>(s_acc) + i;
          const int4 u = *pv;
          if (any) *pv = make_int4(0, 0, 0, 0);
          const float4 P = make_float4(static_cast<float>(u.x) * out_scale, static_cast<float>(u.y) * out_scale,
                                       static_cast<float>(u.z) * out_scale, static_cast<float>(u.w) * out_scale);
          if (pat) store_nt(P, reinterpret_cast<float4 *>(pat + off) + i);
          if (fwd_only) return;
          float4 g;
          if (l2) {
            const float4 d = make_float4(P.x - tv.x, P.y - tv.y, P.z - tv.z, P.w - tv.w);
            lsum += static_cast<double>(d.x * d.x + d.y * d.y + d.z * d.z + d.w * d.w);
            g = make_float4(2.0f * d.x, 2.0f * d.y, 2.0f * d.z, 2.0f * d.w);
          } else {
            g = make_float4(tv.x, tv.y, tv.z, tv.w);
          }
          csum += static_cast<double>(g.x * P.x + g.y * P.y + g.z * P.z + g.w * P.w);
          reinterpret_cast<float4 *>(s_G)[i] =
              make_float4(g.x * gscale, g.y * gscale, g.z * gscale, g.w * gscale);
        };
        const int n4 = cnt / 4;
        if (kLarge || fwd_only) {
          const native_float4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
          for (int i = tid; i < n4; i += kAdjThreads)
            readout4(i, fwd_only ? zero
                                 : __builtin_nontemporal_load(
                                       reinterpret_cast<const native_float4 *>(tgt + off) + i));
        } else {
#pragma unroll
          for (int k = 0; k < kPre; ++k) {
            const int i = tid + k * kAdjThreads;
            if (i < n4) readout4(i, tp[k]);
          }
        }
      } else {
        for (int i = tid; i < cnt; i += kAdjThreads) {
          const aacc_t v = s_acc[i];
          if (any) s_acc[i] = 0;
          const float P = static_cast<float>(v) * out_scale;
          if (pat) __builtin_nontemporal_store(P, pat + off + i);
          if (fwd_only) continue;
          const float tv = tgt[off + i];
          float g;
          if (l2) {
            const float d = P - tv;
            lsum += static_cast<double>(d * d);
            g = 2.0f * d;
          } else {
            g = tv;
          }
          csum += static_cast<double>(g * P);
          s_G[i] = g * gscale;
        }
      }
      prefetch(s + 1);
      lds_sync();
      if (a.want_grad && any) strip_spots(s, true, y0, y1);
      lds_sync();
    }
    lsum = wave_sum_d(lsum);
    csum = wave_sum_d(csum);
    if (lane == 0) {
      sh.red[wid][0] = lsum;
      sh.red[wid][1] = csum;
    }
    lds_sync();
    double ctot = 0.0, ltot = 0.0;
    for (int w = 0; w < kAdjWaves; ++w) {
      ltot += sh.red[w][0];
      ctot += sh.red[w][1];
    }

    if (a.want_grad) {
      // ---- 5. per spot: the normalisation's max term, then the chain rule ----
      const float corr = norm ? static_cast<float>(ctot) * gscale : 0.0f;
      const int px = pstar % W, py = pstar / W;
      if (!kLarge) {
        // sums with the max term to gsum; spot_chain_kernel does the rest
        float4 *gs = a.gsum + (size_t)p * a.slots;
        for (int q = tid; q < ns; q += kAdjThreads) {
          float A = t.A[q], BX = t.BX[q], BY = t.BY[q];
          if (norm)
            max_term(a, corr, px, py, t.X[q], t.Y[q], t.xb[q] >> 16, t.xb[q] & 0xffff, t.ylo[q],
                     t.yhi[q], A, BX, BY);
          gs[q] = make_float4(A, BX, BY, __int_as_float(t.jpos[q]));
        }
        if (tid == 0) a.gcount[p] = ns;
      } else {
        // the max term of each spot whose box holds p*, into its gsum entry;
        // spot_chain_kernel applies the chain rule to the whole row
        if (norm) {
          for (int j = tid; j < n; j += kAdjThreads) {
            const int2 bx = wbox[j];
            const int lo = bx.y >> 16, hi = bx.y & 0xffff, xlo = bx.x >> 16, xhi = bx.x & 0xffff;
            if (py < lo || py > hi || px < xlo || px > xhi) continue;
            const float4 sp = spots[j];
            float X, Y;
            to_pixel(sp, a, X, Y);
            float A = 0.0f, BX = 0.0f, BY = 0.0f;
            max_term(a, corr, px, py, X, Y, xlo, xhi, lo, hi, A, BX, BY);
            float *g = reinterpret_cast<float *>(gsp + wrow[j]);
            unsafeAtomicAdd(g, A);
            unsafeAtomicAdd(g + 1, BX);
            unsafeAtomicAdd(g + 2, BY);
          }
        }
        if (tid == 0) a.gcount[p] = ns_all;
      }
    }
    if (tid == 0 && a.loss) a.loss[p] = static_cast<float>(ltot);
    if (kLarge)
      __syncthreads();  // LDS and workspace state of this pattern is dead before the next one
    else
      lds_sync();  // LDS state of this pattern is dead before the next one
  }
}

// Two blocks per CU: at most 128 VGPRs (four waves per SIMD).
__global__ __launch_bounds__(kAdjThreads, 4) void render_adjoint_kernel(AdjArgs a) {
  extern __shared__ __align__(16) aacc_t s_acc[];  // 16-byte aligned for the 16-byte read-out
  __shared__ AdjShared<1> sh;
  adjoint_body<false>(a, s_acc, sh);
}

// Two blocks per CU, each with half the LDS (at most 128 VGPRs).
__global__ __launch_bounds__(kAdjThreads, 4) void render_adjoint_large_kernel(AdjArgs a) {
  extern __shared__ __align__(16) aacc_t s_acc[];
  __shared__ AdjShared<kWidthClasses> sh;
  adjoint_body<true>(a, s_acc, sh);
}

// Forward mode of the large-pattern kernel (the sub-pixel renderer of long
// spot lists): with the mode and the missing gradient outputs known at
// compile time the gather and the FP64 chain rule drop out, and the kernel
// needs about 100 VGPRs.
__global__ __launch_bounds__(kAdjThreads, 4) void render_adjoint_large_fwd_kernel(AdjArgs a) {
  extern __shared__ __align__(16) aacc_t s_acc[];
  __shared__ AdjShared<kWidthClasses> sh;
  a.mode = DSX_ADJ_FORWARD;
  a.want_grad = 0;
  a.target = nullptr;
  a.dRt = nullptr;
  a.dF2 = nullptr;
  a.drec = nullptr;
  a.loss = nullptr;
  a.gsum = nullptr;
  adjoint_body<true>(a, s_acc, sh);
}

// Chain rule of every spot gathered by the render adjoint: one block per
// pattern (grid-stride), its spots spread over the block's threads; dL/dRt of
// the pattern is a block sum, dL/d|F|^2 of the reflections global atomics,
// dL/drec a per-block partial summed later in a fixed order.
constexpr int kChainThreads = 256;
constexpr int kChainWaves = kChainThreads / kWave;

__global__ __launch_bounds__(kChainThreads) void spot_chain_kernel(AdjArgs a) {
  __shared__ ChainConsts s_cc;
  __shared__ double s_R[9];
  __shared__ double s_part[kChainWaves][9];
  if (threadIdx.x == 0) stage_chain_consts(a, s_cc);
  const int lane = lane_id();
  const int wid = threadIdx.x / kWave;
  double dG[9];  // dL/drec over this thread's spots
#pragma unroll
  for (int i = 0; i < 9; ++i) dG[i] = 0.0;
  for (int p = blockIdx.x; p < a.n; p += gridDim.x) {
    if (threadIdx.x < 9) s_R[threadIdx.x] = a.Rt[(size_t)p * 9 + threadIdx.x];
    __syncthreads();
    double R[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) R[i] = s_R[i];
    double dR[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) dR[i] = 0.0;
    const int ns = a.gcount[p];
    const float4 *gs = a.gsum + (size_t)p * a.slots;
    const int *idx = a.idx + (size_t)p * a.cap;
    const float4 *spots = a.spots + (size_t)p * a.cap;
    for (int q = threadIdx.x; q < ns; q += kChainThreads) {
      const float4 v = gs[q];
      const int j = __float_as_int(v.w);
      spot_chain(a, s_cc, R, idx[j], spots[j].w, v.x, v.y, v.z, dR, dG);
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) dR[i] = wave_sum_d(dR[i]);
    if (lane == 0) {
#pragma unroll
      for (int i = 0; i < 9; ++i) s_part[wid][i] = dR[i];
    }
    __syncthreads();
    if (threadIdx.x < 9) {
      double v = 0.0;
      for (int w = 0; w < kChainWaves; ++w) v += s_part[w][threadIdx.x];
      a.dRt[(size_t)p * 9 + threadIdx.x] = v;
    }
    __syncthreads();  // s_R and s_part are rewritten for the next pattern
  }
  if (!a.drec) return;  // uniform
  // this block's partial dL/drec (summed in a fixed order by sum_partials_kernel)
#pragma unroll
  for (int i = 0; i < 9; ++i) dG[i] = wave_sum_d(dG[i]);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < 9; ++i) s_part[wid][i] = dG[i];
  }
  __syncthreads();
  if (threadIdx.x < 9) {
    double v = 0.0;
    for (int w = 0; w < kChainWaves; ++w) v += s_part[w][threadIdx.x];
    a.dpart[(size_t)blockIdx.x * 9 + threadIdx.x] = v;
  }
}

// drec[i] += sum over blocks of part[b][i]; one wave per entry.
__global__ __launch_bounds__(9 * kWave) void sum_partials_kernel(const double *part, int nb,
                                                                 double *drec) {
  const int i = threadIdx.x / kWave, lane = lane_id();
  double v = 0.0;
  for (int b = lane; b < nb; b += kWave) v += part[(size_t)b * 9 + i];
  v = wave_sum_d(v);
  if (lane == 0) drec[i] += v;
}

// ---------------------------------------------------------------------------
// structure factor adjoint
// ---------------------------------------------------------------------------

constexpr int kSfaThreads = 256;
constexpr int kSfaWaves = kSfaThreads / kWave;

struct SfAdjArgs {
  const int4 *hkl;
  int n;
  double rec[9];
  const uint4 *xyz_fx;
  const int *elem;
  const float *occ;
  const float *B;
  int n_atoms;
  const double *coeff;
  int n_elem;
  int scatter;
  const float *dF2;
  float *dxyz;
  float *dB;
  double *drec;
};

__device__ __forceinline__ float turns_of(uint32_t uh, uint32_t uk, uint32_t ul, const uint4 &x) {
  const uint32_t t = uh * x.x + uk * x.y + ul * x.z;
  return static_cast<float>(static_cast<int32_t>(t)) * 2.3283064365386963e-10f;
}

__global__ __launch_bounds__(kSfaThreads) void structure_factor_adjoint_kernel(SfAdjArgs a) {
  __shared__ uint4 s_xyz[kSfaThreads];
  __shared__ float s_occ[kSfaThreads];
  __shared__ float s_B[kSfaThreads];
  __shared__ int s_el[kSfaThreads];
  __shared__ float s_g[4][kSfaThreads];  // per-atom sums of the chunk: dx, dy, dz, dB
  __shared__ double s_rec[kSfaWaves][9];

  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int wid = tid / kWave;
  const int i = blockIdx.x * kSfaThreads + tid;
  const bool valid = i < a.n;
  const int4 hkl = valid ? a.hkl[i] : make_int4(0, 0, 0, 0);
  const double w = valid ? static_cast<double>(a.dF2[i]) : 0.0;
  const double hx = hkl.x, hy = hkl.y, hz = hkl.z;
  const double gx = a.rec[0] * hx + a.rec[1] * hy + a.rec[2] * hz;
  const double gy = a.rec[3] * hx + a.rec[4] * hy + a.rec[5] * hz;
  const double gz = a.rec[6] * hx + a.rec[7] * hy + a.rec[8] * hz;
  const double g2 = gx * gx + gy * gy + gz * gz;
  const float mq = -0.25f * static_cast<float>(g2);
  const uint32_t uh = static_cast<uint32_t>(hkl.x), uk = static_cast<uint32_t>(hkl.y),
                 ul = static_cast<uint32_t>(hkl.z);

  // pass 1: F and dF/d(g^2)
  double Fr = 0.0, Fi = 0.0, Qr = 0.0, Qi = 0.0;
  int cur_el = -1;
  double f = 0.0, df = 0.0;
  for (int base = 0; base < a.n_atoms; base += kSfaThreads) {
    __syncthreads();
    const int j = base + tid;
    if (j < a.n_atoms) {
      s_xyz[tid] = a.xyz_fx[j];
      s_occ[tid] = a.occ[j];
      s_B[tid] = a.B[j];
      s_el[tid] = a.elem[j];
    }
    __syncthreads();
    const int m = min(kSfaThreads, a.n_atoms - base);
    for (int jj = 0; jj < m; ++jj) {
      const int el = min(max(s_el[jj], 0), a.n_elem - 1);
      if (el != cur_el) {
        f = atomic_f_grad(a.coeff + 10 * el, g2, a.scatter, df);
        cur_el = el;
      }
      float sn, cs;
      sincospif(2.0f * turns_of(uh, uk, ul, s_xyz[jj]), &sn, &cs);
      const double od = static_cast<double>(s_occ[jj] * expf(mq * s_B[jj]));
      const double amp = f * od;
      // d/dq [f occ exp(-B q/4)] = (f' - f B/4) occ exp(-B q/4)
      const double damp = (df - 0.25 * f * static_cast<double>(s_B[jj])) * od;
      Fr = fma(amp, static_cast<double>(cs), Fr);
      Fi = fma(amp, static_cast<double>(sn), Fi);
      Qr = fma(damp, static_cast<double>(cs), Qr);
      Qi = fma(damp, static_cast<double>(sn), Qi);
    }
  }
  // through |F|^2: dL/dg = 2 g dL/d(g^2), dL/d(g^2) = w 2 Re(conj(F) dF/dq)
  const double dq = w * 2.0 * (Fr * Qr + Fi * Qi);
  const double dgx = 2.0 * gx * dq;
  const double dgy = 2.0 * gy * dq;
  const double dgz = 2.0 * gz * dq;
  double dr[9] = {dgx * hx, dgx * hy, dgx * hz, dgy * hx, dgy * hy, dgy * hz,
                  dgz * hx, dgz * hy, dgz * hz};
#pragma unroll
  for (int t = 0; t < 9; ++t) dr[t] = wave_sum_d(dr[t]);
  if (lane == 0) {
#pragma unroll
    for (int t = 0; t < 9; ++t) s_rec[wid][t] = dr[t];
  }

  // pass 2: per-atom terms, dF2/dx_a = -4 pi h Im(conj(F) t_a),
  // dF2/dB_a = -(g^2/2) Re(conj(F) t_a)
  const bool any = __ballot(w != 0.0) != 0;
  cur_el = -1;
  for (int base = 0; base < a.n_atoms; base += kSfaThreads) {
    __syncthreads();
    const int j = base + tid;
    if (j < a.n_atoms) {
      s_xyz[tid] = a.xyz_fx[j];
      s_occ[tid] = a.occ[j];
      s_B[tid] = a.B[j];
      s_el[tid] = a.elem[j];
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) s_g[c][tid] = 0.0f;
    __syncthreads();
    const int m = min(kSfaThreads, a.n_atoms - base);
    if (any) {
      for (int jj = 0; jj < m; ++jj) {
        const int el = min(max(s_el[jj], 0), a.n_elem - 1);
        if (el != cur_el) {
          f = atomic_f(a.coeff + 10 * el, g2, a.scatter);
          cur_el = el;
        }
        float sn, cs;
        sincospif(2.0f * turns_of(uh, uk, ul, s_xyz[jj]), &sn, &cs);
        const double amp = f * static_cast<double>(s_occ[jj] * expf(mq * s_B[jj]));
        const double tr = amp * cs, ti = amp * sn;
        const double im = Fr * ti - Fi * tr;
        const double re = Fr * tr + Fi * ti;
        const float ci = static_cast<float>(-4.0 * M_PI * w * im);
        float vx = ci * static_cast<float>(hx), vy = ci * static_cast<float>(hy),
              vz = ci * static_cast<float>(hz);
        float vb = static_cast<float>(-0.5 * g2 * w * re);
        vx = wave_sum_f(vx);
        vy = wave_sum_f(vy);
        vz = wave_sum_f(vz);
        vb = wave_sum_f(vb);
        if (lane == 0) {
          atomicAdd(&s_g[0][jj], vx);
          atomicAdd(&s_g[1][jj], vy);
          atomicAdd(&s_g[2][jj], vz);
          atomicAdd(&s_g[3][jj], vb);
        }
      }
    }
    __syncthreads();
    if (tid < m) {
      const int atom = base + tid;
      if (s_g[0][tid] != 0.0f) unsafeAtomicAdd(&a.dxyz[3 * atom + 0], s_g[0][tid]);
      if (s_g[1][tid] != 0.0f) unsafeAtomicAdd(&a.dxyz[3 * atom + 1], s_g[1][tid]);
      if (s_g[2][tid] != 0.0f) unsafeAtomicAdd(&a.dxyz[3 * atom + 2], s_g[2][tid]);
      if (s_g[3][tid] != 0.0f) unsafeAtomicAdd(&a.dB[atom], s_g[3][tid]);
    }
  }
  __syncthreads();
  if (tid < 9) {
    double v = 0.0;
    for (int ww = 0; ww < kSfaWaves; ++ww) v += s_rec[ww][tid];
    if (v != 0.0) unsafeAtomicAdd(&a.drec[tid], v);
  }
}

}  // namespace dsx

using namespace dsx;

// Workspace of the large-pattern kernel, per block (E = kWorkEntriesPerSpot *
// cap bin entries): int bins[E] | int2 boxes[cap] | int rows[cap] | int
// q[H W], each part padded to 16 bytes; with gradients the chain-rule hand-over follows the
// blocks.  Chain-rule hand-over (the small-pattern kernel's whole workspace):
// float4 gsum[n * cap] | int gcount[n] | (8-byte aligned) double
// dpart[chain blocks][9].
constexpr int kWorkEntriesPerSpot = 4;
constexpr int kAdjLdsBytes = 160 * 1024;

static bool adjoint_large(int cap) { return cap > kAdjMaxSpots; }
static int chain_blocks_max() { return cu_count() * 8; }
static long long small_work_ints(int cap, int n) {
  return 4ll * n * cap + n + 1 + 18ll * chain_blocks_max();
}
static long long large_block_ints(int cap, long long hw) {
  const long long e = kWorkEntriesPerSpot * static_cast<long long>(cap);
  return ((e + 3) & ~3ll) + ((2ll * cap + 3) & ~3ll) + ((cap + 3ll) & ~3ll) + ((hw + 3) & ~3ll);
}
// blocks of the large-pattern kernels: two per CU
static int large_blocks(int n, int per_cu = 2) { return std::min(n, per_cu * cu_count()); }

extern "C" long long dsx_render_adjoint_workspace_ints(int cap, int n, int height, int width,
                                                       int gradients) {
  if (cap < 1 || n < 0 || height < 1 || width < 1) return 0;
  if (adjoint_large(cap))
    return static_cast<long long>(large_blocks(n)) *
               large_block_ints(cap, static_cast<long long>(height) * width) +
           (gradients ? small_work_ints(cap, n) : 0);
  return small_work_ints(cap, n);
}

// Static LDS of a kernel (queried once) and its dynamic LDS limit raised to
// what is left of the CU's 160 KB.
static int adjoint_static_lds(const void *kernel, int *cache, int *raised) {
  if (*cache < 0) {
    hipFuncAttributes attr;
    if (hipFuncGetAttributes(&attr, kernel) != hipSuccess) {
      (void)hipGetLastError();
      return -1;
    }
    *cache = static_cast<int>(attr.sharedSizeBytes);
  }
  if (!*raised) {
    if (hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            kAdjLdsBytes - *cache) != hipSuccess) {
      (void)hipGetLastError();
      return -1;
    }
    *raised = 1;
  }
  return *cache;
}

extern "C" int dsx_render_adjoint(const int *idx_d, const float *spot_d, const int *count_d, int cap,
                                  int n, const int *hkl_d, const float *F2_d, int n_refl,
                                  const double *Rt_d, const dsx_sim_params *sim,
                                  const dsx_render_params *rp, int mode, const float *target_d,
                                  long long target_stride, float *pattern_d, float *loss_d,
                                  double *dRt_d, float *dF2_d, double *drec_d, int *work_d,
                                  long long work_ints,
                                  int *status_d, void *stream) {
  clear_error();
  if (!sim || !rp || n < 0 || cap < 1 || n_refl < 0 || work_ints < 0 || target_stride < 0) {
    set_error("dsx_render_adjoint: bad arguments (n=%d cap=%d n_refl=%d)", n, cap, n_refl);
    return DSX_EINVAL;
  }
  if (mode != DSX_ADJ_L2 && mode != DSX_ADJ_PATTERN_GRAD && mode != DSX_ADJ_FORWARD) {
    set_error("dsx_render_adjoint: unknown mode %d", mode);
    return DSX_EINVAL;
  }
  if (n == 0) return DSX_OK;
  const bool want_grad = dRt_d != nullptr;
  if (!want_grad && (dF2_d || drec_d)) {
    set_error("dsx_render_adjoint: dF2_d and drec_d need dRt_d");
    return DSX_EINVAL;
  }
  if (mode == DSX_ADJ_FORWARD && (want_grad || !pattern_d)) {
    set_error("dsx_render_adjoint: forward mode writes patterns and no gradients");
    return DSX_EINVAL;
  }
  if (!spot_d || !count_d || (mode != DSX_ADJ_FORWARD && !target_d) ||
      (want_grad && (!idx_d || !Rt_d)) || (want_grad && n_refl > 0 && (!hkl_d || !F2_d))) {
    set_error("dsx_render_adjoint: null device pointer");
    return DSX_EINVAL;
  }
  if (rp->height < 1 || rp->width < 1 || rp->height > 32767 || rp->width > 32767 ||
      static_cast<long long>(rp->height) * rp->width > (1ll << kPixBits) ||
      !(rp->calibration > 0.0f) || !(rp->sigma > 0.0f) || !(rp->clip_threshold > 0.0f)) {
    set_error("dsx_render_adjoint: bad pattern shape %dx%d (at most 2^22 pixels), calibration, "
              "sigma or clip_threshold", rp->height, rp->width);
    return DSX_EINVAL;
  }
  if (sim->shape_model < DSX_SF_LINEAR || sim->shape_model > DSX_SF_LORENTZIAN_PRECESSION ||
      !(sim->k > 0.0) || sim->shape_width == 0.0) {
    set_error("dsx_render_adjoint: bad physical parameters");
    return DSX_EINVAL;
  }
  AdjArgs a;
  a.idx = idx_d;
  a.spots = reinterpret_cast<const float4 *>(spot_d);
  a.count = count_d;
  a.cap = cap;
  a.n = n;
  a.hkl = reinterpret_cast<const int4 *>(hkl_d);
  a.F2 = F2_d;
  a.n_refl = n_refl;
  a.Rt = Rt_d;
  for (int i = 0; i < 9; ++i) a.rec[i] = sim->rec[i];
  fill_physics(a, sim);
  a.H = rp->height;
  a.W = rp->width;
  a.cal = rp->calibration;
  a.cx = rp->center_x;
  a.cy = rp->center_y;
  a.ca = rp->cos_angle;
  a.sa = rp->sin_angle;
  a.mirrored = rp->mirrored;
  a.normalize = rp->normalize;
  const double sg = rp->sigma;
  a.pref = static_cast<float>(1.0 / (2.0 * M_PI * sg * sg));
  a.ep = static_cast<float>(-1.0 / (2.0 * sg * sg));
  a.inv_s2 = static_cast<float>(1.0 / (sg * sg));
  a.e2ep = static_cast<float>(exp(-1.0 / (sg * sg)));
  a.clip = rp->clip_threshold;
  a.mode = mode;
  a.want_grad = want_grad ? 1 : 0;
  a.target = target_d;
  a.target_stride = target_stride;
  a.pattern = pattern_d;
  a.loss = loss_d;
  a.dRt = dRt_d;
  a.dF2 = dF2_d;
  a.drec = drec_d;
  a.dpart = nullptr;
  a.status = status_d;
  a.work = nullptr;
  a.work_per_block = 0;
  a.work_entries = 0;
  a.gsum = nullptr;
  a.gcount = nullptr;

  // LDS per block: the static part, the spot table (slots entries; bins) and
  // two strip buffers of rows x W words.  Patterns up to kAdjMaxSpots spots
  // use two blocks per CU when their strips hold at least 8 rows (or the whole
  // pattern); larger lists use the large-pattern kernel, one block per CU.
  const bool large = adjoint_large(cap);
  static int static_small = -1, static_large = -1, static_fwd = -1;
  static int raised_small = 0, raised_large = 0, raised_fwd = 0;
  const bool fwd_large = large && mode == DSX_ADJ_FORWARD;
  const void *kern = fwd_large ? reinterpret_cast<const void *>(&render_adjoint_large_fwd_kernel)
                     : large   ? reinterpret_cast<const void *>(&render_adjoint_large_kernel)
                               : reinterpret_cast<const void *>(&render_adjoint_kernel);
  const int st = fwd_large ? adjoint_static_lds(kern, &static_fwd, &raised_fwd)
                 : large   ? adjoint_static_lds(kern, &static_large, &raised_large)
                           : adjoint_static_lds(kern, &static_small, &raised_small);
  if (st < 0) {
    set_error("dsx_render_adjoint: cannot query or reserve the kernel's LDS");
    return DSX_ELAUNCH;
  }
  int bpc = 2;  // large kernels: two blocks per CU, one block's barrier waits overlap the other's work
  int rows = 0;
  long long table = 0;
  if (large) {
    // one strip buffer of up to 64 KB (32 KB with two blocks per CU), at least
    // 512 spot slots, the rest slots; one block per CU when a row of the
    // pattern does not fit half the budget
    a.bin_cap = 0;
    auto size_large = [&](int b) {
      const long long per_block = kAdjLdsBytes / b;
      const long long strip = std::min<long long>(kAdjStripBytes / b,
                                                  per_block - st - 16 - 512ll * kSpotTabBytes);
      rows = strip < 4ll * a.W ? 0 : static_cast<int>(std::min<long long>(a.H, strip / (4ll * a.W)));
      a.slots = static_cast<int>(std::min<long long>(
          kAdjLargeMaxSlots, (per_block - st - 16 - 4ll * rows * a.W) / kSpotTabBytes));
      table = static_cast<long long>(kSpotTabBytes) * a.slots + 16;
    };
    size_large(bpc);
    if (rows < 1) {
      bpc = 1;
      size_large(bpc);
    }
  } else {
    a.slots = cap;
    a.bin_cap = std::min(kAdjBinCap, 4 * cap + 256);
    table = static_cast<long long>(kSpotTabBytes) * a.slots + 2ll * a.bin_cap + 16;
    auto rows_for = [&](int per_block) {
      const long long strip = std::min<long long>(kAdjSmallStripBytes, per_block - st - table);
      return strip < 8ll * a.W ? 0 : static_cast<int>(std::min<long long>(a.H, strip / (8ll * a.W)));
    };
    rows = rows_for(kAdjLdsBytes / bpc);
    if (rows < std::min(a.H, 8)) {
      bpc = 1;
      rows = rows_for(kAdjLdsBytes);
    }
  }
  if (rows < 1) {
    set_error("dsx_render_adjoint: pattern width %d exceeds the LDS strip budget", a.W);
    return DSX_EINVAL;
  }
  const int n_strips = (a.H + rows - 1) / rows;
  if (n_strips > kAdjMaxStrips) {
    set_error("dsx_render_adjoint: pattern height %d needs %d strips (> %d)", a.H, n_strips,
              kAdjMaxStrips);
    return DSX_EINVAL;
  }
  a.strip_rows = rows;
  a.n_strips = n_strips;
  a.acc_words = rows * a.W;
  const int blocks = large ? large_blocks(n, bpc) : std::min(n, cu_count() * bpc);
  if (large) {
    const long long per_block = large_block_ints(cap, static_cast<long long>(a.H) * a.W);
    const long long need = per_block * blocks + (want_grad ? small_work_ints(cap, n) : 0);
    if (work_d && reinterpret_cast<uintptr_t>(work_d) % 16 != 0) {
      set_error("dsx_render_adjoint: workspace must be 16-byte aligned");
      return DSX_EINVAL;
    }
    if (want_grad && !work_d) {
      set_error("dsx_render_adjoint: gradients need a workspace of %lld ints "
                "(dsx_render_adjoint_workspace_ints)", need);
      return DSX_EINVAL;
    }
    if (work_d && work_ints >= need &&
        kWorkEntriesPerSpot * static_cast<long long>(cap) < (1ll << 31)) {
      a.work = work_d;
      a.work_per_block = per_block;
      a.work_entries = kWorkEntriesPerSpot * cap;
      if (want_grad) {
        int *hand = work_d + per_block * blocks;  // a multiple of 4 ints: 16-byte aligned
        a.gsum = reinterpret_cast<float4 *>(hand);
        a.gcount = hand + static_cast<size_t>(n) * 4 * cap;
        const size_t off = (static_cast<size_t>(n) * (4ull * cap + 1) + 1) & ~static_cast<size_t>(1);
        a.dpart = reinterpret_cast<double *>(hand + off);
      }
    } else if (work_d) {
      set_error("dsx_render_adjoint: workspace of %lld ints is smaller than %lld "
                "(dsx_render_adjoint_workspace_ints)", work_ints, need);
      return DSX_EINVAL;
    }
  } else if (want_grad) {
    const long long need = small_work_ints(cap, n);
    if (!work_d || work_ints < need) {
      set_error("dsx_render_adjoint: gradients need a workspace of %lld ints, got %lld "
                "(dsx_render_adjoint_workspace_ints)", need, work_d ? work_ints : 0ll);
      return DSX_EINVAL;
    }
    if (reinterpret_cast<uintptr_t>(work_d) % 16 != 0) {
      set_error("dsx_render_adjoint: workspace must be 16-byte aligned");
      return DSX_EINVAL;
    }
    a.gsum = reinterpret_cast<float4 *>(work_d);
    a.gcount = work_d + static_cast<size_t>(n) * 4 * cap;
    const size_t off = (static_cast<size_t>(n) * (4ull * cap + 1) + 1) & ~static_cast<size_t>(1);
    a.dpart = reinterpret_cast<double *>(work_d + off);
  }
  const size_t lds = (((large ? 1ull : 2ull) * a.acc_words * 4 + static_cast<size_t>(table)) + 15) &
                     ~static_cast<size_t>(15);
  hipStream_t s = as_stream(stream);
  if (large) {
    if (fwd_large)
      hipLaunchKernelGGL(render_adjoint_large_fwd_kernel, dim3(blocks), dim3(kAdjThreads), lds, s, a);
    else
      hipLaunchKernelGGL(render_adjoint_large_kernel, dim3(blocks), dim3(kAdjThreads), lds, s, a);
  } else {
    hipLaunchKernelGGL(render_adjoint_kernel, dim3(blocks), dim3(kAdjThreads), lds, s, a);
  }
  int rc = check_launch("dsx_render_adjoint");
  if (rc != DSX_OK || !want_grad) return rc;
  // the chain kernel reads gsum rows of `slots` entries: the large kernel's
  // rows are indexed by spot position (cap entries)
  if (large) a.slots = cap;
  const int cblocks = std::min(n, chain_blocks_max());
  hipLaunchKernelGGL(spot_chain_kernel, dim3(cblocks), dim3(kChainThreads), 0, s, a);
  if (drec_d)
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(9 * kWave), 0, s, a.dpart, cblocks, drec_d);
  return check_launch("dsx_render_adjoint (chain)");
}

// ---------------------------------------------------------------------------
// Allocation-free form of dsx_render_adjoint: the workspace is a file-scope
// device array (one copy per GPU, held until the process ends).  Host code
// never passes the array's name to a kernel (on the host it is only a
// stand-in); its device address is looked up once per device with
// hipGetSymbolAddress and kept here.
// ---------------------------------------------------------------------------

constexpr long long kNoallocScratchInts = 16ll << 20;  // 64 MiB
static __device__ __align__(256) int scratch[kNoallocScratchInts];

static std::mutex g_scratch_mu;
static int *g_scratch_ptr[64] = {};

// Device address of this device's scratch (looked up on the first call).
static int scratch_address(int **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
    (void)hipGetLastError();
    set_error("dsx_render_adjoint_noalloc: no current device (or device id >= 64)");
    return DSX_ELAUNCH;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      (void)hipGetLastError();
      set_error("dsx_render_adjoint_noalloc: hipGetSymbolAddress(scratch) failed: %s",
                hipGetErrorString(e));
      return DSX_ELAUNCH;
    }
    g_scratch_ptr[dev] = static_cast<int *>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return DSX_OK;
}

extern "C" long long dsx_render_adjoint_noalloc_max_ints(void) { return kNoallocScratchInts; }

extern "C" int dsx_render_adjoint_noalloc(const int *idx_d, const float *spot_d, const int *count_d,
                                          int cap, int n, const int *hkl_d, const float *F2_d,
                                          int n_refl, const double *Rt_d, const dsx_sim_params *sim,
                                          const dsx_render_params *rp, int mode,
                                          const float *target_d, long long target_stride,
                                          float *pattern_d, float *loss_d, double *dRt_d,
                                          float *dF2_d, double *drec_d, int *status_d,
                                          void *stream) {
  clear_error();
  if (cap < 1 || n < 0 || !rp) {
    set_error("dsx_render_adjoint_noalloc: bad arguments (n=%d cap=%d)", n, cap);
    return DSX_EINVAL;
  }
  const long long need =
      dsx_render_adjoint_workspace_ints(cap, n, rp->height, rp->width, dRt_d != nullptr);
  if (need > kNoallocScratchInts) {
    set_error("dsx_render_adjoint_noalloc: %d patterns with capacity %d need %lld ints of "
              "scratch; the static scratch holds %lld (use dsx_render_adjoint with a workspace, "
              "or fewer patterns per call)", n, cap, need, kNoallocScratchInts);
    return DSX_EINVAL;
  }
  int *work = nullptr;
  const int rc = scratch_address(&work);
  if (rc != DSX_OK) return rc;
  return dsx_render_adjoint(idx_d, spot_d, count_d, cap, n, hkl_d, F2_d, n_refl, Rt_d, sim, rp, mode,
                            target_d, target_stride, pattern_d, loss_d, dRt_d, dF2_d, drec_d, work,
                            kNoallocScratchInts, status_d, stream);
}

extern "C" int dsx_structure_factor_adjoint(const int *hkl_d, int n, const double rec[9],
                                            const uint32_t *xyz_fx_d, const int *elem_d,
                                            const float *occ_d, const float *B_d, int n_atoms,
                                            const double *coeff_d, int n_elem, int scatter,
                                            const float *dF2_d, float *dxyz_d, float *dB_d,
                                            double *drec_d, void *stream) {
  clear_error();
  if (n < 0 || n_atoms < 0 || n_elem < 0 || !rec) {
    set_error("dsx_structure_factor_adjoint: bad sizes n=%d n_atoms=%d n_elem=%d", n, n_atoms, n_elem);
    return DSX_EINVAL;
  }
  if (scatter != DSX_SCATTER_NONE && scatter != DSX_SCATTER_LOBATO && scatter != DSX_SCATTER_XTABLES) {
    set_error("dsx_structure_factor_adjoint: unknown scattering parameterisation %d", scatter);
    return DSX_EINVAL;
  }
  if (n == 0) return DSX_OK;
  if (n_atoms > 0 && n_elem == 0) {
    set_error("dsx_structure_factor_adjoint: atoms given but no element coefficients");
    return DSX_EINVAL;
  }
  if (!hkl_d || !dF2_d || !drec_d ||
      (n_atoms > 0 && (!xyz_fx_d || !elem_d || !occ_d || !B_d || !dxyz_d || !dB_d)) ||
      (n_elem > 0 && !coeff_d)) {
    set_error("dsx_structure_factor_adjoint: null device pointer");
    return DSX_EINVAL;
  }
  SfAdjArgs a;
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
  a.dF2 = dF2_d;
  a.dxyz = dxyz_d;
  a.dB = dB_d;
  a.drec = drec_d;
  hipLaunchKernelGGL(structure_factor_adjoint_kernel, dim3((n + kSfaThreads - 1) / kSfaThreads),
                     dim3(kSfaThreads), 0, as_stream(stream), a);
  return check_launch("dsx_structure_factor_adjoint");
}
