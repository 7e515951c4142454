// (c) LDS-tiled spot renderer.
//
// fast=True follows Simulation2D.get_diffraction_pattern
// (diffsims/simulations/simulation2d.py:357-442) and the integer path of
// get_pattern_from_pixel_coordinates_and_intensities
// (diffsims/pattern/detector_functions.py:293-300): pixel coordinates are
// truncated, a later spot overwrites an earlier one on the same pixel, the image
// is filtered with scipy.ndimage.gaussian_filter (normalised taps of radius
// int(4 sigma + 0.5), 'reflect' boundaries) and divided by its maximum.  As the
// filtered image is linear in the spot image, it is built here directly as
// sum_spots I * Ky(y) * Kx(x), where K sums the taps over every reflected image
// of the source pixel.
//
// fast=False follows _subpixel_gaussian (detector_functions.py:314-359): a
// clipped box of I/(2 pi s^2) exp(-d^2 / 2 s^2) around each float coordinate.
//
// One block renders one pattern at a time (persistent over patterns):
//   1. spots -> pixel coordinates and the in-frame test; fast mode keeps the
//      last spot per pixel through an LDS hash (atomicMax of the spot's list
//      position), the hash living in the strip buffer before it is needed;
//   2. survivors are binned by the horizontal strips their footprint touches;
//   3. pass A (normalize only): each strip is splatted into LDS and reduced to
//      the pattern maximum;
//   4. pass B: each strip is splatted again and streamed to HBM scaled by
//      1/max with non-temporal 16-byte stores; strips without spots are
//      written as zeros straight from registers.
//
// Accumulation is 32-bit fixed point (ds_add_u32): LDS float atomics run ~20x
// slower than integer ones on gfx950, and 32-bit words let a strip hold twice
// the rows of 64-bit ones (half the barriers per pattern).  Before splatting,
// every spot adds the largest value its footprint can take anywhere (I times
// the largest tap along each axis; I / (2 pi s^2) for the sub-pixel Gaussian)
// to each 16x16 tile it touches; the largest tile sum U bounds every pixel, so
// with one unit = U / (1.9 * 2^30) no pixel sum can overflow.  For the usual
// sparse pattern U is a few times the pattern maximum, so one unit is ~1e-9 of
// it.
#include <algorithm>

#include "dsx_internal.h"
#include "dsx_splat.h"

namespace dsx {

constexpr int kRenderThreads = 512;
// Launch shapes of the splat renderer: at most 80 VGPRs (six waves per SIMD),
// so three 512-thread blocks fit a CU when each holds a 28 KB strip; with two
// blocks per CU the strips are 48 KB (fewer strips, fewer barriers).  Three
// blocks pay off when the normalised fast mode's maximum search leaves gaps
// in a block's HBM writes for the other blocks to fill.
constexpr int kRenderMinWaves = 6;
constexpr int kMaxRenderBlocksPerCU = 3;
constexpr int kMaxSpots = 1024;  // distinct in-frame spots per pattern
constexpr int kHash = 2048;      // LDS hash slots (load factor <= 1/2)
constexpr int kMaxStrips = 256;
constexpr int kBinCap = 4096;    // strip-bin entries per pattern
constexpr int kMaxR = 64;
// bounded search for the pattern maximum (fast mode): 16x16 pixel tiles
constexpr int kTile = 16;
constexpr int kMaxTiles = 1024;   // up to 512x512 patterns
constexpr int kMaxCand = 16;      // candidate tiles evaluated exactly
constexpr int kCandSpots = 64;    // spots touching one candidate tile
constexpr float kScaleRange = 2040109465.6f;  // 1.9 * 2^30: fixed-point units per tile bound

typedef int racc_t;  // strip accumulator word
constexpr int kStripBytes = 48 * 1024;       // one or two blocks per CU
constexpr int kStripBytes3 = 28 * 1024;      // three blocks per CU


enum { kModeFast = 0, kModeSubpixel = 1 };

// Waves per SIMD the splat kernel is compiled for.  The sub-pixel mode always
// runs two blocks per CU (four waves per SIMD), so it may use 128 VGPRs: held
// to 80 it spilled 36 bytes per lane to scratch, written and read back for
// every pattern.
template <int kMode>
constexpr int render_min_waves() {
  return kMode == kModeFast ? kRenderMinWaves : 4;
}

static int splat_blocks_per_cu(int mode, bool normalize, int requested) {
  if (mode == kModeSubpixel) return 2;  // its kernel is compiled for four waves per SIMD
  if (requested > 0) return std::min(requested, kMaxRenderBlocksPerCU);
  return mode == kModeFast && normalize ? 3 : 2;
}

static int strip_bytes_for(int blocks_per_cu) {
  return blocks_per_cu >= 3 ? kStripBytes3 : kStripBytes;
}

struct RenderArgs {
  const float4 *spots;
  const int *count;
  int cap;
  int n;                     // patterns to render (entries of list, or 0..n-1)
  const long long *spot_off; // ragged: first spot of pattern p (NULL: p * cap)
  const int *capv;           // ragged: capacity of pattern p (NULL: cap)
  const int *list;           // ragged: pattern ids to render (NULL: 0..n-1)
  const int *n_dev;          // splat kernel: pattern count in device memory (NULL: n)
  int H, W;
  float cal, cx, cy, ca, sa;
  int mirrored, normalize;
  int R;           // fast: tap radius
  float pref, ep;  // subpixel: 1/(2 pi s^2), -1/(2 s^2)
  float clip;
  int strip_rows;
  int n_strips;
  int lds_words;   // size of the dynamic LDS buffer in racc_t words
  float *out;
  int *status;
  float taps[kMaxR + 1];
};

// Fast mode: total tap weight at output index y from a delta at source c on a
// line of length n with scipy's 'reflect' extension (period 2n; images at
// c + 2n m and -1 - c + 2n m).
__device__ __forceinline__ float reflect_weight(int y, int c, int n, int R, const float *taps) {
  float w = 0.0f;
  const int period = 2 * n;
#pragma unroll
  for (int img = 0; img < 2; ++img) {
    const int base = img == 0 ? c : -1 - c;
    const int d = y - R - base;  // first image t = base + period*m >= y - R
    const int m = d >= 0 ? (d + period - 1) / period : -((-d) / period);
    for (int t = base + period * m; t <= y + R; t += period) {
      const int off = t - y;
      w += taps[off < 0 ? -off : off];
    }
  }
  return w;
}

__device__ __forceinline__ float tap_at(int d, int R, const float *taps) {
  d = d < 0 ? -d : d;
  return d <= R ? taps[d] : 0.0f;
}

// reflect_weight for R < n: only the images c, -1 - c and 2n - 1 - c can reach.
__device__ __forceinline__ float reflect_weight3(int y, int c, int n, int R, const float *taps) {
  return tap_at(c - y, R, taps) + tap_at(-1 - c - y, R, taps) + tap_at(2 * n - 1 - c - y, R, taps);
}

__device__ __forceinline__ unsigned hash_slot(int key) {
  return (static_cast<unsigned>(key) * 2654435761u) >> (32 - 11);  // kHash = 2^11
}

template <int kMode>
struct RenderShared {
  float I[kMaxSpots];
  int xy[kMaxSpots];  // fast: (yi << 16) | xi ; subpixel: (xlo << 16) | xhi
  float fx[kMode == kModeSubpixel ? kMaxSpots : 1], fy[kMode == kModeSubpixel ? kMaxSpots : 1];
  float taps[kMaxR + 1];
  short ylo[kMaxSpots], yhi[kMaxSpots];
  short bin[kBinCap];
  int bin_start[kMaxStrips + 1];
  int bin_fill[kMaxStrips];
  int ns;
  int overflow;
  int cols_max;
  float wsum[kRenderThreads / kWave];
  int wu[kRenderThreads / kWave];
  racc_t wmax[kRenderThreads / kWave];
  // bounded maximum search (fast mode); its per-tile arrays borrow the
  // (zeroed) strip buffer so the block keeps two-blocks-per-CU occupancy
  int cand_tile[kMaxCand];
  int cand_n[kMaxCand];
  int n_cand;
  int fail;
  int qstar;
  long long lfix;
};

// Fast-mode weight of a source pixel c at output index y along one axis, the
// value the splat uses (reflect_weight3 equals the interior tap exactly when
// the footprint does not reach an edge).
__device__ __forceinline__ float fast_weight(int y, int c, int n, int R, bool short_taps,
                                             const float *taps) {
  return short_taps ? reflect_weight3(y, c, n, R, taps) : reflect_weight(y, c, n, R, taps);
}

// Fast-mode fixed-point term: ((I wx) scale) wy truncated (the values are
// non-negative, so this is the floor; one conversion instruction).  The splat
// and the maximum search both use it, so they agree bit for bit.
__device__ __forceinline__ int fast_fixed(float iws, float wy) { return static_cast<int>(iws * wy); }

// Fixed-point contribution of a survivor (I at pixel xi, yi; rows lo..hi) to
// pixel (x, y), rounded exactly as the splat rounds it; 0 outside its footprint.
__device__ __forceinline__ int fast_term(float I, int xi, int yi, int lo, int hi, int x, int y,
                                         int H, int W, int R, bool short_taps, const float *taps,
                                         float scale) {
  if (x < xi - R || x > xi + R || y < lo || y > hi) return 0;
  const float iws = I * fast_weight(x, xi, W, R, short_taps, taps) * scale;
  return fast_fixed(iws, fast_weight(y, yi, H, R, short_taps, taps));
}

__device__ __forceinline__ void racc_add(racc_t *p, float v, float scale) {
  atomicAdd(p, __float2int_rn(v * scale));
}

// Scratch words the tile bounds and the maximum search borrow from the strip
// buffer: int tile_u[n_tiles] | short cand_q[kMaxCand][kCandSpots].
__device__ __forceinline__ int bound_words(int n_tiles) {
  return n_tiles + (kMaxCand * kCandSpots * 2 + 3) / 4;
}

// Upper bound U of every pixel of the pattern: per 16x16 tile, the sum over
// the spots touching it of the largest value each can take there, counted in
// integer units of P / 2^20 (P = the largest spot weight, rounded up) so the
// bound -- and with it the fixed-point scale -- does not depend on the order
// of the atomic adds.  Leaves the per-tile sums in scratch (for the maximum
// search) when they fit; otherwise U is the sum over all spots.  Also notes
// the brightest spot (fast mode).  Returns U in units of P / 2^20.
constexpr float kBoundUnits = 1048576.0f;  // 2^20

template <int kMode>
__device__ int tile_bounds(RenderShared<kMode> &sh, int *tile_u, bool tiles, int ntx, int n_tiles,
                           int ns, int H, int W, int R, bool short_taps, float pref, float pmax,
                           int tid, int lane, int wid) {
  if (tiles)
    for (int t = tid; t < n_tiles; t += kRenderThreads) tile_u[t] = 0;
  if (tid == 0) sh.qstar = 0;
  lds_sync();
  const float w0 = sh.taps[0];
  const float inv_p = pmax > 0.0f ? kBoundUnits / pmax : 0.0f;
  int usum = 0;
  for (int q = tid; q < ns; q += kRenderThreads) {
    const float I = sh.I[q];
    int tx0, tx1;
    float uf;
    bool tight = false;  // interior fast-mode spot: per-tile bound from the nearest taps
    int xi = 0, yi = 0;
    if (kMode == kModeFast) {
      if (fabsf(I) == pmax) sh.qstar = q;
      yi = sh.xy[q] >> 16;
      xi = sh.xy[q] & 0xffff;
      tight = short_taps && xi >= R && xi + R < W && yi >= R && yi + R < H;
      const float bx = !short_taps ? 1.0f : ((xi >= R && xi + R < W) ? w0 : fminf(1.0f, 3.0f * w0));
      const float by = !short_taps ? 1.0f : ((yi >= R && yi + R < H) ? w0 : fminf(1.0f, 3.0f * w0));
      uf = fabsf(I) * bx * by;
      tx0 = max(0, xi - R) / kTile;
      tx1 = min(W - 1, xi + R) / kTile;
    } else {
      uf = fabsf(I) * pref;
      tx0 = (sh.xy[q] >> 16) / kTile;
      tx1 = (sh.xy[q] & 0xffff) / kTile;
    }
    // uf <= P: at most 2^20 units, rounded up (+1 covers the float product)
    const int u = min(static_cast<int>(ceilf(uf * inv_p)) + 1, static_cast<int>(kBoundUnits) + 1);
    usum += u;
    if (tiles) {
      const int ty0 = sh.ylo[q] / kTile, ty1 = sh.yhi[q] / kTile;
      for (int ty = ty0; ty <= ty1; ++ty)
        for (int tx = tx0; tx <= tx1; ++tx) {
          int ut = u;
          if (tight) {
            // the taps fall off from the centre: the tile's largest weight is
            // at its pixel nearest (xi, yi), at most R away along each axis
            const int dx = max(0, max(tx * kTile - xi, xi - (tx * kTile + kTile - 1)));
            const int dy = max(0, max(ty * kTile - yi, yi - (ty * kTile + kTile - 1)));
            const float uft = fabsf(I) * sh.taps[dx] * sh.taps[dy];
            ut = min(static_cast<int>(ceilf(uft * inv_p)) + 1, static_cast<int>(kBoundUnits) + 1);
          }
          atomicAdd(&tile_u[ty * ntx + tx], ut);
        }
    }
  }
  if (!tiles) {
    usum = wave_sum(usum);
    if (lane == 0) sh.wu[wid] = usum;
    lds_sync();
    int U = 0;
    for (int w = 0; w < kRenderThreads / kWave; ++w) U += sh.wu[w];
    lds_sync();
    return U;
  }
  lds_sync();
  int lu = 0;
  for (int t = tid; t < n_tiles; t += kRenderThreads) lu = max(lu, tile_u[t]);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) lu = max(lu, __shfl_xor(lu, off, kWave));
  if (lane == 0) sh.wu[wid] = lu;
  lds_sync();
  int U = 0;
  for (int w = 0; w < kRenderThreads / kWave; ++w) U = max(U, sh.wu[w]);
  lds_sync();
  return U;
}

// Exact pattern maximum (fixed point) without rendering the whole pattern:
// a tile whose bound stays below the value already known at the brightest
// spot's pixel cannot hold the maximum.  Only the few remaining tiles are
// evaluated, pixel by pixel, with the splat's exact fixed-point terms, so the
// result equals a full pass bit for bit.  Returns false (uniformly) when too
// many tiles remain; the caller then runs the full pass.
template <int kMode>
__device__ bool bounded_max(RenderShared<kMode> &sh, int *tile_u, short (*cand_q)[kCandSpots],
                            int ntx, int n_tiles, int ns, int H, int W, int R, bool short_taps,
                            float scale, float unit, int tid, int lane, int wid, racc_t &m_out) {
  if (tid == 0) {
    sh.n_cand = 0;
    sh.fail = 0;
  }
  // lower bound: the exact value at the brightest survivor's pixel (wave 0)
  if (wid == 0) {
    const int qs = sh.qstar;
    const int ys = sh.xy[qs] >> 16, xs = sh.xy[qs] & 0xffff;
    long long acc = 0;
    for (int q = lane; q < ns; q += kWave)
      acc += fast_term(sh.I[q], sh.xy[q] & 0xffff, sh.xy[q] >> 16, sh.ylo[q], sh.yhi[q], xs, ys, H,
                       W, R, short_taps, sh.taps, scale);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, kWave);
    if (lane == 0) sh.lfix = acc;
  }
  lds_sync();
  const long long lfix = sh.lfix;
  // candidate tiles: bound >= lower bound.  Truncated terms never exceed
  // their float products, which the 1e-5 factor covers (the subpixel terms
  // are rounded to nearest: half a unit each); tile_u is reused to hold each
  // tile's candidate slot (-1: pruned)
  int *slot_of = tile_u;
  const float margin = (kMode == kModeFast ? 0.0f : 0.5f * static_cast<float>(ns)) + 1.0f;
  for (int t = tid; t < n_tiles; t += kRenderThreads) {
    const float bound = static_cast<float>(tile_u[t]) * unit * scale * 1.00001f + margin;
    int slot = -1;
    if (bound >= static_cast<float>(lfix)) {
      slot = atomicAdd(&sh.n_cand, 1);
      if (slot < kMaxCand) {
        sh.cand_tile[slot] = t;
        sh.cand_n[slot] = 0;
      } else {
        slot = -1;
        sh.fail = 1;
      }
    }
    slot_of[t] = slot;
  }
  lds_sync();
  if (sh.fail) return false;
  const int n_cand = sh.n_cand;
  for (int q = tid; q < ns; q += kRenderThreads) {
    const int yi = sh.xy[q] >> 16, xi = sh.xy[q] & 0xffff;
    const int tx0 = max(0, xi - R) / kTile, tx1 = min(W - 1, xi + R) / kTile;
    const int ty0 = sh.ylo[q] / kTile, ty1 = sh.yhi[q] / kTile;
    for (int ty = ty0; ty <= ty1; ++ty)
      for (int tx = tx0; tx <= tx1; ++tx) {
        const int sl = slot_of[ty * ntx + tx];
        if (sl < 0) continue;
        const int pos = atomicAdd(&sh.cand_n[sl], 1);
        if (pos < kCandSpots)
          cand_q[sl][pos] = static_cast<short>(q);
        else
          sh.fail = 1;
      }
  }
  lds_sync();
  if (sh.fail) return false;
  long long lmax = lfix;
  for (int i = tid; i < n_cand * kTile * kTile; i += kRenderThreads) {
    const int sl = i / (kTile * kTile);
    const int t = sh.cand_tile[sl];
    const int x = (t % ntx) * kTile + (i % kTile), y = (t / ntx) * kTile + (i / kTile) % kTile;
    if (x >= W || y >= H) continue;
    long long acc = 0;
    const int nq = sh.cand_n[sl];
    for (int k = 0; k < nq; ++k) {
      const int q = cand_q[sl][k];
      acc += fast_term(sh.I[q], sh.xy[q] & 0xffff, sh.xy[q] >> 16, sh.ylo[q], sh.yhi[q], x, y, H, W,
                       R, short_taps, sh.taps, scale);
    }
    lmax = acc > lmax ? acc : lmax;
  }
  // the bound keeps every pixel sum below 2^31
  racc_t lm = static_cast<racc_t>(lmax);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) lm = max(lm, __shfl_xor(lm, off, kWave));
  if (lane == 0) sh.wmax[wid] = lm;
  lds_sync();
  racc_t m = 0;
  for (int w = 0; w < kRenderThreads / kWave; ++w) m = max(m, sh.wmax[w]);
  m_out = m;
  lds_sync();  // wmax and the candidate state are reused
  return true;
}

// kR > 0: the fast-mode tap radius as a compile-time constant (the footprint
// loops unroll and the taps stay in scalar registers); 0: taken from the args.
template <int kMode, int kR>
__global__ __launch_bounds__(kRenderThreads, render_min_waves<kMode>()) void render_kernel(RenderArgs a) {
  // 16-byte aligned: the read-outs use 16-byte LDS accesses, which run at a
  // fraction of their rate on an 8-byte-aligned base
  extern __shared__ __align__(16) racc_t s_acc[];  // strip accumulator; hash table before the passes
  __shared__ RenderShared<kMode> sh;
  // a pattern list in device memory may hold fewer patterns than the grid has blocks
  if (a.n_dev && static_cast<int>(blockIdx.x) >= *a.n_dev) return;
  int *h_key = reinterpret_cast<int *>(s_acc);
  int *h_val = h_key + kHash;
  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int wid = tid / kWave;
  const int W = a.W, H = a.H;
  const int R = kR > 0 ? kR : a.R;
  const bool vec4 = (W & 3) == 0;
  const bool short_taps = R < min(H, W);
  for (int i = tid; i <= kMaxR; i += kRenderThreads) sh.taps[i] = a.taps[i];
  // the strip buffer is kept zeroed between uses (each read-out clears it)
  for (int i = tid; i < a.lds_words; i += kRenderThreads) s_acc[i] = 0;
  // spot tid and the count of the block's next pattern, loaded while the
  // current pattern renders so the scans below start without a memory wait
  float4 next_head = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  int next_count = 0;
  const int n_pat = a.n_dev ? *a.n_dev : a.n;
  auto prefetch = [&](int pi) {
    if (pi >= n_pat) return;
    const int p = a.list ? a.list[pi] : pi;
    const int c = a.capv ? a.capv[p] : a.cap;
    const float4 *sp = a.spots + (a.spot_off ? a.spot_off[p] : (long long)p * a.cap);
    next_count = min(a.count[p], c);
    if (tid < next_count) next_head = sp[tid];
  };
  prefetch(blockIdx.x);
  lds_sync();

  for (int pi = blockIdx.x; pi < n_pat; pi += gridDim.x) {
    const int p = a.list ? a.list[pi] : pi;
    const int n = next_count;
    const float4 head = next_head;  // spots[tid] when tid < n
    next_head = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // dead until the prefetch below
    const float4 *spots = a.spots + (a.spot_off ? a.spot_off[p] : (long long)p * a.cap);
    if (kMode == kModeFast) {
      for (int i = tid; i < kHash; i += kRenderThreads) {
        h_key[i] = -1;
        h_val[i] = -1;
      }
    }
    if (tid == 0) {
      sh.ns = 0;
      sh.overflow = 0;
      sh.cols_max = 0;
    }
    for (int s = tid; s <= kMaxStrips; s += kRenderThreads) sh.bin_start[s] = 0;
    lds_sync();

    // ---- 1a. last write wins: owner of each pixel = largest list position
    if (kMode == kModeFast) {
      for (int j = tid; j < n; j += kRenderThreads) {
        float X, Y;
        if (!to_pixel(j == tid ? head : spots[j], a, X, Y)) continue;
        const int key = min(static_cast<int>(Y), H - 1) * W + min(static_cast<int>(X), W - 1);
        unsigned slot = hash_slot(key);
        bool placed = false;
        for (int probe = 0; probe < kHash; ++probe) {
          const int prev = atomicCAS(&h_key[slot], -1, key);
          if (prev == -1 || prev == key) {
            atomicMax(&h_val[slot], j);
            placed = true;
            break;
          }
          slot = (slot + 1) & (kHash - 1);
        }
        if (!placed) sh.overflow = 1;
      }
      lds_sync();
    }
    // ---- 1b. survivors -> LDS list with their footprint -----------------
    float tmax = 0.0f;  // largest spot weight: bound on any single contribution
    for (int j = tid; j < n; j += kRenderThreads) {
      const float4 sp = j == tid ? head : spots[j];
      float X, Y;
      if (!to_pixel(sp, a, X, Y)) continue;
      int lo, hi, packed, width;
      float fxv = 0.0f, fyv = 0.0f;
      if (kMode == kModeFast) {
        const int yi = min(static_cast<int>(Y), H - 1), xi = min(static_cast<int>(X), W - 1);
        const int key = yi * W + xi;
        unsigned slot = hash_slot(key);
        bool win = false;
        for (int probe = 0; probe < kHash; ++probe) {
          const int k = h_key[slot];
          if (k == key) {
            win = h_val[slot] == j;
            break;
          }
          if (k == -1) break;
          slot = (slot + 1) & (kHash - 1);
        }
        if (!win) continue;
        lo = max(0, yi - R);
        hi = min(H - 1, yi + R);
        packed = (yi << 16) | xi;
        width = 2 * R + 1;
      } else {
        int xlo, xhi;
        if (!subpixel_box(sp.w, X, Y, a.pref, a.ep, a.clip, H, W, lo, hi, xlo, xhi)) continue;
        packed = (xlo << 16) | xhi;
        width = xhi - xlo + 1;
        fxv = X;
        fyv = Y;
      }
      const int q = atomicAdd(&sh.ns, 1);
      if (q >= kMaxSpots) {
        sh.overflow = 1;
        continue;
      }
      tmax = fmaxf(tmax, fabsf(sp.w) * (kMode == kModeFast ? 1.0f : a.pref));
      sh.I[q] = sp.w;
      sh.xy[q] = packed;
      if (kMode == kModeSubpixel) {  // fx/fy have one slot in fast mode
        sh.fx[q] = fxv;
        sh.fy[q] = fyv;
      }
      sh.ylo[q] = static_cast<short>(lo);
      sh.yhi[q] = static_cast<short>(hi);
      atomicMax(&sh.cols_max, width);
      for (int s = lo / a.strip_rows; s <= hi / a.strip_rows; ++s) atomicAdd(&sh.bin_start[s + 1], 1);
    }
    tmax = wave_max(tmax);
    if (lane == 0) sh.wsum[wid] = tmax;
    lds_sync();
    if (kMode == kModeFast) {
      // the hash lived in the strip buffer: clear it for the splats
      for (int i = tid; i < min(2 * kHash, a.lds_words); i += kRenderThreads) s_acc[i] = 0;
    }
    const int ns = min(sh.ns, kMaxSpots);
    if (tid == 0 && a.status) a.status[p] = sh.overflow;
    float *out = a.out + (size_t)p * H * W;

    if (ns == 0) {
      // no spot in frame: np.zeros(shape), not normalised
      if (vec4) {
        const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        for (int i = tid; i < H * W / 4; i += kRenderThreads)
          store_nt(z, reinterpret_cast<float4 *>(out) + i);
      } else {
        for (int i = tid; i < H * W; i += kRenderThreads) __builtin_nontemporal_store(0.0f, out + i);
      }
      prefetch(pi + gridDim.x);
      lds_sync();
      continue;
    }
    float pmax = 0.0f;
    for (int w = 0; w < kRenderThreads / kWave; ++w) pmax = fmaxf(pmax, sh.wsum[w]);

    // ---- 2. strip bins ---------------------------------------------------
    if (tid == 0) {
      int acc = 0;
      for (int s = 0; s < a.n_strips; ++s) {
        acc += sh.bin_start[s + 1];
        sh.bin_start[s + 1] = acc;
      }
    }
    lds_sync();
    const bool binned = sh.bin_start[a.n_strips] <= kBinCap;
    for (int s = tid; s < a.n_strips; s += kRenderThreads) sh.bin_fill[s] = sh.bin_start[s];
    lds_sync();
    if (binned) {
      for (int q = tid; q < ns; q += kRenderThreads) {
        for (int s = sh.ylo[q] / a.strip_rows; s <= sh.yhi[q] / a.strip_rows; ++s) {
          const int pos = atomicAdd(&sh.bin_fill[s], 1);
          sh.bin[pos] = static_cast<short>(q);
        }
      }
    }
    lds_sync();
    // splat work item = (spot in bin, footprint column); each item walks the
    // footprint rows inside the strip, so the lanes of a wave add into
    // consecutive LDS columns
    const int ncols = sh.cols_max;
    const float inv_ncols = 1.0f / static_cast<float>(ncols);

    // ---- 3. fixed-point scale from the tile bounds; maximum (fast mode) -----
    const int ntx = (W + kTile - 1) / kTile;
    const int n_tiles = ntx * ((H + kTile - 1) / kTile);
    const bool tiles = n_tiles <= kMaxTiles && bound_words(n_tiles) <= a.lds_words;
    int *tile_u = s_acc;
    short(*cand_q)[kCandSpots] = reinterpret_cast<short(*)[kCandSpots]>(s_acc + n_tiles);
    const int Ui = tile_bounds(sh, tile_u, tiles, ntx, n_tiles, ns, H, W, R, short_taps, a.pref,
                               pmax, tid, lane, wid);
    const float unit = pmax / kBoundUnits;  // value of one bound unit
    const float U = static_cast<float>(Ui) * unit;
    const float scale = U > 0.0f ? kScaleRange / (U * 1.00001f) : 1.0f;
    float out_scale = 1.0f / scale;  // fixed point -> value (normalize: -> value/max)
    int first_pass = a.normalize ? 0 : 1;
    if (kMode == kModeFast && a.normalize && tiles) {
      racc_t m;
      if (bounded_max(sh, tile_u, cand_q, ntx, n_tiles, ns, H, W, R, short_taps, scale, unit, tid,
                      lane, wid, m)) {
        out_scale = m > 0 ? 1.0f / static_cast<float>(m) : 1.0f / scale;
        first_pass = 1;
      }
    }
    if (tiles) {  // hand the borrowed words back zeroed
      for (int i = tid; i < bound_words(n_tiles); i += kRenderThreads) s_acc[i] = 0;
      lds_sync();
    }
    for (int pass = first_pass; pass < 2; ++pass) {
      racc_t lmax = 0;
      for (int s = 0; s < a.n_strips; ++s) {
        const int y0 = s * a.strip_rows;
        const int y1 = min(H, y0 + a.strip_rows);
        const int cnt = (y1 - y0) * W;
        int b0 = 0, nb = ns;
        if (binned) {
          b0 = sh.bin_start[s];
          nb = sh.bin_start[s + 1] - b0;
        }
        float *dst = out + (size_t)y0 * W;
        // the next pattern's head loads while the last strip is read out (one
        // call per pattern on every path)
        const bool last = pass == 1 && s == a.n_strips - 1;
        if (nb == 0) {
          if (last) prefetch(pi + gridDim.x);
          if (pass == 1) {
            if (vec4) {
              const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
              for (int i = tid; i < cnt / 4; i += kRenderThreads)
                store_nt(z, reinterpret_cast<float4 *>(dst) + i);
            } else {
              for (int i = tid; i < cnt; i += kRenderThreads) __builtin_nontemporal_store(0.0f, dst + i);
            }
          }
          continue;
        }
        const int n_items = nb * ncols;
        for (int it = tid; it < n_items; it += kRenderThreads) {
          int bi = static_cast<int>(static_cast<float>(it) * inv_ncols);
          int c = it - bi * ncols;
          if (c < 0) {
            --bi;
            c += ncols;
          } else if (c >= ncols) {
            ++bi;
            c -= ncols;
          }
          const int q = binned ? sh.bin[b0 + bi] : bi;
          const int ya = max(static_cast<int>(sh.ylo[q]), y0);
          const int yb = min(static_cast<int>(sh.yhi[q]), y1 - 1);
          if (ya > yb) continue;
          const float I = sh.I[q];
          if (kMode == kModeFast) {
            const int yi = sh.xy[q] >> 16, xi = sh.xy[q] & 0xffff;
            const int x = xi - R + c;
            if (x < 0 || x >= W) continue;
            float wx;
            if (!short_taps)
              wx = reflect_weight(x, xi, W, R, sh.taps);
            else if (xi >= R && xi + R < W)
              wx = sh.taps[c > R ? c - R : R - c];
            else
              wx = reflect_weight3(x, xi, W, R, sh.taps);
            const float iws = I * wx * scale;
            racc_t *col = s_acc + x;
            if (short_taps && yi >= R && yi + R < H) {
              // the tap index is uniform across the wave: taps come from
              // scalar registers, not from a dependent LDS load per row
              racc_t *top = col + (yi - R - y0) * W;  // row yi - R (dereferenced only in the strip)
              if (ya == yi - R && yb == yi + R) {
                // the whole footprint lies in this strip: no per-row tests
#pragma unroll
                for (int dd = -R; dd <= R; ++dd)
                  atomicAdd(top + (dd + R) * W, fast_fixed(iws, a.taps[dd < 0 ? -dd : dd]));
              } else {
#pragma unroll
                for (int dd = -R; dd <= R; ++dd) {
                  const int y = yi + dd;
                  if (y >= ya && y <= yb) atomicAdd(top + (dd + R) * W, fast_fixed(iws, a.taps[dd < 0 ? -dd : dd]));
                }
              }
            } else {
              for (int y = ya; y <= yb; ++y) {
                const float wy = short_taps ? reflect_weight3(y, yi, H, R, sh.taps)
                                            : reflect_weight(y, yi, H, R, sh.taps);
                atomicAdd(col + (y - y0) * W, fast_fixed(iws, wy));
              }
            }
          } else {
            const int xlo = sh.xy[q] >> 16, xhi = sh.xy[q] & 0xffff;
            const int x = xlo + c;
            if (x > xhi) continue;
            const float dx = static_cast<float>(x) - sh.fx[q];
            const float amp = I * a.pref;
            const float cy = sh.fy[q];
            racc_t *col = s_acc + x;
            lds_reads_done();
            for (int y = ya; y <= yb; ++y) {
              const float dy = static_cast<float>(y) - cy;
              racc_add(col + (y - y0) * W, amp * fast_exp(a.ep * (dy * dy + dx * dx)), scale);
            }
          }
        }
        lds_sync();
        if (last) prefetch(pi + gridDim.x);
        // read out and clear the strip (16-byte LDS accesses: 4 words)
        if (pass == 0) {
          if ((cnt & 3) == 0) {
#pragma unroll 4
            for (int i = tid; i < cnt / 4; i += kRenderThreads) {
              int4 *pv = reinterpret_cast<int4 *>(s_acc) + i;
              const int4 v = *pv;
              *pv = make_int4(0, 0, 0, 0);
              lmax = max(lmax, max(max(v.x, v.y), max(v.z, v.w)));
            }
          } else {
            for (int i = tid; i < cnt; i += kRenderThreads) {
              const racc_t v = s_acc[i];
              s_acc[i] = 0;
              lmax = max(lmax, v);
            }
          }
        } else if (vec4) {
#pragma unroll 2
          for (int i = tid; i < cnt / 4; i += kRenderThreads) {
            int4 *pv = reinterpret_cast<int4 *>(s_acc) + i;
            const int4 u = *pv;
            *pv = make_int4(0, 0, 0, 0);
            const float4 o = make_float4(static_cast<float>(u.x) * out_scale, static_cast<float>(u.y) * out_scale,
                                         static_cast<float>(u.z) * out_scale, static_cast<float>(u.w) * out_scale);
            store_nt(o, reinterpret_cast<float4 *>(dst) + i);
          }
        } else {
          for (int i = tid; i < cnt; i += kRenderThreads) {
            __builtin_nontemporal_store(static_cast<float>(s_acc[i]) * out_scale, dst + i);
            s_acc[i] = 0;
          }
        }
        lds_sync();
      }
      if (pass == 0) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) lmax = max(lmax, __shfl_xor(lmax, off, kWave));
        if (lane == 0) sh.wmax[wid] = lmax;
        lds_sync();
        racc_t m = 0;
        for (int w = 0; w < kRenderThreads / kWave; ++w) m = max(m, sh.wmax[w]);
        // a zero maximum (every spot has zero intensity) leaves the zeros as
        // they are instead of producing 0/0
        out_scale = m > 0 ? 1.0f / static_cast<float>(m) : 1.0f / scale;
        lds_sync();
      }
    }
    // every thread is done with this pattern's LDS state before the next one
    lds_sync();
  }
}

// ---------------------------------------------------------------------------
// Band renderer (fast mode): the per-pattern spot work and the pattern writes
// run as two kernels, so that the writes come from a one-shot grid of short
// blocks (the store pattern that reaches 6-7 TB/s here, against ~5.9 TB/s
// for blocks streaming whole patterns).
//
//   render_plan_kernel (one wave per pattern): pixel coordinates, last write
//   wins, the 16x16 tile bounds, the fixed-point scale and the exact maximum
//   -- the same integer and float operations as render_kernel, so the output
//   is the same bit for bit -- then the pattern's spots sorted by row, with
//   each band's window of that list, into the workspace;
//
//   render_band_kernel (one block per band of rows, at most 2048 pixels,
//   every band of a pattern on one XCD): the band's window of the row-sorted
//   list (one scalar round for its bounds, one vector round for the spots;
//   neighbouring bands read overlapping windows, which stay in that XCD's
//   L2), the splat into LDS (32-bit fixed point), two scaled non-temporal
//   16-byte stores per thread.
//
// Patterns the plan cannot hold (more than kPlanMax listed spots, too many
// candidate tiles for the maximum) are put on a list that render_kernel
// renders afterwards.
// ---------------------------------------------------------------------------

// patterns per plan block: one, so that the dispatcher refills a finished
// wave's slot at once (the waves share nothing; render stage of 200,000
// library patterns 8.44 ms with one, 8.46 with two, 8.51 with four)
constexpr int kPlanWaves = 1;
constexpr int kPlanThreads = kPlanWaves * kWave;
constexpr int kPlanPerLane = 8;
constexpr int kPlanMax = kPlanPerLane * kWave;  // listed spots per pattern
constexpr int kPlanMaxSide = 1024;              // 10-bit pixel coordinates in the sort key
constexpr int kBandThreads = 256;
constexpr int kBandVec = 2;                              // float4 stores per thread
constexpr int kBandPix = 4 * kBandVec * kBandThreads;    // pixels per band
// patterns per plan/writer round: the plan's hand-over for 16,384 patterns
// stays in the last-level cache while the writer reads it (render stage of
// 200,000 library patterns: 8.50 ms with 16,384, 8.57 with 12,288 or 32,768,
// 8.81 with 8,192 and 11.3 with 65,536)
constexpr int kBandChunk = 16384;

struct BandArgs {
  RenderArgs r;     // spot lists, detector mapping, taps, output
  int *rec;         // per pattern: {listed, spots, scale, out_scale}, first spot of rows 0..H
  int2 *ent;        // per pattern: spots sorted by row, {(y << 16) | x, I}
  int *fallback;    // [0]: patterns listed, [1 + i]: their ids
  int rec_ints;     // stride of rec
  int p0, m;        // first pattern of the round, patterns in it
  int band_rows, nb;
  int ntx, n_tiles;
  int wave_ints;    // LDS words of one plan wave
};

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off, kWave));
  return v;
}

__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = min(v, __shfl_xor(v, off, kWave));
  return v;
}

__device__ __forceinline__ long long wave_sum_ll(long long v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// fast_weight for short taps: an interior source (the footprint clear of both
// edges) has only its own image in reach, and reflect_weight3 then adds two
// zeros to that tap -- the same value.
__device__ __forceinline__ float band_weight(int y, int c, int n, int R, const float *taps) {
  return (c >= R && c + R < n) ? tap_at(c - y, R, taps) : reflect_weight3(y, c, n, R, taps);
}

// Distance from pixel c to the nearest pixel of tile t along one axis.
__device__ __forceinline__ int tile_gap(int t, int c) {
  return max(0, max(t * kTile - c, c - (t * kTile + kTile - 1)));
}

// tile_bounds' integer units of a term |I| wx wy (rounded up, plus one).
__device__ __forceinline__ int bound_units(float aI, float wx, float wy, float inv_p) {
  return min(static_cast<int>(ceilf(aI * wx * wy * inv_p)) + 1, static_cast<int>(kBoundUnits) + 1);
}

// Row-sorted key of the plan: (y << 19) | (x << 9) | list position (y, x < 1024, position < 512).
__device__ __forceinline__ int key_y(int k) { return k >> 19; }
__device__ __forceinline__ int key_x(int k) { return (k >> 9) & 1023; }

// Per-wave LDS of the plan: row starts [H + 1] | tile bounds, later one
// candidate tile's pixels [max(n_tiles, kTile^2)] | row-sorted keys and
// intensities [kPlanMax each] | candidate spots [kCandSpots] | candidate
// tiles and their bounds [kMaxCand each] | taps [kMaxR + 1].
__host__ __device__ __forceinline__ int plan_wave_ints(int H, int n_tiles) {
  return (H + 1) + (n_tiles > kTile * kTile ? n_tiles : kTile * kTile) + 2 * kPlanMax + kCandSpots +
         2 * kMaxCand + (kMaxR + 1);
}

template <int kR>
__global__ __launch_bounds__(kPlanThreads) void render_plan_kernel(BandArgs b) {
  extern __shared__ int s_plan[];
  const RenderArgs &a = b.r;
  constexpr int R = kR;
  const int lane = lane_id();
  const int wid = threadIdx.x / kWave;
  // pattern pl is planned on XCD pl mod 8 (blocks are dispatched round-robin
  // over the 8 XCDs), where its bands are rendered
  const int pl = (blockIdx.x >> 3) * (8 * kPlanWaves) + wid * 8 + (blockIdx.x & 7);
  if (pl >= b.m) return;  // waves are independent: no block barriers below
  const int p = b.p0 + pl;
  const int H = a.H, W = a.W, ntx = b.ntx, n_tiles = b.n_tiles;
  int *rows = s_plan + wid * b.wave_ints;
  int *tile_u = rows + (H + 1);
  // the candidate tiles' pixels take the bounds' place once the candidates
  // and their bounds have been read out
  int *s_tile = tile_u;
  int *s_key = tile_u + max(n_tiles, kTile * kTile);
  float *s_I = reinterpret_cast<float *>(s_key + kPlanMax);
  int *s_cq = s_key + 2 * kPlanMax;
  int *s_ct = s_cq + kCandSpots;
  float *s_cb = reinterpret_cast<float *>(s_ct + kMaxCand);
  float *s_tp = s_cb + kMaxCand;
  int *rec = b.rec + static_cast<size_t>(pl) * b.rec_ints;
  const int n = min(a.count[p], a.cap);
  bool ok = n <= kPlanMax;
  int ns = 0;
  float scale = 1.0f, out_scale = 1.0f;
  if (ok) {
    for (int t = lane; t <= H; t += kWave) rows[t] = 0;
    for (int t = lane; t < n_tiles; t += kWave) tile_u[t] = 0;
    if (lane <= R) s_tp[lane] = a.taps[lane];
    wave_lds_sync();
    // ---- pixel of every listed spot; count per row ------------------------
    const float4 *sp = a.spots + static_cast<size_t>(p) * a.cap;
    int key[kPlanPerLane], pos[kPlanPerLane];
    float Iv[kPlanPerLane];
#pragma unroll
    for (int k = 0; k < kPlanPerLane; ++k) {
      const int j = lane + k * kWave;
      key[k] = -1;
      Iv[k] = 0.0f;
      pos[k] = 0;
      if (j < n) {
        const float4 s = sp[j];
        float X, Y;
        if (to_pixel(s, a, X, Y)) {
          const int yi = min(static_cast<int>(Y), H - 1), xi = min(static_cast<int>(X), W - 1);
          key[k] = (yi << 19) | (xi << 9) | j;
          Iv[k] = s.w;
          pos[k] = atomicAdd(&rows[yi + 1], 1);
        }
      }
    }
    wave_lds_sync();
    // ---- row starts (exclusive scan), spots sorted by row -----------------
    {
      const int per = (H + kWave - 1) / kWave;
      const int i0 = min(H, lane * per), i1 = min(H, i0 + per);
      int sum = 0;
      for (int t = i0; t < i1; ++t) sum += rows[1 + t];
      int run = wave_inclusive_scan(sum) - sum;
      for (int t = i0; t < i1; ++t) {
        run += rows[1 + t];
        rows[1 + t] = run;
      }
    }
    wave_lds_sync();
#pragma unroll
    for (int k = 0; k < kPlanPerLane; ++k) {
      if (key[k] >= 0) {
        const int q = rows[key_y(key[k])] + pos[k];
        pos[k] = q;
        s_key[q] = key[k];
        s_I[q] = Iv[k];
      }
    }
    wave_lds_sync();
    // ---- last write wins: a spot yields to a later one on its pixel -------
#pragma unroll
    for (int k = 0; k < kPlanPerLane; ++k) {
      if (key[k] >= 0) {
        const int yi = key_y(key[k]), px = key[k] >> 9, j = key[k] & 511;
        bool lose = false;
        for (int e = rows[yi]; e < rows[yi + 1]; ++e) {
          const int o = s_key[e];
          lose |= (o >> 9) == px && (o & 511) > j;
        }
        if (lose) {
          key[k] = -1;
          s_I[pos[k]] = 0.0f;  // contributes nothing where the bands read it
        }
      }
    }
    wave_lds_sync();
    float pm = 0.0f;
    int mine = 0;
#pragma unroll
    for (int k = 0; k < kPlanPerLane; ++k) {
      if (key[k] >= 0) {
        pm = fmaxf(pm, fabsf(Iv[k]));
        ++mine;
      }
    }
    const float pmax = wave_max(pm);
    ns = wave_sum(mine);
    if (ns > 0) {
      // ---- tile bounds and the fixed-point scale (as tile_bounds) ---------
      const float w0 = s_tp[0];
      const float inv_p = pmax > 0.0f ? kBoundUnits / pmax : 0.0f;
#pragma unroll
      for (int k = 0; k < kPlanPerLane; ++k) {
        if (key[k] < 0) continue;
        const int yi = key_y(key[k]), xi = key_x(key[k]);
        const float I = Iv[k];
        const bool inx = xi >= R && xi + R < W, iny = yi >= R && yi + R < H;
        const float bx = inx ? w0 : fminf(1.0f, 3.0f * w0);
        const float by = iny ? w0 : fminf(1.0f, 3.0f * w0);
        const float uf = fabsf(I) * bx * by;
        const int u = min(static_cast<int>(ceilf(uf * inv_p)) + 1, static_cast<int>(kBoundUnits) + 1);
        // a footprint of at most 17 pixels reaches at most two tiles per axis
        const int tx0 = max(0, xi - R) / kTile, tx1 = min(W - 1, xi + R) / kTile;
        const int ty0 = max(0, yi - R) / kTile, ty1 = min(H - 1, yi + R) / kTile;
        int u00 = u, u01 = u, u10 = u, u11 = u;
        if (inx && iny) {
          // the taps fall off from the centre: a tile's largest weight is at
          // its pixel nearest (xi, yi)
          const float aI = fabsf(I);
          const float wx0 = s_tp[tile_gap(tx0, xi)], wx1 = s_tp[tile_gap(tx1, xi)];
          const float wy0 = s_tp[tile_gap(ty0, yi)], wy1 = s_tp[tile_gap(ty1, yi)];
          u00 = bound_units(aI, wx0, wy0, inv_p);
          u01 = bound_units(aI, wx1, wy0, inv_p);
          u10 = bound_units(aI, wx0, wy1, inv_p);
          u11 = bound_units(aI, wx1, wy1, inv_p);
        }
        atomicAdd(&tile_u[ty0 * ntx + tx0], u00);
        if (tx1 != tx0) atomicAdd(&tile_u[ty0 * ntx + tx1], u01);
        if (ty1 != ty0) {
          atomicAdd(&tile_u[ty1 * ntx + tx0], u10);
          if (tx1 != tx0) atomicAdd(&tile_u[ty1 * ntx + tx1], u11);
        }
      }
      wave_lds_sync();
      int lu = 0;
      for (int t = lane; t < n_tiles; t += kWave) lu = max(lu, tile_u[t]);
      const int Ui = wave_max_i(lu);
      const float unit = pmax / kBoundUnits;
      const float U = static_cast<float>(Ui) * unit;
      scale = U > 0.0f ? kScaleRange / (U * 1.00001f) : 1.0f;
      out_scale = 1.0f / scale;
      if (a.normalize) {
        // ---- exact maximum (as bounded_max) ------------------------------
        int qc = 0x7fffffff;
#pragma unroll
        for (int k = 0; k < kPlanPerLane; ++k)
          if (key[k] >= 0 && fabsf(Iv[k]) == pmax) qc = min(qc, pos[k]);
        const int qs = wave_min_i(qc);
        const int ys = key_y(s_key[qs]), xs = key_x(s_key[qs]);
        long long acc = 0;
#pragma unroll
        for (int k = 0; k < kPlanPerLane; ++k) {
          if (key[k] < 0) continue;
          const int yi = key_y(key[k]), xi = key_x(key[k]);
          acc += fast_term(Iv[k], xi, yi, max(0, yi - R), min(H - 1, yi + R), xs, ys, H, W, R, true,
                           s_tp, scale);
        }
        const long long lfix = wave_sum_ll(acc);
        int n_cand = 0;
        for (int t0 = 0; t0 < n_tiles; t0 += kWave) {
          const int t = t0 + lane;
          bool c = false;
          if (t < n_tiles) {
            const float bound = static_cast<float>(tile_u[t]) * unit * scale * 1.00001f + 1.0f;
            c = bound >= static_cast<float>(lfix);
          }
          const uint64_t mk = __ballot(c);
          if (c) {
            const int sl = n_cand + mask_rank(mk);
            if (sl < kMaxCand) s_ct[sl] = t;
          }
          n_cand += __popcll(mk);
        }
        ok = n_cand <= kMaxCand;
        wave_lds_sync();
        // the candidates in descending order of their bounds: once a bound
        // is below the largest exact value found, no later tile can hold a
        // larger one (exact comparison in double)
        float cb = 0.0f;
        int ct = 0, rank = 0;
        if (ok && lane < n_cand) {
          ct = s_ct[lane];
          cb = static_cast<float>(tile_u[ct]) * unit * scale * 1.00001f + 1.0f;
        }
        for (int o = 0; ok && o < n_cand; ++o) {
          const float ob = __shfl(cb, o, kWave);
          rank += (ob > cb || (ob == cb && o < lane)) ? 1 : 0;
        }
        wave_lds_sync();
        if (ok && lane < n_cand) {
          s_ct[rank] = ct;
          s_cb[rank] = cb;
        }
        wave_lds_sync();
        // the bound keeps every pixel sum below 2^31
        int best = static_cast<int>(lfix);
        for (int c = 0; ok && c < n_cand; ++c) {
          if (static_cast<double>(s_cb[c]) < static_cast<double>(best)) break;
          const int t = s_ct[c];
          const int ty = t / ntx, tx = t - ty * ntx;
          const int e0 = rows[max(0, ty * kTile - R)];
          const int e1 = rows[min(H - 1, ty * kTile + kTile - 1 + R) + 1];
          int ncq = 0;
          for (int eb = e0; eb < e1; eb += kWave) {
            const int e = eb + lane;
            bool keep = false;
            if (e < e1) {
              const int xi = key_x(s_key[e]);
              keep = s_I[e] != 0.0f && xi + R >= tx * kTile && xi - R <= tx * kTile + kTile - 1;
            }
            const uint64_t mk = __ballot(keep);
            if (keep) {
              const int sl = ncq + mask_rank(mk);
              if (sl < kCandSpots) s_cq[sl] = e;
            }
            ncq += __popcll(mk);
          }
          if (ncq > kCandSpots) {
            ok = false;
            break;
          }
#pragma unroll
          for (int u = 0; u < kTile * kTile / kWave; ++u) s_tile[lane + u * kWave] = 0;
          wave_lds_sync();
          // the tile's pixels from a splat of its spots, item = (spot, column):
          // fast_term's terms ((I wx) scale) wy, truncated, summed in 32 bits
          const int ty0 = ty * kTile, ty1 = min(H, ty0 + kTile) - 1;
          for (int it = lane; it < ncq * kTile; it += kWave) {
            const int e = s_cq[it / kTile], cx = it & (kTile - 1);
            const int x = tx * kTile + cx;
            const int qk = s_key[e];
            const int qy = key_y(qk), qx = key_x(qk);
            if (x >= W || x < qx - R || x > qx + R) continue;
            const float iws = s_I[e] * band_weight(x, qx, W, R, s_tp) * scale;
            int *col = s_tile + cx - ty0 * kTile;
            if (qy >= R && qy + R < H) {
#pragma unroll
              for (int dd = -R; dd <= R; ++dd) {
                const int y = qy + dd;
                if (y >= ty0 && y <= ty1) atomicAdd(col + y * kTile, fast_fixed(iws, a.taps[dd < 0 ? -dd : dd]));
              }
            } else {
              for (int y = max(ty0, qy - R); y <= min(ty1, qy + R); ++y)
                atomicAdd(col + y * kTile, fast_fixed(iws, reflect_weight3(y, qy, H, R, s_tp)));
            }
          }
          wave_lds_sync();
          int lm = best;
#pragma unroll
          for (int u = 0; u < kTile * kTile / kWave; ++u) {
            const int pix = lane + u * kWave;
            if (tx * kTile + (pix & (kTile - 1)) < W && ty0 + pix / kTile <= ty1) lm = max(lm, s_tile[pix]);
          }
          best = wave_max_i(lm);
          wave_lds_sync();  // the candidate list and the tile are rebuilt for the next tile
        }
        if (ok) out_scale = best > 0 ? 1.0f / static_cast<float>(best) : 1.0f / scale;
      }
    }
  }
  if (!ok) {
    if (lane == 0) {
      const int q = atomicAdd(b.fallback, 1);
      b.fallback[1 + q] = p;
      rec[0] = 1;  // the band writer leaves this pattern to render_kernel
    }
    return;
  }
  if (lane == 0) {
    *reinterpret_cast<int4 *>(rec) =
        make_int4(0, ns, __float_as_int(scale), __float_as_int(out_scale));
    if (a.status) a.status[p] = 0;
  }
  // each band's window of the row-sorted list: rows y0 - R .. y1 + R
  for (int bb = lane; bb < b.nb; bb += kWave) {
    const int y0 = bb * b.band_rows, y1 = min(H, y0 + b.band_rows) - 1;
    reinterpret_cast<int2 *>(rec + 4)[bb] = make_int2(rows[max(0, y0 - R)], rows[min(H - 1, y1 + R) + 1]);
  }
  int2 *ent = b.ent + static_cast<size_t>(pl) * kPlanMax;
  const int total = rows[H];
  for (int e = lane; e < total; e += kWave) {
    const int k = s_key[e];
    ent[e] = make_int2((key_y(k) << 16) | key_x(k), __float_as_int(s_I[e]));
  }
}

template <int kR>
__global__ __launch_bounds__(kBandThreads) void render_band_kernel(BandArgs b) {
  __shared__ __align__(16) racc_t s_acc[kBandPix];
  __shared__ int2 s_ent[kPlanMax];
  __shared__ float s_tp[kR + 1];
  const RenderArgs &a = b.r;
  constexpr int R = kR;
  constexpr int ncols = 2 * kR + 1;
  // every band of pattern pl runs on XCD pl mod 8 (blocks are dispatched
  // round-robin over the 8 XCDs), in order: the pattern's header and
  // row-sorted spots come from memory once and from that XCD's L2 after
  const int xcd = blockIdx.x & 7, seq = blockIdx.x >> 3;
  const int grp = seq / b.nb, bb = seq - grp * b.nb;
  const int pl = grp * 8 + xcd;
  if (pl >= b.m) return;
  const int p = b.p0 + pl;
  const int tid = threadIdx.x;
  const int H = a.H, W = a.W;
  const int y0 = bb * b.band_rows;
  const int y1 = min(H, y0 + b.band_rows) - 1;  // last row of the band
  const int nvec = (y1 - y0 + 1) * W / 4;
  // one scalar round: the pattern's header and the band's window of the
  // row-sorted list (rows y0 - R .. y1 + R)
  const int *rec = b.rec + static_cast<size_t>(pl) * b.rec_ints;
  const int4 h = *reinterpret_cast<const int4 *>(rec);
  const int2 win = reinterpret_cast<const int2 *>(rec + 4)[bb];
  const int e0 = win.x, e1 = win.y;
#pragma unroll
  for (int i = 0; i <= R; ++i)
    if (tid == i) s_tp[i] = a.taps[i];
  for (int v = tid; v < nvec; v += kBandThreads) reinterpret_cast<int4 *>(s_acc)[v] = make_int4(0, 0, 0, 0);
  if (h.x != 0) return;  // listed for render_kernel (uniform across the block)
  const int cnt = e1 - e0;
  const int2 *ent = b.ent + static_cast<size_t>(pl) * kPlanMax + e0;
  for (int q = tid; q < cnt; q += kBandThreads) s_ent[q] = ent[q];
  lds_sync();
  const float scale = __int_as_float(h.z), out_scale = __int_as_float(h.w);
  for (int it = tid; it < cnt * ncols; it += kBandThreads) {
    const int q = it / ncols, c = it - q * ncols;
    const int2 en = s_ent[q];
    const int yi = en.x >> 16, xi = en.x & 0xffff;
    const int x = xi - R + c;
    if (x < 0 || x >= W) continue;
    // render_kernel's splat terms: ((I wx) scale) wy, truncated
    const float wx = (xi >= R && xi + R < W) ? s_tp[c > R ? c - R : R - c]
                                             : reflect_weight3(x, xi, W, R, s_tp);
    const float iws = __int_as_float(en.y) * wx * scale;
    const int ya = max(yi - R, y0), yb = min(yi + R, y1);
    racc_t *col = s_acc + (x - y0 * W);
    if (yi >= R && yi + R < H) {
#pragma unroll
      for (int dd = -R; dd <= R; ++dd) {
        const int y = yi + dd;
        if (y >= ya && y <= yb) atomicAdd(col + y * W, fast_fixed(iws, a.taps[dd < 0 ? -dd : dd]));
      }
    } else {
      for (int y = ya; y <= yb; ++y) atomicAdd(col + y * W, fast_fixed(iws, reflect_weight3(y, yi, H, R, s_tp)));
    }
  }
  lds_sync();
  float4 *dst = reinterpret_cast<float4 *>(a.out + static_cast<size_t>(p) * H * W + static_cast<size_t>(y0) * W);
#pragma unroll
  for (int k = 0; k < kBandVec; ++k) {
    const int v = tid + k * kBandThreads;
    if (v < nvec) {
      const int4 u = reinterpret_cast<const int4 *>(s_acc)[v];
      store_nt(make_float4(static_cast<float>(u.x) * out_scale, static_cast<float>(u.y) * out_scale,
                           static_cast<float>(u.z) * out_scale, static_cast<float>(u.w) * out_scale),
               dst + v);
    }
  }
}

// ---------------------------------------------------------------------------
// Dense fast-mode renderer: the whole pattern lives in LDS.  For patterns
// with many spots per pixel area (large cells, wide reciprocal radii) it is
// cheaper than splatting footprints: last write wins through an LDS owner map
// (atomicMax of the spot's list position), then scipy's separable Gaussian
// with 'reflect' boundaries as two LDS passes.  No limit on spots per pattern.
// ---------------------------------------------------------------------------

constexpr int kDenseThreads = 1024;
constexpr int kDenseMaxPix = 16384;  // two FP32 H x W buffers in 128 KB of LDS
// register-blocked separable filter: two padded buffers (transposed with row
// H + 1, then row-major with row W + 1) of at most this many words each
constexpr int kDenseRegionWords = 20352;
constexpr int kDenseSeg = 16;  // outputs per thread and item of the blocked filter
constexpr int kDensePre = 6;   // spots per thread of the next pattern held in registers

__host__ __device__ __forceinline__ bool dense_blocked(int R_tmpl, int H, int W) {
  return R_tmpl > 0 && H * W + (H > W ? H : W) <= kDenseRegionWords;
}

__host__ __device__ __forceinline__ int dense_region_words(int R_tmpl, int H, int W) {
  return dense_blocked(R_tmpl, H, W) ? ((H * W + (H > W ? H : W) + 3) & ~3) : ((H * W + 3) & ~3);
}

// scipy.ndimage 'reflect' (half-sample symmetric) index for any offset
__device__ __forceinline__ int reflect_index(int x, int n) {
  const int period = 2 * n;
  int m = x % period;
  if (m < 0) m += period;
  return m < n ? m : period - 1 - m;
}

// reflect_index for |x - clamp(x)| <= R: one reflection suffices when R < n
// (selects instead of an integer modulo)
__device__ __forceinline__ int reflect_near(int x, int n, int R) {
  if (R >= n) return reflect_index(x, n);
  return x < 0 ? -x - 1 : (x >= n ? 2 * n - 1 - x : x);
}

template <int kR>
__global__ __launch_bounds__(kDenseThreads) void render_dense_kernel(RenderArgs a) {
  extern __shared__ __align__(16) float s_buf[];
  __shared__ float s_taps[kMaxR + 1];
  __shared__ float s_wmax[kDenseThreads / kWave];
  __shared__ int s_any;
  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int wid = tid / kWave;
  const int H = a.H, W = a.W, HW = a.H * a.W;
  const int R = kR > 0 ? kR : a.R;
  const bool blocked = dense_blocked(kR, H, W);
  int *A = reinterpret_cast<int *>(s_buf);  // owner map, then the first filter pass
  float *Af = s_buf;
  float *B = s_buf + dense_region_words(kR, H, W);  // spot image, then the filtered pattern
  for (int i = tid; i <= kMaxR; i += kDenseThreads) s_taps[i] = a.taps[i];

  // The block's next pattern (list entry, count, spot pointer) and the pixel
  // and intensity of its first kDensePre * kDenseThreads spots are loaded
  // while the current pattern is filtered, so a pattern starts without the
  // chain of dependent memory waits; longer lists load the rest in place.
  int nx_p = 0, nx_n = 0;
  long long nx_off = 0;
  int pk[kDensePre];    // pixel of spot tid + k kDenseThreads, -1 when out of frame
  float pw[kDensePre];  // its intensity
  auto prefetch = [&](int pi) {
#pragma unroll
    for (int k = 0; k < kDensePre; ++k) {
      pk[k] = -1;
      pw[k] = 0.0f;
    }
    if (pi >= a.n) return;
    const int p = a.list ? a.list[pi] : pi;
    nx_p = p;
    nx_n = min(a.count[p], a.capv ? a.capv[p] : a.cap);
    nx_off = a.spot_off ? a.spot_off[p] : (long long)p * a.cap;
    const float4 *sp = a.spots + nx_off;
#pragma unroll
    for (int k = 0; k < kDensePre; ++k) {
      const int j = tid + k * kDenseThreads;
      if (j < nx_n) {
        const float4 v = sp[j];
        float X, Y;
        if (to_pixel(v, a, X, Y)) {
          pk[k] = min(static_cast<int>(Y), H - 1) * W + min(static_cast<int>(X), W - 1);
          pw[k] = v.w;
        }
      }
    }
  };
  prefetch(blockIdx.x);

  for (int pi = blockIdx.x; pi < a.n; pi += gridDim.x) {
    const int p = nx_p;
    const int n = nx_n;
    const float4 *spots = a.spots + nx_off;
    for (int i = tid; i < HW; i += kDenseThreads) {
      A[i] = -1;
      B[i] = 0.0f;
    }
    if (tid == 0) s_any = 0;
    lds_sync();
    // owner of each pixel = largest list position (last write wins)
    bool any = false;
#pragma unroll
    for (int k = 0; k < kDensePre; ++k) {
      if (pk[k] >= 0) {
        any = true;
        atomicMax(&A[pk[k]], tid + k * kDenseThreads);
      }
    }
    for (int j = tid + kDensePre * kDenseThreads; j < n; j += kDenseThreads) {
      float X, Y;
      if (!to_pixel(spots[j], a, X, Y)) continue;
      any = true;
      atomicMax(&A[min(static_cast<int>(Y), H - 1) * W + min(static_cast<int>(X), W - 1)], j);
    }
    if (__ballot(any) != 0 && lane == 0) s_any = 1;
    lds_sync();
#pragma unroll
    for (int k = 0; k < kDensePre; ++k)
      if (pk[k] >= 0 && A[pk[k]] == tid + k * kDenseThreads) B[pk[k]] = pw[k];
    for (int j = tid + kDensePre * kDenseThreads; j < n; j += kDenseThreads) {
      const float4 sp = spots[j];
      float X, Y;
      if (!to_pixel(sp, a, X, Y)) continue;
      const int key = min(static_cast<int>(Y), H - 1) * W + min(static_cast<int>(X), W - 1);
      if (A[key] == j) B[key] = sp.w;
    }
    lds_sync();
    prefetch(pi + gridDim.x);  // overlaps the filter passes below
    float *out = a.out + (size_t)p * HW;
    if (tid == 0 && a.status) a.status[p] = 0;
    if (s_any == 0) {  // no spot in frame: np.zeros(shape), not normalised
      for (int i = tid; i < HW; i += kDenseThreads) __builtin_nontemporal_store(0.0f, out + i);
      lds_sync();
      continue;
    }
    float lmax = 0.0f;
    int os = W;  // row stride of the filtered pattern in B
    if (blocked) {
      // scipy's order (axis 0, then axis 1), each pass register-blocked: an
      // item loads kDenseSeg + 2R inputs along its line once and forms
      // kDenseSeg outputs.  Pass 1 reads columns of B (lanes = consecutive x)
      // and writes them transposed (Af[x (H + 1) + y]); pass 2 reads those
      // (lanes = consecutive y) and writes rows of stride W + 1: the odd
      // strides keep both LDS write patterns free of bank conflicts.
      constexpr int kS = kDenseSeg, kN = kDenseSeg + 2 * (kR > 0 ? kR : 1);
      const int Hp = H + 1, Wp = W + 1;
      const int nsy = (H + kS - 1) / kS;
      for (int it = tid; it < W * nsy; it += kDenseThreads) {
        const int x = it % W, y0 = (it / W) * kS;
        float v[kN];
        if (y0 >= kR && y0 + kS + kR <= H) {  // interior segment: no reflection
          const float *col = B + (y0 - kR) * W + x;
#pragma unroll
          for (int k = 0; k < kN; ++k) v[k] = col[k * W];
        } else {
#pragma unroll
          for (int k = 0; k < kN; ++k) {
            const int y = y0 - kR + k;
            v[k] = y < H + kR ? B[reflect_near(y, H, kR) * W + x] : 0.0f;
          }
        }
#pragma unroll
        for (int k = 0; k < kS; ++k) {
          if (y0 + k < H) {
            float o = s_taps[0] * v[k + kR];
#pragma unroll
            for (int d = 1; d <= kR; ++d) o = fmaf(s_taps[d], v[k + kR - d] + v[k + kR + d], o);
            Af[x * Hp + y0 + k] = o;
          }
        }
      }
      lds_sync();
      const int nsx = (W + kS - 1) / kS;
      for (int it = tid; it < H * nsx; it += kDenseThreads) {
        const int y = it % H, x0 = (it / H) * kS;
        float v[kN];
        if (x0 >= kR && x0 + kS + kR <= W) {  // interior segment: no reflection
          const float *row = Af + (x0 - kR) * Hp + y;
#pragma unroll
          for (int k = 0; k < kN; ++k) v[k] = row[k * Hp];
        } else {
#pragma unroll
          for (int k = 0; k < kN; ++k) {
            const int x = x0 - kR + k;
            v[k] = x < W + kR ? Af[reflect_near(x, W, kR) * Hp + y] : 0.0f;
          }
        }
#pragma unroll
        for (int k = 0; k < kS; ++k) {
          if (x0 + k < W) {
            float o = s_taps[0] * v[k + kR];
#pragma unroll
            for (int d = 1; d <= kR; ++d) o = fmaf(s_taps[d], v[k + kR - d] + v[k + kR + d], o);
            B[y * Wp + x0 + k] = o;
            lmax = fmaxf(lmax, o);
          }
        }
      }
      os = Wp;
    } else {
      // rows: Af = B (*) taps along x.  The interior columns run without
      // boundary tests; the 2R edge columns (every column when 2R >= W) go
      // through a separate loop, so no wave mixes the two paths.
      const int Wi = W - 2 * R;
      if (Wi > 0) {
        for (int i = tid; i < H * Wi; i += kDenseThreads) {
          const int y = i / Wi, x = R + (i - y * Wi);
          const float *row = B + y * W;
          float v = s_taps[0] * row[x];
#pragma unroll
          for (int d = 1; d <= (kR > 0 ? kR : kMaxR); ++d) {
            if (kR == 0 && d > R) break;
            v = fmaf(s_taps[d], row[x - d] + row[x + d], v);
          }
          Af[y * W + x] = v;
        }
      }
      const int ne = Wi > 0 ? 2 * R : W;
      for (int i = tid; i < H * ne; i += kDenseThreads) {
        const int y = i / ne, e = i - y * ne;
        const int x = Wi > 0 && e >= R ? W - 2 * R + e : e;
        const float *row = B + y * W;
        float v = s_taps[0] * row[x];
        for (int d = 1; d <= R; ++d)
          v = fmaf(s_taps[d], row[reflect_near(x - d, W, R)] + row[reflect_near(x + d, W, R)], v);
        Af[y * W + x] = v;
      }
      lds_sync();
      // columns: B = Af (*) taps along y, and the maximum
      for (int i = tid; i < HW; i += kDenseThreads) {
        const int y = i / W, x = i - y * W;
        const float *col = Af + x;
        float v = s_taps[0] * col[y * W];
        if (y >= R && y + R < H) {
#pragma unroll
          for (int d = 1; d <= (kR > 0 ? kR : kMaxR); ++d) {
            if (kR == 0 && d > R) break;
            v = fmaf(s_taps[d], col[(y - d) * W] + col[(y + d) * W], v);
          }
        } else {
          for (int d = 1; d <= R; ++d)
            v = fmaf(s_taps[d], col[reflect_near(y - d, H, R) * W] + col[reflect_near(y + d, H, R) * W], v);
        }
        B[i] = v;
        lmax = fmaxf(lmax, v);
      }
    }
    lmax = wave_max(lmax);
    if (lane == 0) s_wmax[wid] = lmax;
    lds_sync();
    float m = 0.0f;
    for (int w = 0; w < kDenseThreads / kWave; ++w) m = fmaxf(m, s_wmax[w]);
    // a zero maximum leaves the zeros as they are instead of producing 0/0
    const float sc = (a.normalize && m > 0.0f) ? 1.0f / m : 1.0f;
    if (os == W && (HW & 3) == 0) {
      for (int i = tid; i < HW / 4; i += kDenseThreads) {
        const float4 v = reinterpret_cast<const float4 *>(B)[i];
        store_nt(make_float4(v.x * sc, v.y * sc, v.z * sc, v.w * sc), reinterpret_cast<float4 *>(out) + i);
      }
    } else if ((W & 3) == 0) {
      for (int i = tid; i < HW / 4; i += kDenseThreads) {
        const int y = (4 * i) / W, x = 4 * i - y * W;
        const float *r = B + y * os + x;
        store_nt(make_float4(r[0] * sc, r[1] * sc, r[2] * sc, r[3] * sc), reinterpret_cast<float4 *>(out) + i);
      }
    } else {
      for (int i = tid; i < HW; i += kDenseThreads) {
        const int y = i / W, x = i - y * W;
        __builtin_nontemporal_store(B[y * os + x] * sc, out + i);
      }
    }
    lds_sync();
  }
}

// ---------------------------------------------------------------------------
// In-place dense fast-mode renderer: the same result as render_dense_kernel
// from one FP32 buffer per pattern, so two 512-thread blocks (two patterns)
// share a CU.  The buffer holds H rows of stride P = W + 1 (odd: lanes along
// x and lanes along y both hit distinct banks) and goes through:
//   owner map   int words, empty = INT_MIN; atomicMax of the spot's list
//               position (last write wins);
//   spot image  the owning spot writes its intensity with the sign bit
//               flipped (I >= 0 gives a negative int, so a later compare of
//               another spot's position never matches it; empty decodes to
//               -0 = 0);
//   column pass scipy's axis 0: each lane filters a segment of one column in
//               place; the 2R rows of halo it needs from the neighbouring
//               segments (or their reflections) are read into registers
//               before a barrier, its own rows as it goes;
//   row pass    axis 1, the same along a segment of one row, and the maximum;
//   output      scaled by 1/max, streamed to HBM with non-temporal stores,
//               every word reset to empty for the next pattern.
// Every output is formed in scipy's order, t0 v0 + sum_d t_d (v_-d + v_d), as
// in render_dense_kernel, so both kernels give the same bits.  The shape is a
// template parameter (every offset an immediate; with runtime strides the
// compiler hoists the halo addresses out of the pattern loop and spills):
// 128 x 128 patterns with R <= 5 (wider taps spill at four waves per SIMD),
// H and W multiples of kInplaceSeg, R < min(H, W), W <= kInplaceThreads.
// ---------------------------------------------------------------------------

constexpr int kInplaceThreads = 512;
constexpr int kInplaceSeg = 16;   // outputs per register block along a line
constexpr int kInplacePre = 4;    // spots per thread of the next pattern held in registers
constexpr int kEmpty = static_cast<int>(0x80000000u);

__host__ __device__ constexpr int inplace_segments(int lanes_per_seg, int len) {
  // segments of a line of `len` samples: as many as the block's lanes allow,
  // each a whole number of register blocks
  const int chunks = len / kInplaceSeg;
  int nseg = kInplaceThreads / lanes_per_seg;
  if (nseg > chunks) nseg = chunks;
  while (nseg > 1 && chunks % nseg != 0) --nseg;
  return nseg < 1 ? 1 : nseg;
}

__host__ __device__ __forceinline__ bool inplace_fits(int R, int H, int W) {
  return R >= 1 && R <= 5 && H % kInplaceSeg == 0 && W % kInplaceSeg == 0 && W <= kInplaceThreads &&
         H <= kInplaceThreads && R < H && R < W;
}

__device__ __forceinline__ float dec_px(int bits) { return __int_as_float(bits ^ kEmpty); }

// Filter the line segment [r0, r0 + len) of one column (stride = P) or one
// row (stride = 1) in place.  top/bot: the kR samples before and after the
// segment (reflected at the pattern edges), read before the caller's barrier.
template <int kR, bool kDecode, int stride, int len>
__device__ __forceinline__ float inplace_line(int *buf, int base, int r0,
                                              const float (&top)[kR], const float (&bot)[kR],
                                              const float *taps, float lmax) {
  constexpr int kS = kInplaceSeg, kN = kS + 2 * kR;
  float v[kN];
#pragma unroll
  for (int m = 0; m < kR; ++m) v[m] = top[m];
#pragma unroll
  for (int m = 0; m < kR; ++m) {
    const int w = buf[base + (r0 + m) * stride];
    v[kR + m] = kDecode ? dec_px(w) : __int_as_float(w);
  }
  constexpr int nc = len / kS;
#pragma unroll
  for (int c = 0; c < nc; ++c) {
    const int y0 = r0 + c * kS;
    if (c + 1 < nc) {
#pragma unroll
      for (int k = 0; k < kS; ++k) {
        const int w = buf[base + (y0 + kR + k) * stride];
        v[2 * kR + k] = kDecode ? dec_px(w) : __int_as_float(w);
      }
    } else {  // last block: the samples past the segment are the halo
#pragma unroll
      for (int k = 0; k < kS - kR; ++k) {
        const int w = buf[base + (y0 + kR + k) * stride];
        v[2 * kR + k] = kDecode ? dec_px(w) : __int_as_float(w);
      }
#pragma unroll
      for (int m = 0; m < kR; ++m) v[kS + kR + m] = bot[m];
    }
#pragma unroll
    for (int k = 0; k < kS; ++k) {
      float o = taps[0] * v[k + kR];
#pragma unroll
      for (int d = 1; d <= kR; ++d) o = fmaf(taps[d], v[k + kR - d] + v[k + kR + d], o);
      buf[base + (y0 + k) * stride] = __float_as_int(o);
      lmax = fmaxf(lmax, o);
    }
#pragma unroll
    for (int m = 0; m < 2 * kR; ++m) v[m] = v[kS + m];
  }
  return lmax;
}

template <int kR, int kH, int kW>
__global__ __launch_bounds__(kInplaceThreads, 4) void render_inplace_kernel(RenderArgs a) {
  extern __shared__ __align__(16) int s_img[];
  __shared__ float s_taps[kR + 1];
  __shared__ float s_wmax[kInplaceThreads / kWave];
  __shared__ int s_any;
  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int wid = tid / kWave;
  constexpr int H = kH, W = kW, P = kW + 1;
  const int words = H * P;
  if (tid <= kR) s_taps[tid] = a.taps[tid];
  for (int i = tid; i < words; i += kInplaceThreads) s_img[i] = kEmpty;
  if (tid == 0) s_any = 0;
  // lane roles (fixed for the launch): column pass lane (x1, segment s1),
  // row pass lane (y2, segment s2), output lane (column xo, first row yo)
  constexpr int ns1 = inplace_segments(W, H), L1 = H / ns1;
  const int x1 = tid % W, s1 = tid / W;
  constexpr int ns2 = inplace_segments(H, W), L2 = W / ns2;
  const int y2 = tid % H, s2 = tid / H;
  const int xo = tid % W, yo = tid / W;
  constexpr int ystep = kInplaceThreads / W;

  int nx_p = 0, nx_n = 0;
  long long nx_off = 0;
  int pk[kInplacePre];
  float pw[kInplacePre];
  auto prefetch = [&](int pi) {
#pragma unroll
    for (int k = 0; k < kInplacePre; ++k) {
      pk[k] = -1;
      pw[k] = 0.0f;
    }
    if (pi >= a.n) return;
    const int p = a.list ? a.list[pi] : pi;
    nx_p = p;
    nx_n = min(a.count[p], a.capv ? a.capv[p] : a.cap);
    nx_off = a.spot_off ? a.spot_off[p] : (long long)p * a.cap;
    const float4 *sp = a.spots + nx_off;
#pragma unroll
    for (int k = 0; k < kInplacePre; ++k) {
      const int j = tid + k * kInplaceThreads;
      if (j < nx_n) {
        const float4 v = sp[j];
        float X, Y;
        if (to_pixel(v, a, X, Y)) {
          pk[k] = min(static_cast<int>(Y), H - 1) * P + min(static_cast<int>(X), W - 1);
          pw[k] = v.w;
        }
      }
    }
  };
  prefetch(blockIdx.x);
  lds_sync();

  for (int pi = blockIdx.x; pi < a.n; pi += gridDim.x) {
    const int p = nx_p;
    const int n = nx_n;
    const float4 *spots = a.spots + nx_off;
    // 1. owner map
    bool any = false;
#pragma unroll
    for (int k = 0; k < kInplacePre; ++k) {
      if (pk[k] >= 0) {
        any = true;
        atomicMax(&s_img[pk[k]], tid + k * kInplaceThreads);
      }
    }
    for (int j = tid + kInplacePre * kInplaceThreads; j < n; j += kInplaceThreads) {
      float X, Y;
      if (!to_pixel(spots[j], a, X, Y)) continue;
      any = true;
      atomicMax(&s_img[min(static_cast<int>(Y), H - 1) * P + min(static_cast<int>(X), W - 1)], j);
    }
    if (__ballot(any) != 0 && lane == 0) s_any = 1;
    lds_sync();
    // 2. spot image: the owner writes its intensity, sign bit flipped
#pragma unroll
    for (int k = 0; k < kInplacePre; ++k)
      if (pk[k] >= 0 && s_img[pk[k]] == tid + k * kInplaceThreads)
        s_img[pk[k]] = __float_as_int(pw[k]) ^ kEmpty;
    for (int j = tid + kInplacePre * kInplaceThreads; j < n; j += kInplaceThreads) {
      const float4 sp = spots[j];
      float X, Y;
      if (!to_pixel(sp, a, X, Y)) continue;
      const int key = min(static_cast<int>(Y), H - 1) * P + min(static_cast<int>(X), W - 1);
      if (s_img[key] == j) s_img[key] = __float_as_int(sp.w) ^ kEmpty;
    }
    lds_sync();
    prefetch(pi + gridDim.x);  // overlaps the filter passes below
    float *out = a.out + (size_t)p * H * W;
    if (tid == 0 && a.status) a.status[p] = 0;
    if (s_any == 0) {  // no spot in frame: np.zeros(shape), not normalised
      for (int i = tid; i < H * W; i += kInplaceThreads) __builtin_nontemporal_store(0.0f, out + i);
      lds_sync();
      continue;
    }
    // 3. column pass (axis 0)
    float top[kR], bot[kR];
    const bool act1 = s1 < ns1;
    if (act1) {
      const int r0 = s1 * L1;
#pragma unroll
      for (int m = 0; m < kR; ++m) {
        top[m] = dec_px(s_img[reflect_near(r0 - kR + m, H, kR) * P + x1]);
        bot[m] = dec_px(s_img[reflect_near(r0 + L1 + m, H, kR) * P + x1]);
      }
    }
    lds_sync();
    if (act1) inplace_line<kR, true, P, L1>(s_img, x1, s1 * L1, top, bot, s_taps, 0.0f);
    lds_sync();
    // 4. row pass (axis 1) and the maximum
    const bool act2 = s2 < ns2;
    if (act2) {
      const int c0 = s2 * L2;
#pragma unroll
      for (int m = 0; m < kR; ++m) {
        top[m] = __int_as_float(s_img[y2 * P + reflect_near(c0 - kR + m, W, kR)]);
        bot[m] = __int_as_float(s_img[y2 * P + reflect_near(c0 + L2 + m, W, kR)]);
      }
    }
    lds_sync();
    float lmax = 0.0f;
    if (act2) lmax = inplace_line<kR, false, 1, L2>(s_img, y2 * P, s2 * L2, top, bot, s_taps, 0.0f);
    lmax = wave_max(lmax);
    if (lane == 0) s_wmax[wid] = lmax;
    lds_sync();
    float m = 0.0f;
    for (int w = 0; w < kInplaceThreads / kWave; ++w) m = fmaxf(m, s_wmax[w]);
    // a zero maximum leaves the zeros as they are instead of producing 0/0
    const float sc = (a.normalize && m > 0.0f) ? 1.0f / m : 1.0f;
    // 5. output, and the buffer back to empty
    if (yo < ystep) {
      for (int y = yo; y < H; y += ystep) {
        const int i = y * P + xo;
        __builtin_nontemporal_store(__int_as_float(s_img[i]) * sc, out + y * W + xo);
        s_img[i] = kEmpty;
      }
    }
    if (tid == 0) s_any = 0;
    lds_sync();
  }
}

// ---------------------------------------------------------------------------
// Windowed dense fast-mode renderer for large patterns with many spots (more
// pixels than the dense kernel holds and more spots than the splat kernel
// holds).  In-frame spots are binned by their source row strip into a
// caller workspace; every output strip then rebuilds, in LDS, the spot image
// of its window (the strip plus R rows of halo each side: all source rows a
// reflected tap can reach), resolves last-write-wins with an owner map,
// filters it along x and then along y (scipy's separable Gaussian with
// 'reflect' boundaries) and writes the strip.  Pass A finds the maximum.
// ---------------------------------------------------------------------------

constexpr int kLargeThreads = 1024;
constexpr int kLargeWindowBytes = 128 * 1024;

struct LargeArgs {
  RenderArgs r;
  int *bins;             // workspace: cap ints per block
  long long per_block;
  int rows;              // output rows per strip
  int n_src;             // source strips (= output strips)
};

template <int kR>
__global__ __launch_bounds__(kLargeThreads) void render_fast_large_kernel(LargeArgs la) {
  extern __shared__ __align__(16) float s_win[];
  __shared__ float s_taps[kMaxR + 1];
  __shared__ int s_start[kMaxStrips + 1];
  __shared__ int s_fill[kMaxStrips];
  __shared__ float s_wmax[kLargeThreads / kWave];
  __shared__ int s_any;
  const RenderArgs &a = la.r;
  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int wid = tid / kWave;
  const int H = a.H, W = a.W;
  const int R = kR > 0 ? kR : a.R;
  const int rows = la.rows;
  const int win_max = rows + 2 * R;
  int *owner = reinterpret_cast<int *>(s_win);  // window spot image (owner, then value)
  float *img = s_win;
  float *tmp = s_win + win_max * W;              // row-filtered window
  int *bins = la.bins + (size_t)blockIdx.x * la.per_block;
  for (int i = tid; i <= kMaxR; i += kLargeThreads) s_taps[i] = a.taps[i];

  for (int pi = blockIdx.x; pi < a.n; pi += gridDim.x) {
    const int p = a.list ? a.list[pi] : pi;
    const int n = min(a.count[p], a.capv ? a.capv[p] : a.cap);
    const float4 *spots = a.spots + (a.spot_off ? a.spot_off[p] : (long long)p * a.cap);
    for (int s = tid; s <= la.n_src; s += kLargeThreads) s_start[s] = 0;
    if (tid == 0) s_any = 0;
    __syncthreads();
    // bins of in-frame spots by source row strip
    bool any = false;
    for (int j = tid; j < n; j += kLargeThreads) {
      float X, Y;
      if (!to_pixel(spots[j], a, X, Y)) continue;
      any = true;
      atomicAdd(&s_start[min(static_cast<int>(Y), H - 1) / rows + 1], 1);
    }
    if (__ballot(any) != 0 && lane == 0) s_any = 1;
    __syncthreads();
    if (tid == 0) {
      int acc = 0;
      for (int s = 0; s < la.n_src; ++s) {
        acc += s_start[s + 1];
        s_start[s + 1] = acc;
      }
    }
    __syncthreads();
    for (int s = tid; s < la.n_src; s += kLargeThreads) s_fill[s] = s_start[s];
    __syncthreads();
    for (int j = tid; j < n; j += kLargeThreads) {
      float X, Y;
      if (!to_pixel(spots[j], a, X, Y)) continue;
      bins[atomicAdd(&s_fill[min(static_cast<int>(Y), H - 1) / rows], 1)] = j;
    }
    __syncthreads();  // bins in global memory, block-visible
    float *out = a.out + (size_t)p * H * W;
    if (tid == 0 && a.status) a.status[p] = 0;
    if (s_any == 0) {  // no spot in frame: np.zeros(shape), not normalised
      for (int i = tid; i < H * W; i += kLargeThreads) __builtin_nontemporal_store(0.0f, out + i);
      __syncthreads();
      continue;
    }
    float sc = 1.0f;
    for (int pass = a.normalize ? 0 : 1; pass < 2; ++pass) {
      float lmax = 0.0f;
      for (int s = 0; s < la.n_src; ++s) {
        const int y0 = s * rows, y1 = min(H, y0 + rows);
        const int wy0 = max(0, y0 - R), wy1 = min(H, y1 + R);
        const int wn = (wy1 - wy0) * W;
        for (int i = tid; i < wn; i += kLargeThreads) owner[i] = -1;
        __syncthreads();
        const int e0 = s_start[wy0 / rows], e1 = s_start[(wy1 - 1) / rows + 1];
        for (int e = e0 + tid; e < e1; e += kLargeThreads) {
          const int j = bins[e];
          float X, Y;
          to_pixel(spots[j], a, X, Y);
          const int yi = min(static_cast<int>(Y), H - 1), xi = min(static_cast<int>(X), W - 1);
          if (yi >= wy0 && yi < wy1) atomicMax(&owner[(yi - wy0) * W + xi], j);
        }
        __syncthreads();
        for (int i = tid; i < wn; i += kLargeThreads) {
          const int o = owner[i];
          img[i] = o >= 0 ? spots[o].w : 0.0f;
        }
        __syncthreads();
        for (int i = tid; i < wn; i += kLargeThreads) {
          const int r = i / W, x = i - r * W;
          const float *row = img + r * W;
          float v = s_taps[0] * row[x];
          if (x >= R && x + R < W) {
            for (int d = 1; d <= R; ++d) v = fmaf(s_taps[d], row[x - d] + row[x + d], v);
          } else {
            for (int d = 1; d <= R; ++d)
              v = fmaf(s_taps[d], row[reflect_index(x - d, W)] + row[reflect_index(x + d, W)], v);
          }
          tmp[i] = v;
        }
        __syncthreads();
        const int cnt = (y1 - y0) * W;
        for (int i = tid; i < cnt; i += kLargeThreads) {
          const int y = y0 + i / W, x = i % W;
          const float *col = tmp + x;
          float v = s_taps[0] * col[(y - wy0) * W];
          for (int d = 1; d <= R; ++d)
            v = fmaf(s_taps[d], col[(reflect_index(y - d, H) - wy0) * W] +
                                    col[(reflect_index(y + d, H) - wy0) * W], v);
          if (pass == 0)
            lmax = fmaxf(lmax, v);
          else
            __builtin_nontemporal_store(v * sc, out + (size_t)y0 * W + i);
        }
        __syncthreads();
      }
      if (pass == 0) {
        lmax = wave_max(lmax);
        if (lane == 0) s_wmax[wid] = lmax;
        __syncthreads();
        float m = 0.0f;
        for (int w = 0; w < kLargeThreads / kWave; ++w) m = fmaxf(m, s_wmax[w]);
        // a zero maximum leaves the zeros as they are instead of producing 0/0
        sc = m > 0.0f ? 1.0f / m : 1.0f;
        __syncthreads();
      }
    }
  }
}

static int fill_render_args(int mode, const dsx_render_params *p, RenderArgs &a, const char *name,
                            int strip_bytes = kStripBytes) {
  if (p->height < 1 || p->width < 1 || p->height > 32767 || p->width > 32767 ||
      !(p->calibration > 0.0f) || !(p->sigma > 0.0f)) {
    set_error("%s: bad pattern shape %dx%d, calibration or sigma", name, p->height, p->width);
    return DSX_EINVAL;
  }
  if (mode == kModeSubpixel && !(p->clip_threshold > 0.0f)) {
    set_error("%s: clip_threshold must be positive", name);
    return DSX_EINVAL;
  }
  if (mode == kModeFast && (p->radius < 0 || p->radius > kMaxR)) {
    set_error("%s: Gaussian radius %d exceeds the supported maximum %d", name, p->radius, kMaxR);
    return DSX_EINVAL;
  }
  a.spot_off = nullptr;
  a.capv = nullptr;
  a.list = nullptr;
  a.n_dev = nullptr;
  a.H = p->height;
  a.W = p->width;
  a.cal = p->calibration;
  a.cx = p->center_x;
  a.cy = p->center_y;
  a.ca = p->cos_angle;
  a.sa = p->sin_angle;
  a.mirrored = p->mirrored;
  a.normalize = p->normalize;
  a.R = mode == kModeFast ? p->radius : 0;
  const double sg = p->sigma;
  a.pref = static_cast<float>(1.0 / (2.0 * M_PI * sg * sg));
  a.ep = static_cast<float>(-1.0 / (2.0 * sg * sg));
  a.clip = p->clip_threshold;
  for (int i = 0; i <= kMaxR; ++i) a.taps[i] = (mode == kModeFast && i <= p->radius) ? p->taps[i] : 0.0f;
  const int budget = strip_bytes / static_cast<int>(sizeof(racc_t));
  if (a.W > budget) {
    set_error("%s: pattern width %d exceeds the LDS strip budget", name, a.W);
    return DSX_EINVAL;
  }
  const int rows = std::max(1, std::min(a.H, budget / a.W));
  const int n_strips = (a.H + rows - 1) / rows;
  if (n_strips > kMaxStrips) {
    set_error("%s: pattern height %d needs %d strips (> %d)", name, a.H, n_strips, kMaxStrips);
    return DSX_EINVAL;
  }
  a.strip_rows = rows;
  a.n_strips = n_strips;
  const size_t lds = std::max(static_cast<size_t>(rows) * a.W * sizeof(racc_t),
                              static_cast<size_t>(2 * kHash) * sizeof(int));
  a.lds_words = static_cast<int>(lds / sizeof(racc_t));
  return DSX_OK;
}

// bpc: blocks per CU as resolved by splat_blocks_per_cu (the strips of `a`
// were sized for it by fill_render_args)
static int launch_splat(int mode, const RenderArgs &a, int bpc, hipStream_t st, const char *name) {
  const size_t lds = static_cast<size_t>(a.lds_words) * sizeof(racc_t);
  const int blocks = std::min(a.n, cu_count() * bpc);
  if (mode == kModeSubpixel) {
    hipLaunchKernelGGL((render_kernel<kModeSubpixel, 0>), dim3(blocks), dim3(kRenderThreads), lds, st, a);
  } else {
    switch (a.R) {  // radii of sigma <= 2 px get an unrolled kernel
#define DSX_RENDER_CASE(r)                                                                     \
  case r:                                                                                      \
    hipLaunchKernelGGL((render_kernel<kModeFast, r>), dim3(blocks), dim3(kRenderThreads), lds, st, a); \
    break;
      DSX_RENDER_CASE(1)
      DSX_RENDER_CASE(2)
      DSX_RENDER_CASE(3)
      DSX_RENDER_CASE(4)
      DSX_RENDER_CASE(5)
      DSX_RENDER_CASE(6)
      DSX_RENDER_CASE(7)
      DSX_RENDER_CASE(8)
#undef DSX_RENDER_CASE
      default:
        hipLaunchKernelGGL((render_kernel<kModeFast, 0>), dim3(blocks), dim3(kRenderThreads), lds, st, a);
    }
  }
  return check_launch(name);
}

template <int kR>
static int launch_dense_r(const RenderArgs &a, int blocks, size_t lds, hipStream_t st) {
  // the LDS limit is raised once per kernel to its maximum, so launches that
  // a caller records in a graph issue no attribute calls
  static int reserved = -1;
  if (reserved < 0) {
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(&render_dense_kernel<kR>),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(2 * kDenseRegionWords * sizeof(float))) != hipSuccess) {
      (void)hipGetLastError();
      set_error("dsx_render_fast_ragged: cannot reserve %zu bytes of LDS", lds);
      return DSX_ELAUNCH;
    }
    reserved = 1;
  }
  hipLaunchKernelGGL(render_dense_kernel<kR>, dim3(blocks), dim3(kDenseThreads), lds, st, a);
  return DSX_OK;
}

template <int kR, int kH, int kW>
static int launch_inplace_r(const RenderArgs &a, int blocks, size_t lds, hipStream_t st) {
  static int reserved = -1;  // raised once, so graph-recorded launches issue no attribute calls
  constexpr int bytes = kH * (kW + 1) * static_cast<int>(sizeof(int));
  if (reserved < 0) {
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(&render_inplace_kernel<kR, kH, kW>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess) {
      (void)hipGetLastError();
      set_error("dsx_render_fast_ragged: cannot reserve %d bytes of LDS", bytes);
      return DSX_ELAUNCH;
    }
    reserved = 1;
  }
  hipLaunchKernelGGL((render_inplace_kernel<kR, kH, kW>), dim3(blocks), dim3(kInplaceThreads), lds, st, a);
  return DSX_OK;
}

static int launch_inplace(const RenderArgs &a, int blocks_per_cu, hipStream_t st, const char *name) {
  if (!(inplace_fits(a.R, a.H, a.W) && a.H == 128 && a.W == 128)) {
    set_error("%s: the in-place dense renderer takes 128x128 patterns with tap radius 1..5 "
              "(got %dx%d, radius %d)", name, a.H, a.W, a.R);
    return DSX_EINVAL;
  }
  const size_t lds = static_cast<size_t>(a.H) * (a.W + 1) * sizeof(int);
  const int per_cu = std::max(1, std::min(blocks_per_cu > 0 ? blocks_per_cu : 2,
                                          static_cast<int>((160 * 1024) / (lds + 1024))));
  const int blocks = std::min(a.n, cu_count() * per_cu);
  int st_ = DSX_OK;
  switch (a.R) {
    case 1: st_ = launch_inplace_r<1, 128, 128>(a, blocks, lds, st); break;
    case 2: st_ = launch_inplace_r<2, 128, 128>(a, blocks, lds, st); break;
    case 3: st_ = launch_inplace_r<3, 128, 128>(a, blocks, lds, st); break;
    case 4: st_ = launch_inplace_r<4, 128, 128>(a, blocks, lds, st); break;
    default: st_ = launch_inplace_r<5, 128, 128>(a, blocks, lds, st);
  }
  if (st_ != DSX_OK) return st_;
  return check_launch(name);
}

static int launch_dense(const RenderArgs &a, int blocks_per_cu, hipStream_t st, const char *name) {
  const int HW = a.H * a.W;
  if (HW > kDenseMaxPix) {
    set_error("%s: the dense renderer holds at most %d pixels (got %dx%d)", name, kDenseMaxPix, a.H, a.W);
    return DSX_EINVAL;
  }
  const size_t lds = 2 * static_cast<size_t>(dense_region_words(a.R <= 8 ? a.R : 0, a.H, a.W)) * sizeof(float);
  const int per_cu = std::max(1, std::min(blocks_per_cu > 0 ? blocks_per_cu : 1,
                                          static_cast<int>((160 * 1024) / (lds + 1024))));
  const int blocks = std::min(a.n, cu_count() * per_cu);
  int st_ = DSX_OK;
  switch (a.R) {
    case 1: st_ = launch_dense_r<1>(a, blocks, lds, st); break;
    case 2: st_ = launch_dense_r<2>(a, blocks, lds, st); break;
    case 3: st_ = launch_dense_r<3>(a, blocks, lds, st); break;
    case 4: st_ = launch_dense_r<4>(a, blocks, lds, st); break;
    case 5: st_ = launch_dense_r<5>(a, blocks, lds, st); break;
    case 6: st_ = launch_dense_r<6>(a, blocks, lds, st); break;
    case 7: st_ = launch_dense_r<7>(a, blocks, lds, st); break;
    case 8: st_ = launch_dense_r<8>(a, blocks, lds, st); break;
    default: st_ = launch_dense_r<0>(a, blocks, lds, st);
  }
  if (st_ != DSX_OK) return st_;
  return check_launch(name);
}

static int render_common(int mode, const float *spot_d, const int *count_d, int cap, int n,
                         const dsx_render_params *p, float *out_d, int *status_d, void *stream,
                         const char *name) {
  clear_error();
  if (!p || n < 0 || cap < 1) {
    set_error("%s: bad arguments (n=%d cap=%d)", name, n, cap);
    return DSX_EINVAL;
  }
  if (n == 0) return DSX_OK;
  if (!spot_d || !count_d || !out_d) {
    set_error("%s: null device pointer", name);
    return DSX_EINVAL;
  }
  RenderArgs a;
  const int bpc = splat_blocks_per_cu(mode, p->normalize != 0, 0);
  const int st = fill_render_args(mode, p, a, name, strip_bytes_for(bpc));
  if (st != DSX_OK) return st;
  a.spots = reinterpret_cast<const float4 *>(spot_d);
  a.count = count_d;
  a.cap = cap;
  a.n = n;
  a.out = out_d;
  a.status = status_d;
  return launch_splat(mode, a, bpc, as_stream(stream), name);
}

// ---- band renderer: host side ----------------------------------------------

static void band_shape(int H, int W, int &band_rows, int &nb) {
  band_rows = std::max(1, std::min(H, kBandPix / std::max(1, W)));
  nb = (H + band_rows - 1) / band_rows;
}

// Shapes the plan and the band writer take: tap radius 1..8 below both
// pattern sides, rows of whole float4 vectors, sides of at most 1024 pixels
// (10-bit coordinates in the plan's sort key), tile bounds in the plan's LDS.
static bool band_fits(const RenderArgs &a) {
  const int n_tiles = ((a.W + kTile - 1) / kTile) * ((a.H + kTile - 1) / kTile);
  return a.R >= 1 && a.R <= 8 && a.R < std::min(a.H, a.W) && (a.W & 3) == 0 &&
         a.W <= kPlanMaxSide && a.H <= kPlanMaxSide && n_tiles <= kMaxTiles;
}

static long long band_list_ints(int n) { return (1LL + n + 3) & ~3LL; }
static int band_rec_ints(int nb) { return (4 + 2 * nb + 3) & ~3; }

// Workspace: the list of patterns left to render_kernel (count + ids, all
// rounds) | per pattern of a round: header and each band's window | row-sorted spots.
static long long band_workspace_ints(int n, int H, int W) {
  int br, nb;
  band_shape(H, W, br, nb);
  const long long chunk = std::min(std::max(n, 1), kBandChunk);
  return band_list_ints(std::max(n, 1)) + chunk * band_rec_ints(nb) + chunk * kPlanMax * 2;
}

__global__ void band_list_reset_kernel(int *count) {
  if (threadIdx.x == 0) *count = 0;
}

template <int kR>
static void launch_bands_r(const BandArgs &b, size_t plan_lds, hipStream_t st) {
  const int plan_blocks = (b.m + 8 * kPlanWaves - 1) / (8 * kPlanWaves) * 8;
  hipLaunchKernelGGL(render_plan_kernel<kR>, dim3(plan_blocks), dim3(kPlanThreads), plan_lds, st, b);
  const unsigned band_blocks = static_cast<unsigned>((b.m + 7) / 8) * 8u * static_cast<unsigned>(b.nb);
  hipLaunchKernelGGL(render_band_kernel<kR>, dim3(band_blocks), dim3(kBandThreads), 0, st, b);
}

}  // namespace dsx

using namespace dsx;

extern "C" int dsx_render_max_spots(void) { return kMaxSpots; }

extern "C" int dsx_render_fast(const float *spot_d, const int *count_d, int cap, int n,
                               const dsx_render_params *params, float *out_d, int *status_d,
                               void *stream) {
  return render_common(kModeFast, spot_d, count_d, cap, n, params, out_d, status_d, stream,
                       "dsx_render_fast");
}

extern "C" int dsx_render_fast_ragged(const float *spot_d, const long long *spot_off_d,
                                      const int *cap_d, const int *count_d, const int *list_d, int n,
                                      const dsx_render_params *params, int method, int blocks_per_cu,
                                      float *out_d, int *status_d, void *stream) {
  const char *name = "dsx_render_fast_ragged";
  clear_error();
  if (!params || n < 0 || blocks_per_cu < 0 ||
      (method != DSX_RENDER_SPLAT && method != DSX_RENDER_DENSE &&
       method != DSX_RENDER_DENSE_INPLACE)) {
    set_error("%s: bad arguments (n=%d method=%d)", name, n, method);
    return DSX_EINVAL;
  }
  if (n == 0) return DSX_OK;
  if (!spot_d || !spot_off_d || !cap_d || !count_d || !list_d || !out_d) {
    set_error("%s: null device pointer", name);
    return DSX_EINVAL;
  }
  RenderArgs a;
  const int bpc = splat_blocks_per_cu(kModeFast, params->normalize != 0, blocks_per_cu);
  const int st = fill_render_args(kModeFast, params, a, name,
                                  method == DSX_RENDER_SPLAT ? strip_bytes_for(bpc) : kStripBytes);
  if (st != DSX_OK) return st;
  a.spots = reinterpret_cast<const float4 *>(spot_d);
  a.count = count_d;
  a.cap = 1;
  a.n = n;
  a.spot_off = spot_off_d;
  a.capv = cap_d;
  a.list = list_d;
  a.out = out_d;
  a.status = status_d;
  if (method == DSX_RENDER_DENSE_INPLACE)
    return launch_inplace(a, blocks_per_cu, as_stream(stream), name);
  return method == DSX_RENDER_DENSE ? launch_dense(a, blocks_per_cu, as_stream(stream), name)
                                    : launch_splat(kModeFast, a, bpc, as_stream(stream), name);
}

extern "C" long long dsx_render_fast_bands_workspace_ints(int n, int height, int width) {
  if (n < 0 || height < 1 || width < 1) return -1;
  return band_workspace_ints(n, height, width);
}

extern "C" int dsx_render_fast_bands(const float *spot_d, const int *count_d, int cap, int n,
                                     const dsx_render_params *params, int *work_d,
                                     long long work_ints, float *out_d, int *status_d,
                                     void *stream) {
  const char *name = "dsx_render_fast_bands";
  clear_error();
  if (!params || n < 0 || cap < 1 || work_ints < 0) {
    set_error("%s: bad arguments (n=%d cap=%d)", name, n, cap);
    return DSX_EINVAL;
  }
  if (n == 0) return DSX_OK;
  if (!spot_d || !count_d || !out_d) {
    set_error("%s: null device pointer", name);
    return DSX_EINVAL;
  }
  RenderArgs a;
  const int bpc = splat_blocks_per_cu(kModeFast, params->normalize != 0, 0);
  const int st = fill_render_args(kModeFast, params, a, name, strip_bytes_for(bpc));
  if (st != DSX_OK) return st;
  a.spots = reinterpret_cast<const float4 *>(spot_d);
  a.count = count_d;
  a.cap = cap;
  a.n = n;
  a.out = out_d;
  a.status = status_d;
  hipStream_t s = as_stream(stream);
  if (!band_fits(a)) return launch_splat(kModeFast, a, bpc, s, name);
  const long long need = band_workspace_ints(n, a.H, a.W);
  if (!work_d || work_ints < need) {
    set_error("%s: workspace of %lld ints is smaller than %lld", name, work_ints, need);
    return DSX_EINVAL;
  }
  BandArgs b;
  b.r = a;
  band_shape(a.H, a.W, b.band_rows, b.nb);
  b.ntx = (a.W + kTile - 1) / kTile;
  b.n_tiles = b.ntx * ((a.H + kTile - 1) / kTile);
  b.wave_ints = plan_wave_ints(a.H, b.n_tiles);
  const int chunk = std::min(n, kBandChunk);
  b.fallback = work_d;
  b.rec_ints = band_rec_ints(b.nb);
  b.rec = work_d + band_list_ints(n);
  b.ent = reinterpret_cast<int2 *>(b.rec + static_cast<long long>(chunk) * b.rec_ints);
  const size_t plan_lds = static_cast<size_t>(kPlanWaves) * b.wave_ints * sizeof(int);
  hipLaunchKernelGGL(band_list_reset_kernel, dim3(1), dim3(64), 0, s, b.fallback);
  for (int c0 = 0; c0 < n; c0 += chunk) {
    b.p0 = c0;
    b.m = std::min(chunk, n - c0);
    switch (a.R) {
      case 1: launch_bands_r<1>(b, plan_lds, s); break;
      case 2: launch_bands_r<2>(b, plan_lds, s); break;
      case 3: launch_bands_r<3>(b, plan_lds, s); break;
      case 4: launch_bands_r<4>(b, plan_lds, s); break;
      case 5: launch_bands_r<5>(b, plan_lds, s); break;
      case 6: launch_bands_r<6>(b, plan_lds, s); break;
      case 7: launch_bands_r<7>(b, plan_lds, s); break;
      default: launch_bands_r<8>(b, plan_lds, s);
    }
  }
  // the patterns the plans listed, by the splat renderer (blocks without a
  // listed pattern exit at once)
  RenderArgs f = a;
  f.list = b.fallback + 1;
  f.n_dev = b.fallback;
  return launch_splat(kModeFast, f, bpc, s, name);
}

extern "C" int dsx_render_dense_max_pixels(void) { return kDenseMaxPix; }

extern "C" int dsx_render_dense_inplace_fits(int height, int width, int radius) {
  return inplace_fits(radius, height, width) && height == 128 && width == 128 ? 1 : 0;
}

extern "C" long long dsx_render_fast_large_workspace_ints(int cap) {
  return cap < 1 ? 0 : static_cast<long long>(cu_count()) * cap;
}

template <int kR>
static int launch_large_r(const LargeArgs &la, int blocks, size_t lds, hipStream_t st) {
  static int reserved = -1;  // raised once, outside any graph capture
  if (reserved < 0) {
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(&render_fast_large_kernel<kR>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, kLargeWindowBytes) != hipSuccess) {
      (void)hipGetLastError();
      set_error("dsx_render_fast_large: cannot reserve %zu bytes of LDS", lds);
      return DSX_ELAUNCH;
    }
    reserved = 1;
  }
  hipLaunchKernelGGL(render_fast_large_kernel<kR>, dim3(blocks), dim3(kLargeThreads), lds, st, la);
  return DSX_OK;
}

extern "C" int dsx_render_fast_large(const float *spot_d, const int *count_d, int cap, int n,
                                     const dsx_render_params *params, int *work_d,
                                     long long work_ints, float *out_d, int *status_d,
                                     void *stream) {
  const char *name = "dsx_render_fast_large";
  clear_error();
  if (!params || n < 0 || cap < 1 || work_ints < 0) {
    set_error("%s: bad arguments (n=%d cap=%d)", name, n, cap);
    return DSX_EINVAL;
  }
  if (n == 0) return DSX_OK;
  if (!spot_d || !count_d || !out_d || !work_d) {
    set_error("%s: null device pointer", name);
    return DSX_EINVAL;
  }
  LargeArgs la;
  RenderArgs &a = la.r;
  const int st = fill_render_args(kModeFast, params, a, name);
  if (st != DSX_OK) return st;
  a.spots = reinterpret_cast<const float4 *>(spot_d);
  a.count = count_d;
  a.cap = cap;
  a.n = n;
  a.out = out_d;
  a.status = status_d;
  // window = strip rows + 2R halo rows, two FP32 buffers of the full width
  const int W = a.W, R = a.R;
  const int win_rows = kLargeWindowBytes / (2 * 4 * W);
  const int rows = win_rows - 2 * R;
  if (rows < 1) {
    set_error("%s: width %d with tap radius %d leaves no room for a strip", name, W, R);
    return DSX_EINVAL;
  }
  if (R >= a.H || R >= W) {
    set_error("%s: tap radius %d must be smaller than the pattern", name, R);
    return DSX_EINVAL;
  }
  la.rows = std::min(rows, a.H);
  la.n_src = (a.H + la.rows - 1) / la.rows;
  if (la.n_src > kMaxStrips) {
    set_error("%s: pattern height %d needs %d strips (> %d)", name, a.H, la.n_src, kMaxStrips);
    return DSX_EINVAL;
  }
  const int blocks = std::min(n, cu_count());
  if (work_ints < static_cast<long long>(blocks) * cap) {
    set_error("%s: workspace of %lld ints is smaller than %lld", name, work_ints,
              static_cast<long long>(blocks) * cap);
    return DSX_EINVAL;
  }
  la.bins = work_d;
  la.per_block = cap;
  const size_t lds = 2 * static_cast<size_t>(la.rows + 2 * R) * W * sizeof(float);
  hipStream_t s = as_stream(stream);
  int rc;
  switch (R) {
    case 1: rc = launch_large_r<1>(la, blocks, lds, s); break;
    case 2: rc = launch_large_r<2>(la, blocks, lds, s); break;
    case 3: rc = launch_large_r<3>(la, blocks, lds, s); break;
    case 4: rc = launch_large_r<4>(la, blocks, lds, s); break;
    default: rc = launch_large_r<0>(la, blocks, lds, s);
  }
  if (rc != DSX_OK) return rc;
  return check_launch(name);
}

extern "C" int dsx_render_subpixel(const float *spot_d, const int *count_d, int cap, int n,
                                   const dsx_render_params *params, float *out_d, int *status_d,
                                   void *stream) {
  return render_common(kModeSubpixel, spot_d, count_d, cap, n, params, out_d, status_d, stream,
                       "dsx_render_subpixel");
}
