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
//   3. pass A (normalize only): each strip is splatted into LDS with ds_add and
//      reduced to the pattern maximum;
//   4. pass B: each strip is splatted again and streamed to HBM scaled by
//      1/max with non-temporal 16-byte stores; strips without spots are
//      written as zeros straight from registers.
#include <algorithm>

#include "dsx_internal.h"

namespace dsx {

constexpr int kRenderThreads = 256;
constexpr int kMaxSpots = 1024;  // distinct in-frame spots per pattern
constexpr int kHash = 2048;      // LDS hash slots (load factor <= 1/2)
constexpr int kMaxStrips = 256;
constexpr int kBinCap = 4096;    // strip-bin entries per pattern
constexpr int kMaxR = 64;
constexpr int kStripBytes = 32 * 1024;

enum { kModeFast = 0, kModeSubpixel = 1 };

struct RenderArgs {
  const float4 *spots;
  const int *count;
  int cap;
  int n;
  int H, W;
  float cal, cx, cy, ca, sa;
  int mirrored, normalize;
  int R;           // fast: tap radius
  float pref, ep;  // subpixel: 1/(2 pi s^2), -1/(2 s^2)
  float clip;
  int strip_rows;
  int n_strips;
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

// Pixel position of a lab-frame spot (Simulation2D._get_transformed_coordinates,
// simulation2d.py:261-285, in pixel units).
__device__ __forceinline__ bool to_pixel(const float4 &sp, const RenderArgs &a, float &X, float &Y) {
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
  int rows_max;
  float red[kRenderThreads / kWave];
};

template <int kMode>
__global__ __launch_bounds__(kRenderThreads) void render_kernel(RenderArgs a) {
  extern __shared__ float s_dyn[];  // strip buffer; hash table before the passes
  __shared__ RenderShared<kMode> sh;
  int *h_key = reinterpret_cast<int *>(s_dyn);
  int *h_val = h_key + kHash;
  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int wid = tid / kWave;
  const int W = a.W, H = a.H;
  const int R = a.R;
  const bool vec4 = (W & 3) == 0;
  const bool short_taps = R < min(H, W);
  for (int i = tid; i <= kMaxR; i += kRenderThreads) sh.taps[i] = a.taps[i];

  for (int p = blockIdx.x; p < a.n; p += gridDim.x) {
    const int n = min(a.count[p], a.cap);
    const float4 *spots = a.spots + (size_t)p * a.cap;
    if (kMode == kModeFast) {
      for (int i = tid; i < kHash; i += kRenderThreads) {
        h_key[i] = -1;
        h_val[i] = -1;
      }
    }
    if (tid == 0) {
      sh.ns = 0;
      sh.overflow = 0;
      sh.rows_max = 0;
    }
    for (int s = tid; s <= kMaxStrips; s += kRenderThreads) sh.bin_start[s] = 0;
    lds_sync();

    // ---- 1a. last write wins: owner of each pixel = largest list position
    if (kMode == kModeFast) {
      for (int j = tid; j < n; j += kRenderThreads) {
        float X, Y;
        if (!to_pixel(spots[j], a, X, Y)) continue;
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
    for (int j = tid; j < n; j += kRenderThreads) {
      const float4 sp = spots[j];
      float X, Y;
      if (!to_pixel(sp, a, X, Y)) continue;
      int lo, hi, packed;
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
      } else {
        const double amp = static_cast<double>(sp.w) * static_cast<double>(a.pref);
        // radius = sqrt(log(clip / amp) / ep); NaN (amp < clip) skips the spot
        const double rad2 = log(static_cast<double>(a.clip) / amp) / static_cast<double>(a.ep);
        if (!(amp > 0.0) || !(rad2 >= 0.0)) continue;
        const double rad = sqrt(fmin(rad2, 1e12));
        lo = max(0, static_cast<int>(ceil(static_cast<double>(Y) - rad)));
        hi = min(H, static_cast<int>(floor(static_cast<double>(Y) + rad + 1.0))) - 1;
        const int xlo = max(0, static_cast<int>(ceil(static_cast<double>(X) - rad)));
        const int xhi = min(W, static_cast<int>(floor(static_cast<double>(X) + rad + 1.0))) - 1;
        if (hi < lo || xhi < xlo) continue;
        packed = (xlo << 16) | xhi;
        fxv = X;
        fyv = Y;
      }
      const int q = atomicAdd(&sh.ns, 1);
      if (q >= kMaxSpots) {
        sh.overflow = 1;
        continue;
      }
      sh.I[q] = sp.w;
      sh.xy[q] = packed;
      if (kMode == kModeSubpixel) {  // fx/fy have one slot in fast mode
        sh.fx[q] = fxv;
        sh.fy[q] = fyv;
      }
      sh.ylo[q] = static_cast<short>(lo);
      sh.yhi[q] = static_cast<short>(hi);
      atomicMax(&sh.rows_max, hi - lo + 1);
      for (int s = lo / a.strip_rows; s <= hi / a.strip_rows; ++s) atomicAdd(&sh.bin_start[s + 1], 1);
    }
    lds_sync();
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
      lds_sync();
      continue;
    }

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
    // footprint rows per item slot: next power of two of the tallest footprint
    int rows_shift = 0;
    while ((1 << rows_shift) < sh.rows_max) ++rows_shift;
    const int rows_mask = (1 << rows_shift) - 1;

    float scale = 1.0f;
    for (int pass = a.normalize ? 0 : 1; pass < 2; ++pass) {
      float lmax = 0.0f;
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
        if (nb == 0) {
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
        // zero the LDS strip
        if (vec4) {
          for (int i = tid; i < cnt / 4; i += kRenderThreads)
            reinterpret_cast<float4 *>(s_dyn)[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        } else {
          for (int i = tid; i < cnt; i += kRenderThreads) s_dyn[i] = 0.0f;
        }
        lds_sync();
        // splat: item = (spot in bin, footprint row)
        const int n_items = nb << rows_shift;
        for (int it = tid; it < n_items; it += kRenderThreads) {
          const int bi = it >> rows_shift;
          const int q = binned ? sh.bin[b0 + bi] : bi;
          const int y = sh.ylo[q] + (it & rows_mask);
          if (y > sh.yhi[q] || y < y0 || y >= y1) continue;
          const float I = sh.I[q];
          float *row = s_dyn + (y - y0) * W;
          if (kMode == kModeFast) {
            const int yi = sh.xy[q] >> 16, xi = sh.xy[q] & 0xffff;
            const int xlo = max(0, xi - R), xhi = min(W - 1, xi + R);
            if (short_taps) {
              const float wy = I * reflect_weight3(y, yi, H, R, sh.taps);
              if (xi >= R && xi + R < W) {
                for (int x = xlo; x <= xhi; ++x) atomicAdd(row + x, wy * tap_at(x - xi, R, sh.taps));
              } else {
                for (int x = xlo; x <= xhi; ++x)
                  atomicAdd(row + x, wy * reflect_weight3(x, xi, W, R, sh.taps));
              }
            } else {
              const float wy = I * reflect_weight(y, yi, H, R, sh.taps);
              for (int x = xlo; x <= xhi; ++x) atomicAdd(row + x, wy * reflect_weight(x, xi, W, R, sh.taps));
            }
          } else {
            const float cy = sh.fy[q], cx = sh.fx[q];
            const float dy = static_cast<float>(y) - cy;
            const int xlo = sh.xy[q] >> 16, xhi = sh.xy[q] & 0xffff;
            const float amp = I * a.pref;
            for (int x = xlo; x <= xhi; ++x) {
              const float dx = static_cast<float>(x) - cx;
              atomicAdd(row + x, amp * expf(a.ep * (dy * dy + dx * dx)));
            }
          }
        }
        lds_sync();
        if (pass == 0) {
          if (vec4) {
            for (int i = tid; i < cnt / 4; i += kRenderThreads) {
              const float4 v = reinterpret_cast<const float4 *>(s_dyn)[i];
              lmax = fmaxf(lmax, fmaxf(fmaxf(v.x, v.y), fmaxf(v.z, v.w)));
            }
          } else {
            for (int i = tid; i < cnt; i += kRenderThreads) lmax = fmaxf(lmax, s_dyn[i]);
          }
        } else {
          if (vec4) {
            for (int i = tid; i < cnt / 4; i += kRenderThreads) {
              float4 v = reinterpret_cast<const float4 *>(s_dyn)[i];
              v.x *= scale;
              v.y *= scale;
              v.z *= scale;
              v.w *= scale;
              store_nt(v, reinterpret_cast<float4 *>(dst) + i);
            }
          } else {
            for (int i = tid; i < cnt; i += kRenderThreads) __builtin_nontemporal_store(s_dyn[i] * scale, dst + i);
          }
        }
        lds_sync();
      }
      if (pass == 0) {
        lmax = wave_max(lmax);
        if (lane == 0) sh.red[wid] = lmax;
        lds_sync();
        float m = 0.0f;
        for (int w = 0; w < kRenderThreads / kWave; ++w) m = fmaxf(m, sh.red[w]);
        // a zero maximum (every spot has zero intensity) leaves the zeros as
        // they are instead of producing 0/0
        scale = m > 0.0f ? 1.0f / m : 1.0f;
        lds_sync();
      }
    }
    // every thread is done with this pattern's LDS state before the next one
    lds_sync();
  }
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
  RenderArgs a;
  a.spots = reinterpret_cast<const float4 *>(spot_d);
  a.count = count_d;
  a.cap = cap;
  a.n = n;
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
  const int budget = kStripBytes / 4;
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
  a.out = out_d;
  a.status = status_d;
  const size_t lds = std::max(static_cast<size_t>(rows) * a.W * sizeof(float),
                              static_cast<size_t>(2 * kHash) * sizeof(int));
  const int blocks = std::min(n, cu_count() * 2);
  if (mode == kModeFast)
    hipLaunchKernelGGL(render_kernel<kModeFast>, dim3(blocks), dim3(kRenderThreads), lds,
                       as_stream(stream), a);
  else
    hipLaunchKernelGGL(render_kernel<kModeSubpixel>, dim3(blocks), dim3(kRenderThreads), lds,
                       as_stream(stream), a);
  return check_launch(name);
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

extern "C" int dsx_render_subpixel(const float *spot_d, const int *count_d, int cap, int n,
                                   const dsx_render_params *params, float *out_d, int *status_d,
                                   void *stream) {
  return render_common(kModeSubpixel, spot_d, count_d, cap, n, params, out_d, status_d, stream,
                       "dsx_render_subpixel");
}
