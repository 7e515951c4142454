// Physics shared by the per-orientation kernel and the adjoint: the shape
// factor models of diffsims/utils/shape_factor_models.py, their derivatives,
// the Ewald-sphere membership test of get_intersecting_reflections
// (diffsims/generators/simulation_generator.py:355-375) and the atomic
// scattering factor parameterisations (diffsims/utils/sim_utils.py:227-253).
//
// The functions take any argument struct that carries the fields filled by
// fill_physics() (k, k2, me, width, sigma, fac, model, precession, phi, Pz, Pt).
#pragma once

#include <math.h>

#include "diffsims_amd.h"

namespace dsx {

template <class Args>
inline void fill_physics(Args &a, const dsx_sim_params *p) {
  const double k = p->k;
  a.k = k;
  a.k2 = k * k;
  a.me = p->max_excitation_error;
  const double w = p->shape_width;
  a.width = w;
  a.sigma = M_PI / w;
  const bool atanc = p->shape_model == DSX_SF_ATANC || p->shape_model == DSX_SF_PRECESSION_AVG + DSX_SF_ATANC;
  a.fac = M_PI * p->minima_number / (atanc ? fabs(w) : w);
  a.model = p->shape_model;
  const double phi = p->precession_angle;
  a.precession = phi > 0.0;
  a.phi = phi;
  a.Pz = k * cos(phi);
  a.Pt = k * sin(phi);
}

// The models of shape_factor_models.py in plain FP64 (the inner function of
// the full precession average).
template <class Args>
__device__ __forceinline__ double base_model_f64(int m, double x, const Args &a) {
  switch (m) {
    case DSX_SF_LINEAR:
      return fmax(1.0 - fabs(x) / a.width, 0.0);
    case DSX_SF_LORENTZIAN: {
      const double t = a.sigma * x;
      return 1.0 / (1.0 + t * t);
    }
    case DSX_SF_SINC:
    case DSX_SF_SIN2C: {
      const double y = a.fac * x;
      if (y == 0.0) return 0.0;  // the reference's where=denom != 0 leaves 0
      const double v = fabs(sin(y) / y);
      return m == DSX_SF_SINC ? v : v * v;
    }
    case DSX_SF_ATANC: {
      const double y = a.fac * x;
      return y == 0.0 ? 1.0 : atan(y) / y;
    }
    case DSX_SF_BINARY:
      return 1.0;
    default:
      return 0.0;
  }
}

// (1/pi) * integral over [t0, t1] of f(s + A cos t) dt by 8-point Gauss-Legendre.
template <class Args>
__device__ __forceinline__ double gl8_piece(int m, double s, double A, double t0, double t1,
                                            const Args &a) {
  const double xs[4] = {0.18343464249564980, 0.52553240991632899, 0.79666647741362674,
                        0.96028985649753623};
  const double ws[4] = {0.36268378337836198, 0.31370664587788729, 0.22238103445337447,
                        0.10122853629037626};
  const double c = 0.5 * (t0 + t1), h = 0.5 * (t1 - t0);
  double acc = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    acc += ws[i] * (base_model_f64(m, s + A * cos(c - h * xs[i]), a) +
                    base_model_f64(m, s + A * cos(c + h * xs[i]), a));
  }
  return acc * h;
}

// Full precession average (1/2pi) * integral_0^2pi f(s + A cos t) dt, A = r phi
// (_shape_factor_precession, shape_factor_models.py:222-268, which integrates
// with scipy.integrate.quad).  By symmetry it is (1/pi) * integral_0^pi.
//  * lorentzian: closed form Re[1 / sqrt((1 + i sigma s)^2 + sigma^2 A^2)]
//    (the average of Re[1 / (1 + i sigma x)] with x = s + A cos t);
//  * linear: exact, piece by piece between the angles where x = -w, 0, w;
//  * sinc, sin2c: Gauss-Legendre on each lobe between consecutive zeros of
//    sin(fac x) (the kinks of |sinc| and the oscillation of both);
//  * atanc: Gauss-Legendre on 8 equal pieces (smooth integrand);
//  * binary: 1.
template <class Args>
__device__ double precession_average(int m, double s, double A, const Args &a) {
  if (!(A > 0.0)) return base_model_f64(m, s, a);
  switch (m) {
    case DSX_SF_BINARY:
      return 1.0;
    case DSX_SF_LORENTZIAN: {
      const double sg = a.sigma;
      const double re = 1.0 - sg * sg * s * s + sg * sg * A * A, im = 2.0 * sg * s;
      const double mod = hypot(re, im);
      // principal square root w of re + i im without cancellation; Re(1/w) = Re(w) / |z|
      double wr;
      if (re >= 0.0) {
        wr = sqrt(0.5 * (mod + re));
      } else {
        const double wi = sqrt(0.5 * (mod - re));
        wr = fabs(im) / (2.0 * wi);
      }
      return wr / mod;
    }
    case DSX_SF_LINEAR: {
      const double w = fabs(a.width);
      double brk[3];
      int nb = 0;
      const double xb[3] = {w, 0.0, -w};  // x decreases as t grows: t ascending
      for (int i = 0; i < 3; ++i) {
        const double c = (xb[i] - s) / A;
        if (c > -1.0 && c < 1.0) brk[nb++] = acos(c);
      }
      double total = 0.0, t0 = 0.0;
      for (int i = 0; i <= nb; ++i) {
        const double t1 = i < nb ? brk[i] : M_PI;
        if (t1 > t0) {
          const double xm = s + A * cos(0.5 * (t0 + t1));
          if (fabs(xm) < w) {
            // 1 -+ x / w with x = s + A cos t: c0 + c1 cos t
            const double sgn = xm >= 0.0 ? -1.0 : 1.0;
            const double c0 = 1.0 + sgn * s / w, c1 = sgn * A / w;
            total += c0 * (t1 - t0) + c1 * (sin(t1) - sin(t0));
          }
        }
        t0 = t1;
      }
      return total / M_PI;
    }
    case DSX_SF_SINC:
    case DSX_SF_SIN2C: {
      const double fa = fabs(a.fac);
      const double kmax = floor((s + A) * fa / M_PI), kmin = ceil((s - A) * fa / M_PI);
      double total = 0.0, t0 = 0.0;
      if (kmax - kmin > 8192.0) {  // an extreme excursion: uniform pieces
        for (int i = 0; i < 4096; ++i) {
          const double t1 = M_PI * (i + 1) / 4096.0;
          total += gl8_piece(m, s, A, t0, t1, a);
          t0 = t1;
        }
        return total / M_PI;
      }
      for (double k = kmax; k >= kmin; k -= 1.0) {
        const double c = fmin(1.0, fmax(-1.0, (k * M_PI / fa - s) / A));
        const double t1 = acos(c);
        if (t1 > t0) total += gl8_piece(m, s, A, t0, t1, a);
        t0 = fmax(t0, t1);
      }
      if (M_PI > t0) total += gl8_piece(m, s, A, t0, M_PI, a);
      return total / M_PI;
    }
    case DSX_SF_ATANC: {
      double total = 0.0;
      for (int i = 0; i < 8; ++i) total += gl8_piece(m, s, A, M_PI * i / 8.0, M_PI * (i + 1) / 8.0, a);
      return total / M_PI;
    }
    default:
      return 0.0;
  }
}

// shape_factor_models.py in FP64 (evaluated for candidates only).  kAvg
// compiles in the full precession averages (DSX_SF_PRECESSION_AVG + m), kept
// out of the common kernels for their register pressure.
template <bool kAvg = false, class Args>
__device__ __forceinline__ double shape_factor(double s, double r, const Args &a) {
  switch (a.model) {
    case DSX_SF_LINEAR:
      return fmax(1.0 - fabs(s) / a.width, 0.0);
    case DSX_SF_LORENTZIAN: {
      const double t = a.sigma * s;
      return 1.0 / (1.0 + t * t);
    }
    // the transcendental models take their FP64 argument rounded to FP32
    // (relative error ~1e-7 of the factor)
    case DSX_SF_SINC:
    case DSX_SF_SIN2C: {
      const float x = static_cast<float>(a.fac * s);
      if (x == 0.0f) return 0.0;  // the reference's where=denom != 0 leaves 0
      const double v = fabsf(sinf(x) / x);
      return a.model == DSX_SF_SINC ? v : v * v;
    }
    case DSX_SF_ATANC: {
      const float x = static_cast<float>(a.fac * s);
      return x == 0.0f ? 1.0 : static_cast<double>(atanf(x) / x);  // nan_to_num(0/0) = 1
    }
    case DSX_SF_BINARY:
      return 1.0;
    case DSX_SF_LORENTZIAN_PRECESSION: {
      // (sigma/pi) sqrt(2 (u + z) / z^2), u = sigma^2 (r^2 phi^2 - s^2) + 1,
      // z = sqrt(u^2 + 4 sigma^2 s^2); u + z rewritten for u < 0.
      const double A = r * a.phi;
      const double sg2 = a.sigma * a.sigma;
      const double u = sg2 * ((A - s) * (A + s)) + 1.0;
      const double ss = 4.0 * sg2 * s * s;
      const double z = sqrt(u * u + ss);
      const double upz = u >= 0.0 ? u + z : ss / (z - u);
      return (a.sigma * 0.31830988618379067) * sqrt(2.0 * upz) / z;
    }
    default:
      if (kAvg && a.model >= DSX_SF_PRECESSION_AVG)
        return precession_average(a.model - DSX_SF_PRECESSION_AVG, s, r * a.phi, a);
      return 0.0;
  }
}

// Membership test of simulation_generator.py:355-375 in FP64 with the
// cancellations removed:
//   k - sqrt(k^2 - r^2)            = r^2 / (k + sqrt(k^2 - r^2))
//   Pz - sqrt(k^2 - (r +- Pt)^2)   = r (r +- 2 Pt) / (Pz + sqrt(k^2 - (r +- Pt)^2))
template <class Args>
__device__ __forceinline__ bool intersects(double x, double y, double z, const Args &a,
                                           double &s, double &r) {
  const double r2 = x * x + y * y;
  r = sqrt(r2);
  if (!(r2 < a.k2)) return false;
  s = r2 / (a.k + sqrt(a.k2 - r2)) - z;
  if (!a.precession) return fabs(s) < a.me;
  const double rp = r + a.Pt, rm = r - a.Pt;
  const double qp = a.k2 - rp * rp, qm = a.k2 - rm * rm;
  if (!(qp >= 0.0) || !(qm >= 0.0)) return false;
  const double up = r * (r + 2.0 * a.Pt) / (a.Pz + sqrt(qp));
  const double dn = r * (r - 2.0 * a.Pt) / (a.Pz + sqrt(qm));
  return (z - a.me <= up) && (z + a.me >= dn);
}

// Shape factor f(s, r) with its partial derivatives df/ds and df/dr (r only
// enters the precession model).  The value matches shape_factor(); the
// derivatives are the closed forms of each model, in FP64.
template <class Args>
__device__ __forceinline__ double shape_factor_grad(double s, double r, const Args &a, double &fs,
                                                    double &fr) {
  fs = 0.0;
  fr = 0.0;
  switch (a.model) {
    case DSX_SF_LINEAR: {
      const double v = 1.0 - fabs(s) / a.width;
      if (v > 0.0) fs = (s > 0.0 ? -1.0 : (s < 0.0 ? 1.0 : 0.0)) / a.width;
      return fmax(v, 0.0);
    }
    case DSX_SF_LORENTZIAN: {
      const double t = a.sigma * s;
      const double f = 1.0 / (1.0 + t * t);
      fs = -2.0 * a.sigma * t * f * f;
      return f;
    }
    case DSX_SF_SINC:
    case DSX_SF_SIN2C: {
      const double x = a.fac * s;
      if (static_cast<float>(x) == 0.0f) return 0.0;
      const double sx = sin(x), cx = cos(x);
      const double v = sx / x;
      const double dv = (x * cx - sx) / (x * x) * a.fac;  // dv/ds
      if (a.model == DSX_SF_SINC) {
        fs = v < 0.0 ? -dv : dv;
      } else {
        fs = 2.0 * v * dv;
      }
      return shape_factor(s, r, a);
    }
    case DSX_SF_ATANC: {
      const double x = a.fac * s;
      if (static_cast<float>(x) != 0.0f) fs = (x / (1.0 + x * x) - atan(x)) / (x * x) * a.fac;
      return shape_factor(s, r, a);
    }
    case DSX_SF_BINARY:
      return 1.0;
    case DSX_SF_LORENTZIAN_PRECESSION: {
      // ln f = const + ln(u + z)/2 - ln z; with d(u + z) = du (u + z)/z + 4 sg2 s ds / z
      // the u < 0 branch stays free of cancellation.
      const double A = r * a.phi;
      const double sg2 = a.sigma * a.sigma;
      const double u = sg2 * ((A - s) * (A + s)) + 1.0;
      const double ss = 4.0 * sg2 * s * s;
      const double z = sqrt(u * u + ss);
      const double upz = u >= 0.0 ? u + z : ss / (z - u);
      const double f = (a.sigma * 0.31830988618379067) * sqrt(2.0 * upz) / z;
      if (upz > 0.0) {
        const double du_s = -2.0 * sg2 * s, du_r = 2.0 * sg2 * A * a.phi;
        const double dz_s = (u * du_s + 4.0 * sg2 * s) / z, dz_r = u * du_r / z;
        const double dl_s = 0.5 * (du_s / z + 4.0 * sg2 * s / (z * upz)) - dz_s / z;
        const double dl_r = 0.5 * du_r / z - dz_r / z;
        fs = f * dl_s;
        fr = f * dl_r;
      }
      return f;
    }
    default:
      return 0.0;
  }
}

// f(g^2) in FP64: the Lobato parameterisation sums terms of opposite sign that
// are up to ~100x larger than the result (e.g. He: a2 = -62.0, a3 = +64.0), so
// FP32 evaluation loses 2+ digits.  Once per element per reflection; cheap.
__device__ __forceinline__ double atomic_f(const double *c, double g2, int scatter) {
  if (scatter == DSX_SCATTER_NONE) return 1.0;
  double f = 0.0;
  if (scatter == DSX_SCATTER_LOBATO) {
#pragma unroll
    for (int t = 0; t < 5; ++t) {
      const double bg = c[2 * t + 1] * g2;
      const double d = 1.0 + bg;
      f += c[2 * t] * (2.0 + bg) / (d * d);
    }
  } else {
#pragma unroll
    for (int t = 0; t < 5; ++t) f += c[2 * t] * exp(-0.25 * c[2 * t + 1] * g2);
  }
  return f;
}

// atomic_f and its derivative with respect to g^2:
//   lobato  d/dq a (2 + b q)/(1 + b q)^2 = -a b (3 + b q)/(1 + b q)^3
//   xtables d/dq a exp(-b q/4)         = -(a b / 4) exp(-b q/4)
__device__ __forceinline__ double atomic_f_grad(const double *c, double g2, int scatter, double &df) {
  df = 0.0;
  if (scatter == DSX_SCATTER_NONE) return 1.0;
  double f = 0.0;
  if (scatter == DSX_SCATTER_LOBATO) {
#pragma unroll
    for (int t = 0; t < 5; ++t) {
      const double b = c[2 * t + 1];
      const double bg = b * g2;
      const double d = 1.0 + bg;
      const double id2 = 1.0 / (d * d);
      f += c[2 * t] * (2.0 + bg) * id2;
      df -= c[2 * t] * b * (3.0 + bg) * id2 / d;
    }
  } else {
#pragma unroll
    for (int t = 0; t < 5; ++t) {
      const double e = exp(-0.25 * c[2 * t + 1] * g2);
      f += c[2 * t] * e;
      df -= 0.25 * c[2 * t] * c[2 * t + 1] * e;
    }
  }
  return f;
}

}  // namespace dsx
