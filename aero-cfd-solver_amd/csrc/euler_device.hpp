// euler_device.hpp -- pointwise Euler numerics for gfx950 kernels (2D), templated on the scalar.
//
// Each function follows the reference CPU function it cites, operation by operation, so that with
// -ffp-contract=off a kernel reproduces the FP64 evaluation of the reference algorithm.  The reference
// carries a third velocity component w; in 2D it is exactly zero everywhere (w = rhow/rho = 0,
// n_z = 0, limiter(0, 0) = 0), and every place it enters adds an exact +0 or multiplies by 0, so it is
// dropped here.  States are 4-vectors: primitive (rho, u, v, p), conservative (rho, rhou, rhov, rhoE).
//
// The scalar T is `double` for the residual and `Dual` (dual.hpp) for tangents: the same code then
// carries one directional derivative alongside every value, and branches (limiter cases, HLLC
// fallbacks, max / min / clamp) follow the values, so the result is the exact derivative of the
// branch the value took.
#pragma once

#include <hip/hip_runtime.h>

#include "dual.hpp"

namespace aero_dev {

template <class T> struct P4T { T r, u, v, p; };
template <class T> struct C4T { T r, ru, rv, E; };
using P4 = P4T<double>;
using C4 = C4T<double>;

constexpr double kFloor = 1.0e-8;      // rho, p and wave-speed floors (euler_flux.cpp:9-11)
constexpr double kMinScale = 1.0e-4;   // acoustic-scale clamp (euler_flux.cpp:12)
constexpr double kAsMachFloor = 1.0e-6;  // all-speed Mach floor (euler_flux.cpp:13)

// std::max / std::min / std::clamp (libstdc++ operand order).
__host__ __device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }
__host__ __device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__host__ __device__ __forceinline__ double dclamp(double v, double lo, double hi) {
  return (v < lo) ? lo : (hi < v) ? hi : v;
}
__host__ __device__ __forceinline__ double dmax4(double a, double b, double c, double d) {
  double r = a;
  if (r < b) r = b;
  if (r < c) r = c;
  if (r < d) r = d;
  return r;
}
__host__ __device__ __forceinline__ Dual dmax(const Dual& a, const Dual& b) { return (a < b) ? b : a; }
__host__ __device__ __forceinline__ Dual dmin(const Dual& a, const Dual& b) { return (b < a) ? b : a; }
__host__ __device__ __forceinline__ Dual dclamp(const Dual& v, const Dual& lo, const Dual& hi) {
  return (v < lo) ? lo : (hi < v) ? hi : v;
}
__host__ __device__ __forceinline__ Dual dmax4(const Dual& a, const Dual& b, const Dual& c, const Dual& d) {
  Dual r = a;
  if (r < b) r = b;
  if (r < c) r = c;
  if (r < d) r = d;
  return r;
}
__device__ __forceinline__ bool fin(double x) { return __builtin_isfinite(x); }
__device__ __forceinline__ bool fin(const Dual& x) { return __builtin_isfinite(x.v); }
__device__ __forceinline__ double vsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ double vabs(double x) { return fabs(x); }

// max / min for the FAST face fluxes: there every operand is finite (the lanes whose operands could be
// NaN fail the flux's bounds and are discarded) and no two operands are zeros of opposite sign, so the
// single-instruction IEEE maxNum / minNum give the bits of std::max / std::min (dmax / dmin).
template <bool FAST, class T> __device__ __forceinline__ T fmx(T a, T b) {
  if constexpr (FAST) return __builtin_fmax(a, b);
  else return dmax(a, b);
}
template <bool FAST, class T> __device__ __forceinline__ T fmn(T a, T b) {
  if constexpr (FAST) return __builtin_fmin(a, b);
  else return dmin(a, b);
}

// Quotients that share a denominator.  On gfx950, x / y in double is the compiler's IEEE expansion:
// v_div_scale of both operands, a reciprocal of the scaled denominator refined by two Newton steps, the
// scaled quotient, one residual correction (v_div_fmas) and v_div_fixup.  v_div_scale returns its
// operand unchanged, and v_div_fmas is a plain fma, unless an operand is zero, denormal or tiny, 1 / y
// or x / y leaves the normal range, or the exponents of x and y lie 768 or more apart.  For |x| and |y|
// in [2^-300, 2^300] none of that happens, so the refined reciprocal depends on y alone: quot_den forms
// it once, quot takes each numerator through the remaining steps, and the result is the correctly
// rounded x / y, the same bits as the division.  Callers establish the operand range.
constexpr double kQuotHi = 0x1p300;
struct QuotDen { double y, r; };
__device__ __forceinline__ QuotDen quot_den(double y) {
  double r = __builtin_amdgcn_rcp(y);
  double e = __builtin_fma(-y, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-y, r, 1.0);
  r = __builtin_fma(r, e, r);
  return {y, r};
}
__device__ __forceinline__ double quot(double x, const QuotDen& d) {
  const double q = x * d.r;
  const double e = __builtin_fma(-d.y, q, x);
  return __builtin_amdgcn_div_fixup(__builtin_fma(e, d.r, q), d.y, x);
}

// Residual-path options that vary per call (EulerResidualAssemblyConfig, residual_assembly.hpp:12-31).
// The fields that depend on the freestream (state, stabilisation floor, preconditioner bounds) take the
// scalar type, so a tangent can carry their derivatives with respect to Mach and angle of attack.
template <class T> struct FaceOptsT {
  C4T<T> ff;             // freestream conservative state
  double g;              // gamma
  int low_mach_fix;
  double low_mach_cutoff;
  T stab_floor;
  int char_bc;
  int precond;
  T mach_ref, mach_min, beta_min, beta_max;
  int as_fix;
  double as_cut, as_fmin, as_w;
  const double* frozen;  // frozen limiter ratios [8][F] (beyond the reference) or null
  const unsigned short* frozen_slots;  // the same per face slot [S][8] in units of kRatioStep (flux kernels)
};

// Frozen limiter ratios (beyond the reference) are held at a resolution of 2^-15: a captured ratio
// r in [0, 1] is stored as the nearest multiple of 2^-15 (ties to even), so the flux kernels can stream it
// as a 16-bit integer q with ratio = q * 2^-15 exactly.
constexpr double kRatioUnits = 32768.0;
constexpr double kRatioStep = 1.0 / 32768.0;
__host__ __device__ __forceinline__ double quantize_ratio(double r) {
  const double c = (r < 0.0) ? 0.0 : (1.0 < r) ? 1.0 : r;  // NaN -> 0
  return (c == c ? rint(c * kRatioUnits) : 0.0) * kRatioStep;
}
using FaceOpts = FaceOptsT<double>;

// The face functions below take a compile-time PLAIN switch: PLAIN instantiations are the ones for
// options with the low-Mach preconditioner, the low-Mach dissipation fix and the all-speed HLLC fix all
// off (plain_options), where those branches are known not taken; the arithmetic is the same.  They also
// require 1 < gamma <= 30, the range the exact shortcuts of the PLAIN face flux (flux_of_states) assume.
template <class T> __host__ __device__ __forceinline__ bool plain_options(const FaceOptsT<T>& o) {
  return !o.precond && !o.low_mach_fix && !o.as_fix && o.g > 1.0 && o.g <= 30.0;
}

template <class T> __device__ __forceinline__ C4T<T> p2c(const P4T<T>& w, double g) {  // euler_flux.cpp:101-113
  C4T<T> c;
  c.r = dmax(w.r, kFloor);
  c.ru = c.r * w.u;
  c.rv = c.r * w.v;
  const T ke = 0.5 * c.r * (w.u * w.u + w.v * w.v);
  const T p = dmax(w.p, kFloor);
  c.E = p / (g - 1.0) + ke;
  return c;
}

template <class T> __device__ __forceinline__ P4T<T> c2p(const C4T<T>& c, double g) {  // euler_flux.cpp:115-127
  P4T<T> w;
  w.r = dmax(c.r, kFloor);
  const T inv = 1.0 / w.r;
  w.u = c.ru * inv;
  w.v = c.rv * inv;
  const T ke = 0.5 * w.r * (w.u * w.u + w.v * w.v);
  w.p = dmax((g - 1.0) * (c.E - ke), kFloor);
  return w;
}

__device__ __forceinline__ void enforce_physical(C4* s, double g) {  // euler_flux.cpp:129-144
  s->r = dmax(s->r, kFloor);
  const double inv = 1.0 / s->r;
  const double u = s->ru * inv;
  const double v = s->rv * inv;
  const double ke = 0.5 * s->r * (u * u + v * v);
  s->E = dmax(s->E, ke + kFloor / (g - 1.0));
  const double p = (g - 1.0) * (s->E - ke);
  if (p < kFloor) s->E = ke + kFloor / (g - 1.0);
}

template <class T> __device__ __forceinline__ T sound(const P4T<T>& w, double g) {  // euler_flux.cpp:150-152
  return vsqrt(dmax(g * w.p / dmax(w.r, kFloor), T(0.0)));
}

// The same three maps with their quotients taken through quot (exact shortcuts of PLAIN face fluxes,
// whose callers bound the operands): p2c_q divides by gamma - 1 (dg), c2p_q by the floored density and
// returns that density's QuotDen, sound_q divides by it again (dmax(w.r, kFloor) is w.r there).
__device__ __forceinline__ C4 p2c_q(const P4& w, double g, const QuotDen& dg) {
  C4 c;
  c.r = dmax(w.r, kFloor);
  c.ru = c.r * w.u;
  c.rv = c.r * w.v;
  const double ke = 0.5 * c.r * (w.u * w.u + w.v * w.v);
  const double p = dmax(w.p, kFloor);
  c.E = quot(p, dg) + ke;
  return c;
}
__device__ __forceinline__ P4 c2p_q(const C4& c, double g, QuotDen* d) {
  P4 w;
  w.r = fmx<true>(c.r, kFloor);
  *d = quot_den(w.r);
  const double inv = quot(1.0, *d);
  w.u = c.ru * inv;
  w.v = c.rv * inv;
  const double ke = 0.5 * w.r * (w.u * w.u + w.v * w.v);
  w.p = fmx<true>((g - 1.0) * (c.E - ke), kFloor);
  return w;
}
// c2p of a cell state with 1 / rho through quot: exact for densities up to 2^300 (at least kFloor
// after the floor); *ok is cleared for one above the bound or NaN.
__device__ __forceinline__ P4 c2p_cell_q(const C4& c, double g, bool* ok) {
  P4 w;
  w.r = dmax(c.r, kFloor);
  *ok = *ok & (w.r <= kQuotHi);
  const double inv = quot(1.0, quot_den(w.r));
  w.u = c.ru * inv;
  w.v = c.rv * inv;
  const double ke = 0.5 * w.r * (w.u * w.u + w.v * w.v);
  w.p = dmax((g - 1.0) * (c.E - ke), kFloor);
  return w;
}
__device__ __forceinline__ double sound_q(const P4& w, double g, const QuotDen& d) {
  return vsqrt(fmx<true>(quot(g * w.p, d), 0.0));
}

template <class T> struct SpeedsT { T lm, lp, sr; };

// FAST (PLAIN face fluxes, where s is exactly 1): b = b2 = 1, so ut = (0.5 * 2) * un = un and disc =
// (+-0) + 4 RN(a * a), whose square root halved is RN(sqrt(RN(a * a))) = a in binary floating point while
// a * a stays normal and un is finite, which the callers' operand bounds guarantee: lm = un - a and
// lp = un + a exactly.
template <bool FAST = false, class T>
__device__ __forceinline__ SpeedsT<T> pspeeds(T un, T a, T s) {  // euler_flux.cpp:58-74
  if constexpr (FAST) {
    SpeedsT<T> sp;
    sp.lm = un - a;
    sp.lp = un + a;
    sp.sr = fmx<true>(fmx<true>(fmx<true>(vabs(un), vabs(sp.lm)), vabs(sp.lp)), T(kFloor));
    return sp;
  }
  const T b = dclamp(s, T(kMinScale), T(1.0));
  const T b2 = b * b;
  const T omb2 = 1.0 - b2;
  const T ut = 0.5 * (1.0 + b2) * un;
  const T disc = omb2 * omb2 * un * un + 4.0 * b2 * a * a;
  const T ct = 0.5 * vsqrt(dmax(disc, T(0.0)));
  SpeedsT<T> sp;
  sp.lm = ut - ct;
  sp.lp = ut + ct;
  sp.sr = dmax4(vabs(un), vabs(sp.lm), vabs(sp.lp), T(kFloor));
  return sp;
}

template <class T> __device__ __forceinline__ T local_mach(const P4T<T>& p, double g) {  // preconditioning.cpp:141-146
  const T a = dmax(sound(p, g), kFloor);
  const T vm = vsqrt(p.u * p.u + p.v * p.v);
  return vm / a;
}

template <class T>
__device__ __forceinline__ T cell_beta(const P4T<T>& p, const FaceOptsT<T>& o) {  // preconditioning.cpp:134-156
  const T lm = local_mach(p, o.g);
  const T t = dmax4(lm, o.mach_ref, o.mach_min, T(kFloor));
  const T bmin = dmax(o.beta_min, 1.0e-4);
  const T bmax = dmax(bmin, o.beta_max);
  return dclamp(t, bmin, bmax);
}

template <bool PLAIN = false, class T>
__device__ __forceinline__ T face_beta(const C4T<T>& l, const C4T<T>& r, const FaceOptsT<T>& o) {  // :158-173
  if (PLAIN || !o.precond) return T(1.0);
  const P4T<T> pl = c2p(l, o.g);
  const P4T<T> pr = c2p(r, o.g);
  P4T<T> a;
  a.r = 0.5 * (pl.r + pr.r);
  a.u = 0.5 * (pl.u + pr.u);
  a.v = 0.5 * (pl.v + pr.v);
  a.p = 0.5 * (pl.p + pr.p);
  return cell_beta(a, o);
}

template <class T>
__device__ __forceinline__ T low_mach_scale(const C4T<T>& l, const C4T<T>& r, const FaceOptsT<T>& o) {  // euler_flux.cpp:154-175
  if (o.low_mach_cutoff <= 0.0) return T(1.0);
  const P4T<T> pl = c2p(l, o.g);
  const P4T<T> pr = c2p(r, o.g);
  const T al = dmax(sound(pl, o.g), kFloor);
  const T ar = dmax(sound(pr, o.g), kFloor);
  const T vl = vsqrt(pl.u * pl.u + pl.v * pl.v);
  const T vr = vsqrt(pr.u * pr.u + pr.v * pr.v);
  const T ml = vl / al;
  const T mr = vr / ar;
  const T fm = 0.5 * (ml + mr);
  const T ms = dmax(fm, dmax(o.mach_ref, kFloor));
  const T ratio = ms / dmax(o.low_mach_cutoff, kFloor);
  return dclamp(ratio, T(kMinScale), T(1.0));
}

template <class T> __device__ __forceinline__ T smooth_max(T a, T b) {  // residual_assembly.cpp:12-20
  const T d = a - b;
  return 0.5 * (a + b + vsqrt(d * d + 1.0e-6 * 1.0e-6));
}

// Face acoustic scale (residual_assembly.cpp:31-51).
template <bool PLAIN = false, class T>
__device__ __forceinline__ T acoustic_scale(const C4T<T>& l, const C4T<T>& r, T fb, const FaceOptsT<T>& o) {
  // With the preconditioner and the low-Mach fix off, s stays 1 and max(1, clamp(m, 1e-4, 1)) is 1 (a NaN
  // m also gives 1): the stabilisation floor acts only through those two options.
  if (PLAIN) return T(1.0);
  T s = 1.0;
  if (!PLAIN && o.precond) s = dclamp(fb, T(1.0e-4), T(1.0));
  if (!PLAIN && o.low_mach_fix) s = dmin(s, low_mach_scale(l, r, o));
  s = dclamp(s, T(1.0e-4), T(1.0));
  const P4T<T> pl = c2p(l, o.g);
  const P4T<T> pr = c2p(r, o.g);
  const T fm = 0.5 * (local_mach(pl, o.g) + local_mach(pr, o.g));
  const T fl = dclamp(o.stab_floor, T(0.0), T(1.0));
  const T me = smooth_max(fm, fl);
  const T st = dclamp(me, T(1.0e-4), T(1.0));
  return dmax(s, st);
}

// All-speed HLLC pressure scale (euler_flux.cpp:81-98, :177-201).
template <bool PLAIN = false, class T>
__device__ __forceinline__ T all_speed_scale(const P4T<T>& pl, const P4T<T>& pr, T nx, T ny, const FaceOptsT<T>& o) {
  if (PLAIN || !o.as_fix) return T(1.0);
  const T unl = pl.u * nx + pl.v * ny;
  const T unr = pr.u * nx + pr.v * ny;
  const T al = dmax(sound(pl, o.g), kFloor);
  const T ar = dmax(sound(pr, o.g), kFloor);
  const T vl = vsqrt(pl.u * pl.u + pl.v * pl.v);
  const T vr = vsqrt(pr.u * pr.u + pr.v * pr.v);
  const T mach = dmax(0.5 * (vl / al + vr / ar), kAsMachFloor);
  T mn = dmax(0.5 * (vabs(unl) / al + vabs(unr) / ar), kAsMachFloor);
  mn = dmin(mn, mach);
  mn = dmax(mn, kAsMachFloor);
  const double c = dmax(o.as_cut, kAsMachFloor);
  const double tf = dclamp(o.as_fmin, kAsMachFloor, 1.0);
  const T tn = dclamp(mn / c, T(tf), T(1.0));
  const T mr = dclamp(mach / c, T(0.0), T(1.0));
  const T mg = mr * mr;
  const T tc = (1.0 - mg) * tn + mg;
  const double w = dclamp(o.as_w, 0.0, 1.0);
  const T te = (1.0 - w) + w * tc;
  return dclamp(te, T(tf), T(1.0));
}

template <class T> struct FluxT { T f0, f1, f2, f3, lam; int fallback; };
using Flux = FluxT<double>;

// Primitive states and sound speeds of a face's two conservative states.  FAST (PLAIN face fluxes,
// T = double, operands bounded by flux_of_states) takes the quotients by each side's density through
// quot and keeps its QuotDen for the HLLC star states.
template <class T> struct SidesT { P4T<T> pl, pr; T al, ar; QuotDen dl, dr; };
template <bool FAST, class T>
__device__ __forceinline__ SidesT<T> face_sides(const C4T<T>& l, const C4T<T>& r, double g) {
  SidesT<T> s;
  if constexpr (FAST) {
    s.pl = c2p_q(l, g, &s.dl);
    s.pr = c2p_q(r, g, &s.dr);
    s.al = sound_q(s.pl, g, s.dl);
    s.ar = sound_q(s.pr, g, s.dr);
  } else {
    s.pl = c2p(l, g);
    s.pr = c2p(r, g);
    s.al = sound(s.pl, g);
    s.ar = sound(s.pr, g);
  }
  return s;
}
template <bool FAST, class T> __device__ __forceinline__ T div_side(T x, T rho, const QuotDen& d) {
  if constexpr (FAST) return quot(x, d);
  return x / rho;
}

template <bool FAST = false, class T>
__device__ __forceinline__ FluxT<T> rusanov(const C4T<T>& l, const C4T<T>& r, T nx, T ny, T s, double g) {  // euler_flux.cpp:203-239
  const SidesT<T> sides = face_sides<FAST>(l, r, g);
  const P4T<T>& pl = sides.pl;
  const P4T<T>& pr = sides.pr;
  const T unl = pl.u * nx + pl.v * ny;
  const T unr = pr.u * nx + pr.v * ny;
  const T al = sides.al;
  const T ar = sides.ar;
  const T sphys = fmx<FAST>(vabs(unl) + al, vabs(unr) + ar);
  const SpeedsT<T> sl = pspeeds<FAST>(unl, al, s);
  const SpeedsT<T> sr = pspeeds<FAST>(unr, ar, s);
  const T sd = fmx<FAST>(sl.sr, sr.sr);
  FluxT<T> F;
  F.lam = fmx<FAST>(sphys, T(kFloor));
  F.fallback = 0;
  // physical fluxes (euler_flux.cpp:28-39)
  const T fl0 = l.r * unl, fr0 = r.r * unr;
  const T fl1 = l.ru * unl + nx * pl.p, fr1 = r.ru * unr + nx * pr.p;
  const T fl2 = l.rv * unl + ny * pl.p, fr2 = r.rv * unr + ny * pr.p;
  const T fl3 = (l.E + pl.p) * unl, fr3 = (r.E + pr.p) * unr;
  F.f0 = 0.5 * (fl0 + fr0) - 0.5 * sd * (r.r - l.r);
  F.f1 = 0.5 * (fl1 + fr1) - 0.5 * sd * (r.ru - l.ru);
  F.f2 = 0.5 * (fl2 + fr2) - 0.5 * sd * (r.rv - l.rv);
  F.f3 = 0.5 * (fl3 + fr3) - 0.5 * sd * (r.E - l.E);
  return F;
}

template <class T> __device__ __forceinline__ T sign_floor(T v) {  // euler_flux.cpp:45-50
  if (vabs(v) >= kFloor) return v;
  return (v >= 0.0) ? T(kFloor) : T(-kFloor);
}

template <bool PLAIN = false, bool FAST = false, class T>
__device__ __forceinline__ FluxT<T> hllc(const C4T<T>& l, const C4T<T>& r, T nx, T ny, T s,
                                         const FaceOptsT<T>& o) {  // euler_flux.cpp:241-364
  const double g = o.g;
  const SidesT<T> sides = face_sides<FAST>(l, r, g);
  const P4T<T>& pl = sides.pl;
  const P4T<T>& pr = sides.pr;
  const T unl = pl.u * nx + pl.v * ny;
  const T unr = pr.u * nx + pr.v * ny;
  const T al = sides.al;
  const T ar = sides.ar;
  const T ps = all_speed_scale<PLAIN>(pl, pr, nx, ny, o);
  const T pavg = 0.5 * (pl.p + pr.p);
  const T plc = pavg + ps * (pl.p - pavg);
  const T prc = pavg + ps * (pr.p - pavg);
  const SpeedsT<T> spl = pspeeds<FAST>(unl, al, s);
  const SpeedsT<T> spr = pspeeds<FAST>(unr, ar, s);
  const T sL = fmn<FAST>(spl.lm, spr.lm);
  const T sR = fmx<FAST>(spl.lp, spr.lp);
  const T sLp = fmn<FAST>(unl - al, unr - ar);
  const T sRp = fmx<FAST>(unl + al, unr + ar);
  FluxT<T> F;
  F.lam = fmx<FAST>(fmx<FAST>(vabs(sLp), vabs(sRp)), T(kFloor));
  F.fallback = 0;
  if (vabs(sR - sL) < kFloor) {
    F = rusanov<FAST>(l, r, nx, ny, s, g);
    F.fallback = 1;
    return F;
  }
  const T fl0 = l.r * unl, fr0 = r.r * unr;
  const T fl1 = l.ru * unl + nx * pl.p, fr1 = r.ru * unr + nx * pr.p;
  const T fl2 = l.rv * unl + ny * pl.p, fr2 = r.rv * unr + ny * pr.p;
  const T fl3 = (l.E + pl.p) * unl, fr3 = (r.E + pr.p) * unr;
  if (sL >= 0.0) {
    F.f0 = fl0; F.f1 = fl1; F.f2 = fl2; F.f3 = fl3;
    return F;
  }
  if (sR <= 0.0) {
    F.f0 = fr0; F.f1 = fr1; F.f2 = fr2; F.f3 = fr3;
    return F;
  }
  const T rl = fmx<FAST>(pl.r, T(kFloor));
  const T rr = fmx<FAST>(pr.r, T(kFloor));
  const T den = rl * (sL - unl) - rr * (sR - unr);
  if (vabs(den) < kFloor) {
    F = rusanov<FAST>(l, r, nx, ny, s, g);
    F.fallback = 1;
    return F;
  }
  const T sM = (prc - plc + rl * unl * (sL - unl) - rr * unr * (sR - unr)) / den;
  // star states (euler_flux.cpp:314-335)
  T ls0 = l.r, ls1 = l.ru, ls2 = l.rv, ls3 = l.E;
  {
    const T skmu = sL - unl, skms = sL - sM;
    if (!(vabs(skmu) < kFloor || vabs(skms) < kFloor)) {
      const T rs = rl * skmu / skms;
      const T sh = sM - unl;
      ls0 = rs;
      ls1 = rs * (pl.u + sh * nx);
      ls2 = rs * (pl.v + sh * ny);
      const T et = div_side<FAST>(l.E, rl, sides.dl);  // rl is pl.r
      const T pt = plc / sign_floor(rl * skmu);
      ls3 = rs * (et + sh * (sM + pt));
    }
  }
  T rs0 = r.r, rs1 = r.ru, rs2 = r.rv, rs3 = r.E;
  {
    const T skmu = sR - unr, skms = sR - sM;
    if (!(vabs(skmu) < kFloor || vabs(skms) < kFloor)) {
      const T rs = rr * skmu / skms;
      const T sh = sM - unr;
      rs0 = rs;
      rs1 = rs * (pr.u + sh * nx);
      rs2 = rs * (pr.v + sh * ny);
      const T et = div_side<FAST>(r.E, rr, sides.dr);  // rr is pr.r
      const T pt = prc / sign_floor(rr * skmu);
      rs3 = rs * (et + sh * (sM + pt));
    }
  }
  if (!(fin(ls0) && fin(ls1) && fin(ls2) && fin(ls3)) || !(fin(rs0) && fin(rs1) && fin(rs2) && fin(rs3)) ||
      ls0 <= 0.0 || rs0 <= 0.0) {
    F = rusanov<FAST>(l, r, nx, ny, s, g);
    F.fallback = 1;
    return F;
  }
  if (sM >= 0.0) {
    F.f0 = fl0 + sL * (ls0 - l.r);
    F.f1 = fl1 + sL * (ls1 - l.ru);
    F.f2 = fl2 + sL * (ls2 - l.rv);
    F.f3 = fl3 + sL * (ls3 - l.E);
  } else {
    F.f0 = fr0 + sR * (rs0 - r.r);
    F.f1 = fr1 + sR * (rs1 - r.ru);
    F.f2 = fr2 + sR * (rs2 - r.rv);
    F.f3 = fr3 + sR * (rs3 - r.E);
  }
  if (!(fin(F.f0) && fin(F.f1) && fin(F.f2) && fin(F.f3))) {
    F = rusanov<FAST>(l, r, nx, ny, s, g);
    F.fallback = 1;
  }
  return F;
}

template <int SCHEME, bool PLAIN = false, bool FAST = false, class T>
__device__ __forceinline__ FluxT<T> face_flux(const C4T<T>& l, const C4T<T>& r, T nx, T ny, T s,
                                              const FaceOptsT<T>& o) {  // euler_flux.cpp:366-390
  if (SCHEME == 1) return hllc<PLAIN, FAST>(l, r, nx, ny, s, o);
  return rusanov<FAST>(l, r, nx, ny, s, o.g);
}

// Characteristic farfield state (boundary_conditions.cpp:27-78).
template <class T>
__device__ __forceinline__ C4T<T> char_farfield_state(const C4T<T>& in, T nx, T ny, T s, const FaceOptsT<T>& o) {
  const double g = o.g;
  const P4T<T> pi = c2p(in, g);
  const P4T<T> pf = c2p(o.ff, g);
  const T uni = pi.u * nx + pi.v * ny;
  const T unf = pf.u * nx + pf.v * ny;
  const T ai = dmax(sound(pi, g), kFloor);
  const T af = dmax(sound(pf, g), kFloor);
  const T rref = dmax(0.5 * (pi.r + pf.r), kFloor);
  const T aref = dmax(0.5 * (ai + af), kFloor);
  const T at = dmax(dclamp(s, T(1.0e-4), T(1.0)) * aref, 1.0e-6);
  if (vabs(uni) >= at) {
    C4T<T> r;  // supersonic normal flow: upwind state, selected per component
    const bool out = uni >= 0.0;
    r.r = out ? in.r : o.ff.r;
    r.ru = out ? in.ru : o.ff.ru;
    r.rv = out ? in.rv : o.ff.rv;
    r.E = out ? in.E : o.ff.E;
    return r;
  }
  const T rp = uni + pi.p / (rref * at);
  const T rm = unf - pf.p / (rref * at);
  const T unb = 0.5 * (rp + rm);
  const T pb = dmax(0.5 * rref * at * (rp - rm), kFloor);
  const bool inflow = unb < 0.0;
  const T tx = inflow ? (pf.u - unf * nx) : (pi.u - uni * nx);
  const T ty = inflow ? (pf.v - unf * ny) : (pi.v - uni * ny);
  const T rb = inflow ? pf.r : pi.r;
  const T pbase = inflow ? pf.p : pi.p;
  const T ab = inflow ? af : ai;
  P4T<T> b;
  b.r = dmax(rb + (pb - pbase) / dmax(ab * ab, kFloor), kFloor);
  b.u = unb * nx + tx;
  b.v = unb * ny + ty;
  b.p = pb;
  return p2c(b, g);
}

// Limiter (reconstruction.cpp:121-146).
//
// FAST (venkat, double): past a * b > 0 every term of den is positive, so den >= e2 >= 1e-12, and
// |num| >= e2 |a|.  With den <= 2^300 and |num| >= 2^-700 (*ok reports both; the caller evaluates a face
// that fails them again without FAST) the quotient's exponents lie less than 768 apart, 1 / den and
// num / den (at least 2^-1000) are normal and |num| is far above 2^-969, so v_div_scale scales nothing
// and quot gives the division's bits; x is then finite and non-zero, the den and finiteness tests cannot
// fire, and with a non-zero the clamp to [min(a, 0), max(a, 0)] is one maxNum and one minNum.
constexpr double kVenkatNumLo = 0x1p-700;
template <int LIMITER, bool FAST = false, class T> __device__ __forceinline__ T limiter(T a, T b, bool* ok = nullptr) {
  if (LIMITER == 1) {
    if (a * b <= 0.0) return T(0.0);
    const T e2 = 1.0e-12 * (1.0 + a * a + b * b);
    const T num = (b * b + 2.0 * a * b + e2) * a;
    const T den = a * a + 2.0 * b * b + a * b + e2;
    if constexpr (FAST) {
      static_assert(std::is_same_v<T, double>, "FAST limiters are in double");
      *ok = *ok & (den <= kQuotHi) & (vabs(num) >= kVenkatNumLo);
      const double x = quot(num, quot_den(den));
      return fmn<true>(fmx<true>(x, fmn<true>(a, 0.0)), fmx<true>(a, 0.0));
    }
    if (vabs(den) <= 1.0e-20) return T(0.0);
    const T x = num / den;
    if (!fin(x)) return T(0.0);
    if (a > 0.0) return dclamp(x, T(0.0), a);
    return dclamp(x, a, T(0.0));
  }
  if (a * b <= 0.0) return T(0.0);
  return (vabs(a) < vabs(b)) ? a : b;
}

template <class T> __device__ __forceinline__ bool valid_prim(const P4T<T>& p) {  // residual_assembly.cpp:53-57
  return fin(p.r) && fin(p.u) && fin(p.v) && fin(p.p) && p.r > kFloor && p.p > kFloor;
}

}  // namespace aero_dev
