// euler_device.hpp -- FP64 pointwise Euler numerics for gfx950 kernels (2D).
//
// Each function follows the reference CPU function it cites, operation by operation, so that with
// -ffp-contract=off a kernel reproduces the FP64 evaluation of the reference algorithm.  The reference
// carries a third velocity component w; in 2D it is exactly zero everywhere (w = rhow/rho = 0,
// n_z = 0, limiter(0, 0) = 0), and every place it enters adds an exact +0 or multiplies by 0, so it is
// dropped here.  States are 4-vectors: primitive (rho, u, v, p), conservative (rho, rhou, rhov, rhoE).
#pragma once

#include <hip/hip_runtime.h>

namespace aero_dev {

struct P4 { double r, u, v, p; };
struct C4 { double r, ru, rv, E; };

constexpr double kFloor = 1.0e-8;      // rho, p and wave-speed floors (euler_flux.cpp:9-11)
constexpr double kMinScale = 1.0e-4;   // acoustic-scale clamp (euler_flux.cpp:12)
constexpr double kAsMachFloor = 1.0e-6;  // all-speed Mach floor (euler_flux.cpp:13)

// std::max / std::min / std::clamp (libstdc++ operand order).
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dclamp(double v, double lo, double hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }
__device__ __forceinline__ double dmax4(double a, double b, double c, double d) {
  double r = a;
  if (r < b) r = b;
  if (r < c) r = c;
  if (r < d) r = d;
  return r;
}
__device__ __forceinline__ bool fin(double x) { return __builtin_isfinite(x); }

// Residual-path options that vary per call (EulerResidualAssemblyConfig, residual_assembly.hpp:12-31).
struct FaceOpts {
  C4 ff;                 // freestream conservative state
  double g;              // gamma
  int low_mach_fix;
  double low_mach_cutoff;
  double stab_floor;
  int char_bc;
  int precond;
  double mach_ref, mach_min, beta_min, beta_max;
  int as_fix;
  double as_cut, as_fmin, as_w;
};

__device__ __forceinline__ C4 p2c(const P4& w, double g) {  // euler_flux.cpp:101-113
  C4 c;
  c.r = dmax(w.r, kFloor);
  c.ru = c.r * w.u;
  c.rv = c.r * w.v;
  const double ke = 0.5 * c.r * (w.u * w.u + w.v * w.v);
  const double p = dmax(w.p, kFloor);
  c.E = p / (g - 1.0) + ke;
  return c;
}

__device__ __forceinline__ P4 c2p(const C4& c, double g) {  // euler_flux.cpp:115-127
  P4 w;
  w.r = dmax(c.r, kFloor);
  const double inv = 1.0 / w.r;
  w.u = c.ru * inv;
  w.v = c.rv * inv;
  const double ke = 0.5 * w.r * (w.u * w.u + w.v * w.v);
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

__device__ __forceinline__ double sound(const P4& w, double g) {  // euler_flux.cpp:150-152
  return sqrt(dmax(g * w.p / dmax(w.r, kFloor), 0.0));
}

struct Speeds { double lm, lp, sr; };

__device__ __forceinline__ Speeds pspeeds(double un, double a, double s) {  // euler_flux.cpp:58-74
  const double b = dclamp(s, kMinScale, 1.0);
  const double b2 = b * b;
  const double omb2 = 1.0 - b2;
  const double ut = 0.5 * (1.0 + b2) * un;
  const double disc = omb2 * omb2 * un * un + 4.0 * b2 * a * a;
  const double ct = 0.5 * sqrt(dmax(disc, 0.0));
  Speeds sp;
  sp.lm = ut - ct;
  sp.lp = ut + ct;
  sp.sr = dmax4(fabs(un), fabs(sp.lm), fabs(sp.lp), kFloor);
  return sp;
}

__device__ __forceinline__ double local_mach(const P4& p, double g) {  // preconditioning.cpp:141-146
  const double a = dmax(sound(p, g), kFloor);
  const double vm = sqrt(p.u * p.u + p.v * p.v);
  return vm / a;
}

__device__ __forceinline__ double cell_beta(const P4& p, const FaceOpts& o) {  // preconditioning.cpp:134-156
  const double lm = local_mach(p, o.g);
  const double t = dmax4(lm, o.mach_ref, o.mach_min, kFloor);
  const double bmin = dmax(o.beta_min, 1.0e-4);
  const double bmax = dmax(bmin, o.beta_max);
  return dclamp(t, bmin, bmax);
}

__device__ __forceinline__ double face_beta(const C4& l, const C4& r, const FaceOpts& o) {  // :158-173
  if (!o.precond) return 1.0;
  const P4 pl = c2p(l, o.g);
  const P4 pr = c2p(r, o.g);
  P4 a;
  a.r = 0.5 * (pl.r + pr.r);
  a.u = 0.5 * (pl.u + pr.u);
  a.v = 0.5 * (pl.v + pr.v);
  a.p = 0.5 * (pl.p + pr.p);
  return cell_beta(a, o);
}

__device__ __forceinline__ double low_mach_scale(const C4& l, const C4& r, const FaceOpts& o) {  // euler_flux.cpp:154-175
  if (o.low_mach_cutoff <= 0.0) return 1.0;
  const P4 pl = c2p(l, o.g);
  const P4 pr = c2p(r, o.g);
  const double al = dmax(sound(pl, o.g), kFloor);
  const double ar = dmax(sound(pr, o.g), kFloor);
  const double vl = sqrt(pl.u * pl.u + pl.v * pl.v);
  const double vr = sqrt(pr.u * pr.u + pr.v * pr.v);
  const double ml = vl / al;
  const double mr = vr / ar;
  const double fm = 0.5 * (ml + mr);
  const double ms = dmax(fm, dmax(o.mach_ref, kFloor));
  const double ratio = ms / dmax(o.low_mach_cutoff, kFloor);
  return dclamp(ratio, kMinScale, 1.0);
}

__device__ __forceinline__ double smooth_max(double a, double b) {  // residual_assembly.cpp:12-20
  const double d = a - b;
  return 0.5 * (a + b + sqrt(d * d + 1.0e-6 * 1.0e-6));
}

// Face acoustic scale (residual_assembly.cpp:31-51).
__device__ __forceinline__ double acoustic_scale(const C4& l, const C4& r, double fb, const FaceOpts& o) {
  double s = 1.0;
  if (o.precond) s = dclamp(fb, 1.0e-4, 1.0);
  if (o.low_mach_fix) s = dmin(s, low_mach_scale(l, r, o));
  s = dclamp(s, 1.0e-4, 1.0);
  const P4 pl = c2p(l, o.g);
  const P4 pr = c2p(r, o.g);
  const double fm = 0.5 * (local_mach(pl, o.g) + local_mach(pr, o.g));
  const double fl = dclamp(o.stab_floor, 0.0, 1.0);
  const double me = smooth_max(fm, fl);
  const double st = dclamp(me, 1.0e-4, 1.0);
  return dmax(s, st);
}

// All-speed HLLC pressure scale (euler_flux.cpp:81-98, :177-201).
__device__ __forceinline__ double all_speed_scale(const P4& pl, const P4& pr, double nx, double ny,
                                                  const FaceOpts& o) {
  if (!o.as_fix) return 1.0;
  const double unl = pl.u * nx + pl.v * ny;
  const double unr = pr.u * nx + pr.v * ny;
  const double al = dmax(sound(pl, o.g), kFloor);
  const double ar = dmax(sound(pr, o.g), kFloor);
  const double vl = sqrt(pl.u * pl.u + pl.v * pl.v);
  const double vr = sqrt(pr.u * pr.u + pr.v * pr.v);
  const double mach = dmax(0.5 * (vl / al + vr / ar), kAsMachFloor);
  double mn = dmax(0.5 * (fabs(unl) / al + fabs(unr) / ar), kAsMachFloor);
  mn = dmin(mn, mach);
  mn = dmax(mn, kAsMachFloor);
  const double c = dmax(o.as_cut, kAsMachFloor);
  const double tf = dclamp(o.as_fmin, kAsMachFloor, 1.0);
  const double tn = dclamp(mn / c, tf, 1.0);
  const double mr = dclamp(mach / c, 0.0, 1.0);
  const double mg = mr * mr;
  const double tc = (1.0 - mg) * tn + mg;
  const double w = dclamp(o.as_w, 0.0, 1.0);
  const double te = (1.0 - w) + w * tc;
  return dclamp(te, tf, 1.0);
}

struct Flux { double f0, f1, f2, f3, lam; int fallback; };

__device__ __forceinline__ Flux rusanov(const C4& l, const C4& r, double nx, double ny, double s, double g) {  // euler_flux.cpp:203-239
  const P4 pl = c2p(l, g);
  const P4 pr = c2p(r, g);
  const double unl = pl.u * nx + pl.v * ny;
  const double unr = pr.u * nx + pr.v * ny;
  const double al = sound(pl, g);
  const double ar = sound(pr, g);
  const double sphys = dmax(fabs(unl) + al, fabs(unr) + ar);
  const Speeds sl = pspeeds(unl, al, s);
  const Speeds sr = pspeeds(unr, ar, s);
  const double sd = dmax(sl.sr, sr.sr);
  Flux F;
  F.lam = dmax(sphys, kFloor);
  F.fallback = 0;
  // physical fluxes (euler_flux.cpp:28-39)
  const double fl0 = l.r * unl, fr0 = r.r * unr;
  const double fl1 = l.ru * unl + nx * pl.p, fr1 = r.ru * unr + nx * pr.p;
  const double fl2 = l.rv * unl + ny * pl.p, fr2 = r.rv * unr + ny * pr.p;
  const double fl3 = (l.E + pl.p) * unl, fr3 = (r.E + pr.p) * unr;
  F.f0 = 0.5 * (fl0 + fr0) - 0.5 * sd * (r.r - l.r);
  F.f1 = 0.5 * (fl1 + fr1) - 0.5 * sd * (r.ru - l.ru);
  F.f2 = 0.5 * (fl2 + fr2) - 0.5 * sd * (r.rv - l.rv);
  F.f3 = 0.5 * (fl3 + fr3) - 0.5 * sd * (r.E - l.E);
  return F;
}

__device__ __forceinline__ double sign_floor(double v) {  // euler_flux.cpp:45-50
  if (fabs(v) >= kFloor) return v;
  return (v >= 0.0) ? kFloor : -kFloor;
}

__device__ __forceinline__ Flux hllc(const C4& l, const C4& r, double nx, double ny, double s,
                                     const FaceOpts& o) {  // euler_flux.cpp:241-364
  const double g = o.g;
  const P4 pl = c2p(l, g);
  const P4 pr = c2p(r, g);
  const double unl = pl.u * nx + pl.v * ny;
  const double unr = pr.u * nx + pr.v * ny;
  const double al = sound(pl, g);
  const double ar = sound(pr, g);
  const double ps = all_speed_scale(pl, pr, nx, ny, o);
  const double pavg = 0.5 * (pl.p + pr.p);
  const double plc = pavg + ps * (pl.p - pavg);
  const double prc = pavg + ps * (pr.p - pavg);
  const Speeds spl = pspeeds(unl, al, s);
  const Speeds spr = pspeeds(unr, ar, s);
  const double sL = dmin(spl.lm, spr.lm);
  const double sR = dmax(spl.lp, spr.lp);
  const double sLp = dmin(unl - al, unr - ar);
  const double sRp = dmax(unl + al, unr + ar);
  Flux F;
  F.lam = dmax(dmax(fabs(sLp), fabs(sRp)), kFloor);
  F.fallback = 0;
  if (fabs(sR - sL) < kFloor) {
    F = rusanov(l, r, nx, ny, s, g);
    F.fallback = 1;
    return F;
  }
  const double fl0 = l.r * unl, fr0 = r.r * unr;
  const double fl1 = l.ru * unl + nx * pl.p, fr1 = r.ru * unr + nx * pr.p;
  const double fl2 = l.rv * unl + ny * pl.p, fr2 = r.rv * unr + ny * pr.p;
  const double fl3 = (l.E + pl.p) * unl, fr3 = (r.E + pr.p) * unr;
  if (sL >= 0.0) {
    F.f0 = fl0; F.f1 = fl1; F.f2 = fl2; F.f3 = fl3;
    return F;
  }
  if (sR <= 0.0) {
    F.f0 = fr0; F.f1 = fr1; F.f2 = fr2; F.f3 = fr3;
    return F;
  }
  const double rl = dmax(pl.r, kFloor);
  const double rr = dmax(pr.r, kFloor);
  const double den = rl * (sL - unl) - rr * (sR - unr);
  if (fabs(den) < kFloor) {
    F = rusanov(l, r, nx, ny, s, g);
    F.fallback = 1;
    return F;
  }
  const double sM = (prc - plc + rl * unl * (sL - unl) - rr * unr * (sR - unr)) / den;
  // star states (euler_flux.cpp:314-335)
  double ls0 = l.r, ls1 = l.ru, ls2 = l.rv, ls3 = l.E;
  {
    const double skmu = sL - unl, skms = sL - sM;
    if (!(fabs(skmu) < kFloor || fabs(skms) < kFloor)) {
      const double rs = rl * skmu / skms;
      const double sh = sM - unl;
      ls0 = rs;
      ls1 = rs * (pl.u + sh * nx);
      ls2 = rs * (pl.v + sh * ny);
      const double et = l.E / rl;
      const double pt = plc / sign_floor(rl * skmu);
      ls3 = rs * (et + sh * (sM + pt));
    }
  }
  double rs0 = r.r, rs1 = r.ru, rs2 = r.rv, rs3 = r.E;
  {
    const double skmu = sR - unr, skms = sR - sM;
    if (!(fabs(skmu) < kFloor || fabs(skms) < kFloor)) {
      const double rs = rr * skmu / skms;
      const double sh = sM - unr;
      rs0 = rs;
      rs1 = rs * (pr.u + sh * nx);
      rs2 = rs * (pr.v + sh * ny);
      const double et = r.E / rr;
      const double pt = prc / sign_floor(rr * skmu);
      rs3 = rs * (et + sh * (sM + pt));
    }
  }
  if (!(fin(ls0) && fin(ls1) && fin(ls2) && fin(ls3)) || !(fin(rs0) && fin(rs1) && fin(rs2) && fin(rs3)) ||
      ls0 <= 0.0 || rs0 <= 0.0) {
    F = rusanov(l, r, nx, ny, s, g);
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
    F = rusanov(l, r, nx, ny, s, g);
    F.fallback = 1;
  }
  return F;
}

template <int SCHEME>
__device__ __forceinline__ Flux face_flux(const C4& l, const C4& r, double nx, double ny, double s,
                                          const FaceOpts& o) {  // euler_flux.cpp:366-390
  if (SCHEME == 1) return hllc(l, r, nx, ny, s, o);
  return rusanov(l, r, nx, ny, s, o.g);
}

// Characteristic farfield state (boundary_conditions.cpp:27-78).
__device__ __forceinline__ C4 char_farfield_state(const C4& in, double nx, double ny, double s, const FaceOpts& o) {
  const double g = o.g;
  const P4 pi = c2p(in, g);
  const P4 pf = c2p(o.ff, g);
  const double uni = pi.u * nx + pi.v * ny;
  const double unf = pf.u * nx + pf.v * ny;
  const double ai = dmax(sound(pi, g), kFloor);
  const double af = dmax(sound(pf, g), kFloor);
  const double rref = dmax(0.5 * (pi.r + pf.r), kFloor);
  const double aref = dmax(0.5 * (ai + af), kFloor);
  const double at = dmax(dclamp(s, 1.0e-4, 1.0) * aref, 1.0e-6);
  if (fabs(uni) >= at) return (uni >= 0.0) ? in : o.ff;
  const double rp = uni + pi.p / (rref * at);
  const double rm = unf - pf.p / (rref * at);
  const double unb = 0.5 * (rp + rm);
  const double pb = dmax(0.5 * rref * at * (rp - rm), kFloor);
  const bool inflow = unb < 0.0;
  const double tx = inflow ? (pf.u - unf * nx) : (pi.u - uni * nx);
  const double ty = inflow ? (pf.v - unf * ny) : (pi.v - uni * ny);
  const double rb = inflow ? pf.r : pi.r;
  const double pbase = inflow ? pf.p : pi.p;
  const double ab = inflow ? af : ai;
  P4 b;
  b.r = dmax(rb + (pb - pbase) / dmax(ab * ab, kFloor), kFloor);
  b.u = unb * nx + tx;
  b.v = unb * ny + ty;
  b.p = pb;
  return p2c(b, g);
}

// Limiter (reconstruction.cpp:121-146).
template <int LIMITER> __device__ __forceinline__ double limiter(double a, double b) {
  if (LIMITER == 1) {
    if (a * b <= 0.0) return 0.0;
    const double e2 = 1.0e-12 * (1.0 + a * a + b * b);
    const double num = (b * b + 2.0 * a * b + e2) * a;
    const double den = a * a + 2.0 * b * b + a * b + e2;
    if (fabs(den) <= 1.0e-20) return 0.0;
    const double x = num / den;
    if (!fin(x)) return 0.0;
    if (a > 0.0) return dclamp(x, 0.0, a);
    return dclamp(x, a, 0.0);
  }
  if (a * b <= 0.0) return 0.0;
  return (fabs(a) < fabs(b)) ? a : b;
}

__device__ __forceinline__ bool valid_prim(const P4& p) {  // residual_assembly.cpp:53-57
  return fin(p.r) && fin(p.u) && fin(p.v) && fin(p.p) && p.r > kFloor && p.p > kFloor;
}

}  // namespace aero_dev
