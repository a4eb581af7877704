// aero_oracle.hpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// CPU restatement of the AeroCFD reference (AstaWasTaken/Aero-CFD-Solver v0.1.1) 2D finite-volume
// Euler residual path, templated on the scalar type:
//   * Real = float  reproduces the reference's own FP32 arithmetic operation by operation.  It is
//     pinned bit-for-bit against golden vectors written by the compiled reference (tests/golden/).
//   * Real = double is the FP64 parity oracle the HIP backend is checked against.
// Constants are the reference's FP32 literals; the double instantiation uses the same decimal values
// (AO_K below picks `x##f` for float and `x` for double, so neither instantiation double-rounds).
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this code, and only
// as the checker / CPU baseline.  The product path (aero-cfd-solver_amd/) never includes it.
//
// Every function cites the reference file:line it follows (paths relative to the reference root).
// All functions here are pinned: see tests/test_oracle_golden.py (golden vectors from the compiled
// reference) and tests/test_oracle_invariants.py (closed forms / invariants for what the reference
// has no function for, e.g. the density-only norm).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <deque>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace aero_oracle {

template <class R> constexpr R kc(float f, double d);
template <> constexpr float kc<float>(float f, double) { return f; }
template <> constexpr double kc<double>(float, double d) { return d; }
#define AO_K(x) ::aero_oracle::kc<R>(x##f, x)

// sqrt / abs / isfinite, overloadable for other scalar types (oracle_dual.hpp).
template <class T> inline T ao_sqrt(T x) { return std::sqrt(x); }
template <class T> inline T ao_abs(T x) { return std::abs(x); }
template <class T> inline bool ao_isfinite(T x) { return std::isfinite(x); }
template <class T> inline double ao_value(T x) { return static_cast<double>(x); }

// std::max / std::min / std::clamp semantics (libstdc++), spelled out so the operand order is explicit.
template <class T> inline T smax(T a, T b) { return (a < b) ? b : a; }
template <class T> inline T smin(T a, T b) { return (b < a) ? b : a; }
template <class T> inline T sclamp(T v, T lo, T hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }
// std::max({a, b, c, d}): first largest element.
template <class T> inline T smax4(T a, T b, T c, T d) {
  T r = a;
  if (r < b) r = b;
  if (r < c) r = c;
  if (r < d) r = d;
  return r;
}

template <class R> struct Prim { R rho = R(1), u = R(0), v = R(0), w = R(0), p = R(1); };
template <class R> struct Cons { R rho = R(1), rhou = R(0), rhov = R(0), rhow = R(0), rhoE = R(1); };
template <class R> using A3 = std::array<R, 3>;
template <class R> using A5 = std::array<R, 5>;

enum Scheme : int { kRusanov = 0, kHllc = 1 };
enum Limiter : int { kMinmod = 0, kVenkat = 1 };
enum Bc : int { kInterior = 0, kFarfield = 1, kSlipWall = 2 };

// ---------------------------------------------------------------------------------------------
// State maps (src/cfd_core/numerics/euler_flux.cpp:101-152)
// ---------------------------------------------------------------------------------------------
template <class R> Cons<R> prim_to_cons(const Prim<R>& w, R g) {  // euler_flux.cpp:101-113
  Cons<R> c;
  c.rho = smax(w.rho, AO_K(1.0e-8));
  c.rhou = c.rho * w.u;
  c.rhov = c.rho * w.v;
  c.rhow = c.rho * w.w;
  const R ke = AO_K(0.5) * c.rho * (w.u * w.u + w.v * w.v + w.w * w.w);
  const R p = smax(w.p, AO_K(1.0e-8));
  c.rhoE = p / (g - AO_K(1.0)) + ke;
  return c;
}

template <class R> Prim<R> cons_to_prim(const Cons<R>& c, R g) {  // euler_flux.cpp:115-127
  Prim<R> w;
  w.rho = smax(c.rho, AO_K(1.0e-8));
  const R inv = AO_K(1.0) / w.rho;
  w.u = c.rhou * inv;
  w.v = c.rhov * inv;
  w.w = c.rhow * inv;
  const R ke = AO_K(0.5) * w.rho * (w.u * w.u + w.v * w.v + w.w * w.w);
  w.p = smax((g - AO_K(1.0)) * (c.rhoE - ke), AO_K(1.0e-8));
  return w;
}

template <class R> void enforce_physical_state(Cons<R>* s, R g) {  // euler_flux.cpp:129-144
  s->rho = smax(s->rho, AO_K(1.0e-8));
  const R inv = AO_K(1.0) / s->rho;
  const R u = s->rhou * inv;
  const R v = s->rhov * inv;
  const R w = s->rhow * inv;
  const R ke = AO_K(0.5) * s->rho * (u * u + v * v + w * w);
  s->rhoE = smax(s->rhoE, ke + AO_K(1.0e-8) / (g - AO_K(1.0)));
  const R p = (g - AO_K(1.0)) * (s->rhoE - ke);
  if (p < AO_K(1.0e-8)) s->rhoE = ke + AO_K(1.0e-8) / (g - AO_K(1.0));
}

template <class R> R speed_of_sound(const Prim<R>& w, R g) {  // euler_flux.cpp:150-152
  return ao_sqrt(smax(g * w.p / smax(w.rho, AO_K(1.0e-8)), R(0)));
}

template <class R> A5<R> cons_array(const Cons<R>& c) { return {c.rho, c.rhou, c.rhov, c.rhow, c.rhoE}; }

template <class R> bool finite5(const A5<R>& a) {  // euler_flux.cpp:19-26
  for (R x : a)
    if (!ao_isfinite(x)) return false;
  return true;
}

template <class R> A5<R> physical_flux(const Cons<R>& c, const Prim<R>& w, const A3<R>& n) {  // euler_flux.cpp:28-39
  const R un = w.u * n[0] + w.v * n[1] + w.w * n[2];
  return {c.rho * un, c.rhou * un + n[0] * w.p, c.rhov * un + n[1] * w.p, c.rhow * un + n[2] * w.p,
          (c.rhoE + w.p) * un};
}

template <class R> R with_sign_floor(R v) {  // euler_flux.cpp:45-50
  if (ao_abs(v) >= AO_K(1.0e-8)) return v;
  return (v >= R(0)) ? AO_K(1.0e-8) : -AO_K(1.0e-8);
}

template <class R> struct Speeds { R lm = 0, lp = 0, sr = 0; };

template <class R> Speeds<R> precond_speeds(R un, R a, R s) {  // euler_flux.cpp:58-74
  const R beta = sclamp(s, AO_K(1.0e-4), AO_K(1.0));
  const R b2 = beta * beta;
  const R omb2 = AO_K(1.0) - b2;
  const R ut = AO_K(0.5) * (AO_K(1.0) + b2) * un;
  const R disc = omb2 * omb2 * un * un + AO_K(4.0) * b2 * a * a;
  const R ct = AO_K(0.5) * ao_sqrt(smax(disc, R(0)));
  Speeds<R> sp;
  sp.lm = ut - ct;
  sp.lp = ut + ct;
  sp.sr = smax4(ao_abs(un), ao_abs(sp.lm), ao_abs(sp.lp), AO_K(1.0e-8));
  return sp;
}

template <class R> struct MachSensors { R mach = 1, mn = 1; };

template <class R>
MachSensors<R> face_mach_sensors(const Prim<R>& pl, const Prim<R>& pr, const A3<R>& n, R g) {  // euler_flux.cpp:81-98
  const R unl = pl.u * n[0] + pl.v * n[1] + pl.w * n[2];
  const R unr = pr.u * n[0] + pr.v * n[1] + pr.w * n[2];
  const R al = smax(speed_of_sound(pl, g), AO_K(1.0e-8));
  const R ar = smax(speed_of_sound(pr, g), AO_K(1.0e-8));
  const R vl = ao_sqrt(pl.u * pl.u + pl.v * pl.v + pl.w * pl.w);
  const R vr = ao_sqrt(pr.u * pr.u + pr.v * pr.v + pr.w * pr.w);
  MachSensors<R> s;
  s.mach = smax(AO_K(0.5) * (vl / al + vr / ar), AO_K(1.0e-6));
  s.mn = smax(AO_K(0.5) * (ao_abs(unl) / al + ao_abs(unr) / ar), AO_K(1.0e-6));
  s.mn = smin(s.mn, s.mach);
  s.mn = smax(s.mn, AO_K(1.0e-6));
  return s;
}

template <class R>
R low_mach_dissipation_scale(const Cons<R>& l, const Cons<R>& r, R g, bool enable, R cutoff,
                             R mref) {  // euler_flux.cpp:154-175
  if (!enable || cutoff <= R(0)) return AO_K(1.0);
  const Prim<R> pl = cons_to_prim(l, g);
  const Prim<R> pr = cons_to_prim(r, g);
  const R al = smax(speed_of_sound(pl, g), AO_K(1.0e-8));
  const R ar = smax(speed_of_sound(pr, g), AO_K(1.0e-8));
  const R vl = ao_sqrt(pl.u * pl.u + pl.v * pl.v + pl.w * pl.w);
  const R vr = ao_sqrt(pr.u * pr.u + pr.v * pr.v + pr.w * pr.w);
  const R ml = vl / al;
  const R mr = vr / ar;
  const R fm = AO_K(0.5) * (ml + mr);
  const R mstar = smax(fm, smax(mref, AO_K(1.0e-8)));
  const R ratio = mstar / smax(cutoff, AO_K(1.0e-8));
  return sclamp(ratio, AO_K(1.0e-4), AO_K(1.0));
}

template <class R>
R all_speed_pressure_scale(const Cons<R>& l, const Cons<R>& r, const A3<R>& n, R g, bool enable,
                           R cutoff, R fmin, R ramp) {  // euler_flux.cpp:177-201
  if (!enable) return AO_K(1.0);
  const Prim<R> pl = cons_to_prim(l, g);
  const Prim<R> pr = cons_to_prim(r, g);
  const MachSensors<R> s = face_mach_sensors(pl, pr, n, g);
  const R c = smax(cutoff, AO_K(1.0e-6));
  const R tf = sclamp(fmin, AO_K(1.0e-6), AO_K(1.0));
  const R tn = sclamp(s.mn / c, tf, AO_K(1.0));
  const R mr = sclamp(s.mach / c, R(0), AO_K(1.0));
  const R mg = mr * mr;
  const R tc = (AO_K(1.0) - mg) * tn + mg;
  const R w = sclamp(ramp, R(0), AO_K(1.0));
  const R te = (AO_K(1.0) - w) + w * tc;
  return sclamp(te, tf, AO_K(1.0));
}

template <class R>
A5<R> rusanov_flux(const Cons<R>& l, const Cons<R>& r, const A3<R>& n, R g, R* lam,
                   R s) {  // euler_flux.cpp:203-239
  const Prim<R> pl = cons_to_prim(l, g);
  const Prim<R> pr = cons_to_prim(r, g);
  const R unl = pl.u * n[0] + pl.v * n[1] + pl.w * n[2];
  const R unr = pr.u * n[0] + pr.v * n[1] + pr.w * n[2];
  const R al = speed_of_sound(pl, g);
  const R ar = speed_of_sound(pr, g);
  const R smax_phys = smax(ao_abs(unl) + al, ao_abs(unr) + ar);
  const Speeds<R> sl = precond_speeds(unl, al, s);
  const Speeds<R> sr = precond_speeds(unr, ar, s);
  const R smax_d = smax(sl.sr, sr.sr);
  if (lam) *lam = smax(smax_phys, AO_K(1.0e-8));
  const A5<R> fl = physical_flux(l, pl, n);
  const A5<R> fr = physical_flux(r, pr, n);
  const A5<R> du = {r.rho - l.rho, r.rhou - l.rhou, r.rhov - l.rhov, r.rhow - l.rhow, r.rhoE - l.rhoE};
  A5<R> f{};
  for (int k = 0; k < 5; ++k) f[k] = AO_K(0.5) * (fl[k] + fr[k]) - AO_K(0.5) * smax_d * du[k];
  return f;
}

template <class R>
A5<R> hllc_flux(const Cons<R>& l, const Cons<R>& r, const A3<R>& n, R g, R* lam, R s, bool as_fix,
                R as_cut, R as_fmin, R as_w, bool* fallback) {  // euler_flux.cpp:241-364
  if (fallback) *fallback = false;
  const Prim<R> pl = cons_to_prim(l, g);
  const Prim<R> pr = cons_to_prim(r, g);
  const R unl = pl.u * n[0] + pl.v * n[1] + pl.w * n[2];
  const R unr = pr.u * n[0] + pr.v * n[1] + pr.w * n[2];
  const R al = speed_of_sound(pl, g);
  const R ar = speed_of_sound(pr, g);
  const R ps = all_speed_pressure_scale(l, r, n, g, as_fix, as_cut, as_fmin, as_w);
  const R pavg = AO_K(0.5) * (pl.p + pr.p);
  const R plc = pavg + ps * (pl.p - pavg);
  const R prc = pavg + ps * (pr.p - pavg);
  const Speeds<R> spl = precond_speeds(unl, al, s);
  const Speeds<R> spr = precond_speeds(unr, ar, s);
  const R sL = smin(spl.lm, spr.lm);
  const R sR = smax(spl.lp, spr.lp);
  const R sLp = smin(unl - al, unr - ar);
  const R sRp = smax(unl + al, unr + ar);
  if (lam) *lam = smax(smax(ao_abs(sLp), ao_abs(sRp)), AO_K(1.0e-8));

  if (ao_abs(sR - sL) < AO_K(1.0e-8)) {
    if (fallback) *fallback = true;
    return rusanov_flux(l, r, n, g, lam, s);
  }
  const A5<R> ul = cons_array(l);
  const A5<R> ur = cons_array(r);
  const A5<R> fl = physical_flux(l, pl, n);
  const A5<R> fr = physical_flux(r, pr, n);
  if (sL >= R(0)) return fl;
  if (sR <= R(0)) return fr;

  const R rl = smax(pl.rho, AO_K(1.0e-8));
  const R rr = smax(pr.rho, AO_K(1.0e-8));
  const R den = rl * (sL - unl) - rr * (sR - unr);
  if (ao_abs(den) < AO_K(1.0e-8)) {
    if (fallback) *fallback = true;
    return rusanov_flux(l, r, n, g, lam, s);
  }
  const R sM = (prc - plc + rl * unl * (sL - unl) - rr * unr * (sR - unr)) / den;

  auto star = [&](const Cons<R>& st, const Prim<R>& pm, R pc, R sk, R unk) {  // euler_flux.cpp:314-335
    A5<R> x = cons_array(st);
    const R rk = smax(pm.rho, AO_K(1.0e-8));
    const R skmu = sk - unk;
    const R skms = sk - sM;
    if (ao_abs(skmu) < AO_K(1.0e-8) || ao_abs(skms) < AO_K(1.0e-8)) return x;
    const R rs = rk * skmu / skms;
    const R sh = sM - unk;
    x[0] = rs;
    x[1] = rs * (pm.u + sh * n[0]);
    x[2] = rs * (pm.v + sh * n[1]);
    x[3] = rs * (pm.w + sh * n[2]);
    const R et = st.rhoE / rk;
    const R pt = pc / with_sign_floor(rk * skmu);
    x[4] = rs * (et + sh * (sM + pt));
    return x;
  };
  const A5<R> usl = star(l, pl, plc, sL, unl);
  const A5<R> usr = star(r, pr, prc, sR, unr);
  if (!finite5(usl) || !finite5(usr) || usl[0] <= R(0) || usr[0] <= R(0)) {
    if (fallback) *fallback = true;
    return rusanov_flux(l, r, n, g, lam, s);
  }
  A5<R> f{};
  if (sM >= R(0)) {
    for (int k = 0; k < 5; ++k) f[k] = fl[k] + sL * (usl[k] - ul[k]);
  } else {
    for (int k = 0; k < 5; ++k) f[k] = fr[k] + sR * (usr[k] - ur[k]);
  }
  if (!finite5(f)) {
    if (fallback) *fallback = true;
    return rusanov_flux(l, r, n, g, lam, s);
  }
  return f;
}

template <class R>
A5<R> face_flux(const Cons<R>& l, const Cons<R>& r, const A3<R>& n, R g, int scheme, R* lam, R s,
                bool as_fix, R as_cut, R as_fmin, R as_w, bool* fallback) {  // euler_flux.cpp:366-390
  if (scheme == kHllc) return hllc_flux(l, r, n, g, lam, s, as_fix, as_cut, as_fmin, as_w, fallback);
  if (fallback) *fallback = false;
  return rusanov_flux(l, r, n, g, lam, s);
}

// ---------------------------------------------------------------------------------------------
// Limiter (src/cfd_core/numerics/reconstruction.cpp:121-146)
// ---------------------------------------------------------------------------------------------
template <class R> R apply_limiter(int lim, R a, R b) {
  if (lim == kVenkat) {
    if (a * b <= R(0)) return R(0);
    const R e2 = AO_K(1.0e-12) * (AO_K(1.0) + a * a + b * b);
    const R num = (b * b + AO_K(2.0) * a * b + e2) * a;
    const R den = a * a + AO_K(2.0) * b * b + a * b + e2;
    if (ao_abs(den) <= AO_K(1.0e-20)) return R(0);
    const R x = num / den;
    if (!ao_isfinite(x)) return R(0);
    if (a > R(0)) return sclamp(x, R(0), a);
    return sclamp(x, a, R(0));
  }
  if (a * b <= R(0)) return R(0);
  return (ao_abs(a) < ao_abs(b)) ? a : b;
}

// ---------------------------------------------------------------------------------------------
// Mesh view: the reference UnstructuredMesh arrays (src/cfd_core/include/cfd_core/mesh.hpp:14-36),
// xyz-interleaved like the reference, non-owning.
// ---------------------------------------------------------------------------------------------
template <class R> struct MeshView {
  int num_cells = 0;
  int num_faces = 0;
  const int* owner = nullptr;
  const int* neighbor = nullptr;
  const int* bc = nullptr;             // per face: kInterior / kFarfield / kSlipWall
  const R* normal = nullptr;           // 3F, unit, out of owner
  const R* face_center = nullptr;      // 3F
  const R* area = nullptr;             // F
  const R* volume = nullptr;           // N
  const R* cell_center = nullptr;      // 3N
  int wall_start = 0, wall_count = 0;  // the "wall" patch range
  int gradient = 0;                    // 0: Green-Gauss, 1: weighted least squares
};

template <class R> A3<R> normal_of(const MeshView<R>& m, int f) {  // mesh_geometry.cpp:33-39
  return {m.normal[3 * f + 0], m.normal[3 * f + 1], m.normal[3 * f + 2]};
}

template <class R> R prim_var(const Prim<R>& p, int k) {  // reconstruction.cpp:16-31
  switch (k) {
    case 0: return p.rho;
    case 1: return p.u;
    case 2: return p.v;
    case 3: return p.w;
    case 4: return p.p;
    default: return R(0);
  }
}
template <class R> void set_prim_var(Prim<R>* p, int k, R x) {  // reconstruction.cpp:33-56
  switch (k) {
    case 0: p->rho = x; break;
    case 1: p->u = x; break;
    case 2: p->v = x; break;
    case 3: p->w = x; break;
    case 4: p->p = x; break;
    default: break;
  }
}

// Green-Gauss primitive gradients, layout [cell][var 5][comp 2] (reconstruction.cpp:73-119).
template <class R>
void green_gauss(const MeshView<R>& m, const std::vector<Prim<R>>& prim, std::vector<R>* grad) {
  const int nc = 2;
  grad->assign(static_cast<std::size_t>(m.num_cells) * 5 * nc, R(0));
  auto G = [&](int c, int k, int d) -> R& { return (*grad)[static_cast<std::size_t>((c * 5 + k) * nc + d)]; };
  for (int f = 0; f < m.num_faces; ++f) {
    const int o = m.owner[f];
    const int nb = m.neighbor[f];
    for (int k = 0; k < 5; ++k) {
      const R qo = prim_var(prim[o], k);
      R qf = qo;
      if (nb >= 0) {
        const R qn = prim_var(prim[nb], k);
        qf = AO_K(0.5) * (qo + qn);
      }
      for (int d = 0; d < nc; ++d) {
        const R nA = m.normal[3 * f + d] * m.area[f];
        G(o, k, d) += qf * nA;
        if (nb >= 0) G(nb, k, d) -= qf * nA;
      }
    }
  }
  for (int c = 0; c < m.num_cells; ++c) {
    const R iv = AO_K(1.0) / smax(m.volume[c], AO_K(1.0e-12));
    for (int k = 0; k < 5; ++k)
      for (int d = 0; d < nc; ++d) G(c, k, d) *= iv;
  }
}

// Weighted least-squares primitive gradients, layout [cell][var 5][comp 2].  The reference has no such
// function (its roadmap lists one, docs/project_outline.md:49); this is the definition both backends
// follow.  For every interior face in increasing face index, with d = x_neighbour - x_owner,
// w = 1 / (d.d) and dq_k = q_k(neighbour) - q_k(owner), both adjacent cells add
//   a += (w dx) dx,  b += (w dx) dy,  c += (w dy) dy,  r_kx += (w dx) dq_k,  r_ky += (w dy) dq_k
// (the terms are the same seen from either side: d and dq both change sign).  Boundary faces add
// nothing.  The 2x2 normal equations are solved by Cramer's rule; when det = a c - b b is not above
// 1e-12 (a + c)^2 (collinear neighbours) the cell's gradient is zero.  Exact for linear fields.
template <class R>
void least_squares(const MeshView<R>& m, const std::vector<Prim<R>>& prim, std::vector<R>* grad) {
  const int nc = 2;
  const std::size_t N = static_cast<std::size_t>(m.num_cells);
  std::vector<R> acc(N * 13, R(0));  // a, b, c, then r_kx, r_ky for k = 0..4
  for (int f = 0; f < m.num_faces; ++f) {
    const int o = m.owner[f];
    const int nb = m.neighbor[f];
    if (nb < 0) continue;
    const R dx = m.cell_center[3 * nb + 0] - m.cell_center[3 * o + 0];
    const R dy = m.cell_center[3 * nb + 1] - m.cell_center[3 * o + 1];
    const R w = AO_K(1.0) / (dx * dx + dy * dy);
    const R wdx = w * dx;
    const R wdy = w * dy;
    R t[13];
    t[0] = wdx * dx;
    t[1] = wdx * dy;
    t[2] = wdy * dy;
    for (int k = 0; k < 5; ++k) {
      const R dq = prim_var(prim[nb], k) - prim_var(prim[o], k);
      t[3 + 2 * k] = wdx * dq;
      t[4 + 2 * k] = wdy * dq;
    }
    for (int j = 0; j < 13; ++j) {
      acc[static_cast<std::size_t>(o) * 13 + j] += t[j];
      acc[static_cast<std::size_t>(nb) * 13 + j] += t[j];
    }
  }
  grad->assign(N * 5 * nc, R(0));
  for (std::size_t c = 0; c < N; ++c) {
    const R* q = &acc[c * 13];
    const R a = q[0], b = q[1], cc = q[2];
    const R det = a * cc - b * b;
    const R tr = a + cc;
    if (!(det > AO_K(1.0e-12) * (tr * tr))) continue;
    const R inv = AO_K(1.0) / det;
    for (int k = 0; k < 5; ++k) {
      const R rx = q[3 + 2 * k], ry = q[4 + 2 * k];
      (*grad)[(c * 5 + k) * nc + 0] = (cc * rx - b * ry) * inv;
      (*grad)[(c * 5 + k) * nc + 1] = (a * ry - b * rx) * inv;
    }
  }
}

// The cell gradients the mesh asks for.
template <class R>
void cell_gradients(const MeshView<R>& m, const std::vector<Prim<R>>& prim, std::vector<R>* grad) {
  if (m.gradient == 1)
    least_squares(m, prim, grad);
  else
    green_gauss(m, prim, grad);
}

// MUSCL face states (reconstruction.cpp:148-219).
// Frozen limiter (an option beyond the reference, off unless ratios are given): `frozen` holds, per
// face f, the ratios phi / a of the limited to the unlimited increment captured once (layout
// [j][F], j = 0..3 owner side rho, u, v, p and 4..7 the neighbour side); the reconstruction then uses
// q + ratio * a, a smooth function of the state.  `capture` (may be null) receives those ratios.
inline int frozen_slot(int k) { return k == 4 ? 3 : k; }  // rho, u, v, -, p -> 0, 1, 2, -, 3

template <class R>
void reconstruct(const MeshView<R>& m, const std::vector<Prim<R>>& prim, const std::vector<R>& grad, int f,
                 int lim, Prim<R>* L, Prim<R>* Rt, const double* frozen = nullptr, double* capture = nullptr) {
  const int nc = 2;
  const int o = m.owner[f];
  const int nb = m.neighbor[f];
  const A3<R> fc = {m.face_center[3 * f + 0], m.face_center[3 * f + 1], m.face_center[3 * f + 2]};
  const A3<R> oc = {m.cell_center[3 * o + 0], m.cell_center[3 * o + 1], m.cell_center[3 * o + 2]};
  const A3<R> nbc = {m.cell_center[3 * nb + 0], m.cell_center[3 * nb + 1], m.cell_center[3 * nb + 2]};
  const A3<R> dO = {fc[0] - oc[0], fc[1] - oc[1], fc[2] - oc[2]};
  const A3<R> dN = {fc[0] - nbc[0], fc[1] - nbc[1], fc[2] - nbc[2]};
  *L = prim[o];
  *Rt = prim[nb];
  for (int k = 0; k < 5; ++k) {
    const R qo = prim_var(prim[o], k);
    const R qn = prim_var(prim[nb], k);
    const R dc = qn - qo;
    R da = R(0), db = R(0);
    for (int d = 0; d < nc; ++d) {
      da += grad[static_cast<std::size_t>((o * 5 + k) * nc + d)] * dO[d];
      db += grad[static_cast<std::size_t>((nb * 5 + k) * nc + d)] * dN[d];
    }
    R la, lb;
    if (frozen && k != 3) {
      const std::size_t F = static_cast<std::size_t>(m.num_faces);
      la = R(frozen[frozen_slot(k) * F + f]) * da;
      lb = R(frozen[(4 + frozen_slot(k)) * F + f]) * db;
    } else {
      la = apply_limiter(lim, da, dc);
      lb = apply_limiter(lim, db, -dc);
    }
    if (capture && k != 3) {
      const std::size_t F = static_cast<std::size_t>(m.num_faces);
      capture[frozen_slot(k) * F + f] = da != R(0) ? ao_value(la / da) : 0.0;
      capture[(4 + frozen_slot(k)) * F + f] = db != R(0) ? ao_value(lb / db) : 0.0;
    }
    const R ql = qo + la;
    const R qr = qn + lb;
    set_prim_var(L, k, ql);
    set_prim_var(Rt, k, qr);
  }
  L->rho = smax(L->rho, AO_K(1.0e-8));
  Rt->rho = smax(Rt->rho, AO_K(1.0e-8));
  L->p = smax(L->p, AO_K(1.0e-8));
  Rt->p = smax(Rt->p, AO_K(1.0e-8));
}

// ---------------------------------------------------------------------------------------------
// Low-Mach preconditioning (src/cfd_core/solvers/euler/preconditioning.cpp)
// ---------------------------------------------------------------------------------------------
template <class R> struct PrecondCfg {  // preconditioning.hpp LowMachPreconditionConfig
  bool enabled = true;
  R mach_ref = AO_K(0.15);
  R mach_min = AO_K(1.0e-3);
  R beta_min = AO_K(0.2);
  R beta_max = AO_K(1.0);
};

template <class R> R local_mach(const Prim<R>& p, R g) {  // preconditioning.cpp:141-146
  const R a = smax(speed_of_sound(p, g), AO_K(1.0e-8));
  const R vm = ao_sqrt(p.u * p.u + p.v * p.v + p.w * p.w);
  return vm / a;
}

template <class R> R cell_beta(const Prim<R>& p, R g, const PrecondCfg<R>& c) {  // preconditioning.cpp:134-156
  if (!c.enabled) return AO_K(1.0);
  const R lm = local_mach(p, g);
  const R t = smax4(lm, c.mach_ref, c.mach_min, AO_K(1.0e-8));
  const R bmin = smax(c.beta_min, AO_K(1.0e-4));
  const R bmax = smax(bmin, c.beta_max);
  return sclamp(t, bmin, bmax);
}

template <class R>
R face_beta(const Cons<R>& l, const Cons<R>& r, R g, const PrecondCfg<R>& c) {  // preconditioning.cpp:158-173
  if (!c.enabled) return AO_K(1.0);
  const Prim<R> pl = cons_to_prim(l, g);
  const Prim<R> pr = cons_to_prim(r, g);
  Prim<R> a;
  a.rho = AO_K(0.5) * (pl.rho + pr.rho);
  a.u = AO_K(0.5) * (pl.u + pr.u);
  a.v = AO_K(0.5) * (pl.v + pr.v);
  a.w = AO_K(0.5) * (pl.w + pr.w);
  a.p = AO_K(0.5) * (pl.p + pr.p);
  return cell_beta(a, g, c);
}

// R_E' = beta^2 R_E + (beta^2 - 1)(u.R_mom)/rho, residual AoS [cell][5] (preconditioning.cpp:175-200).
template <class R>
void precond_residual(const std::vector<Prim<R>>& prim, R g, const PrecondCfg<R>& c, std::vector<R>* res) {
  if (!c.enabled) return;
  const int n = static_cast<int>(prim.size());
  for (int i = 0; i < n; ++i) {
    const Prim<R>& p = prim[i];
    const R b = cell_beta(p, g, c);
    const R b2 = b * b;
    const R rho = smax(p.rho, AO_K(1.0e-8));
    const R inv = AO_K(1.0) / rho;
    R* r = &(*res)[static_cast<std::size_t>(i) * 5];
    const R udr = p.u * r[1] + p.v * r[2] + p.w * r[3];
    r[4] = b2 * r[4] + (b2 - AO_K(1.0)) * inv * udr;
  }
}

// ---------------------------------------------------------------------------------------------
// Face acoustic scale (src/cfd_core/solvers/euler/residual_assembly.cpp:16-51)
// ---------------------------------------------------------------------------------------------
template <class R> struct AssemblyCfg {  // residual_assembly.hpp:12-31 (pointer members split out)
  Cons<R> farfield;
  R gamma = AO_K(1.4);
  int scheme = kHllc;
  int limiter = kMinmod;
  bool second_order = true;
  bool low_mach_fix = false;
  R low_mach_cutoff = AO_K(0.3);
  R stab_mach_floor = R(0);
  bool precond_farfield_bc = false;
  PrecondCfg<R> precond;
  bool as_fix = false;
  R as_cutoff = AO_K(0.25);
  R as_fmin = AO_K(0.05);
  R as_ramp = AO_K(1.0);
  const double* frozen = nullptr;  // frozen limiter ratios (see reconstruct), or null
};

template <class R> R smooth_max(R a, R b) {  // residual_assembly.cpp:12-20
  const R d = a - b;
  return AO_K(0.5) * (a + b + ao_sqrt(d * d + AO_K(1.0e-6) * AO_K(1.0e-6)));
}

template <class R> R face_mach(const Cons<R>& l, const Cons<R>& r, R g) {  // residual_assembly.cpp:22-29
  const Prim<R> pl = cons_to_prim(l, g);
  const Prim<R> pr = cons_to_prim(r, g);
  return AO_K(0.5) * (local_mach(pl, g) + local_mach(pr, g));
}

template <class R>
R acoustic_scale(const AssemblyCfg<R>& c, const Cons<R>& l, const Cons<R>& r, R fb) {  // residual_assembly.cpp:31-51
  R s = AO_K(1.0);
  if (c.precond.enabled) s = sclamp(fb, AO_K(1.0e-4), AO_K(1.0));
  if (c.low_mach_fix) {
    const R lm = low_mach_dissipation_scale(l, r, c.gamma, true, c.low_mach_cutoff, c.precond.mach_ref);
    s = smin(s, lm);
  }
  s = sclamp(s, AO_K(1.0e-4), AO_K(1.0));
  const R fm = face_mach(l, r, c.gamma);
  const R fl = sclamp(c.stab_mach_floor, R(0), AO_K(1.0));
  const R me = smooth_max(fm, fl);
  const R st = sclamp(me, AO_K(1.0e-4), AO_K(1.0));
  return smax(s, st);
}

template <class R> bool valid_prim(const Prim<R>& p) {  // residual_assembly.cpp:53-57
  return ao_isfinite(p.rho) && ao_isfinite(p.u) && ao_isfinite(p.v) && ao_isfinite(p.w) &&
         ao_isfinite(p.p) && p.rho > AO_K(1.0e-8) && p.p > AO_K(1.0e-8);
}

// ---------------------------------------------------------------------------------------------
// Boundary fluxes (src/cfd_core/solvers/euler/boundary_conditions.cpp)
// ---------------------------------------------------------------------------------------------
template <class R> R precond_spectral_radius(R un, R a, R s) {  // boundary_conditions.cpp:14-25
  return precond_speeds(un, a, s).sr;
}

template <class R>
Cons<R> precond_farfield_state(const Cons<R>& in, const Cons<R>& ff, const A3<R>& n, R g,
                               R s) {  // boundary_conditions.cpp:27-78
  const Prim<R> pi = cons_to_prim(in, g);
  const Prim<R> pf = cons_to_prim(ff, g);
  const R uni = pi.u * n[0] + pi.v * n[1] + pi.w * n[2];
  const R unf = pf.u * n[0] + pf.v * n[1] + pf.w * n[2];
  const R ai = smax(speed_of_sound(pi, g), AO_K(1.0e-8));
  const R af = smax(speed_of_sound(pf, g), AO_K(1.0e-8));
  const R rref = smax(AO_K(0.5) * (pi.rho + pf.rho), AO_K(1.0e-8));
  const R aref = smax(AO_K(0.5) * (ai + af), AO_K(1.0e-8));
  const R at = smax(sclamp(s, AO_K(1.0e-4), AO_K(1.0)) * aref, AO_K(1.0e-6));
  if (ao_abs(uni) >= at) return (uni >= R(0)) ? in : ff;
  const R rp = uni + pi.p / (rref * at);
  const R rm = unf - pf.p / (rref * at);
  const R unb = AO_K(0.5) * (rp + rm);
  const R pb = smax(AO_K(0.5) * rref * at * (rp - rm), AO_K(1.0e-8));
  const A3<R> ti = {pi.u - uni * n[0], pi.v - uni * n[1], pi.w - uni * n[2]};
  const A3<R> tf = {pf.u - unf * n[0], pf.v - unf * n[1], pf.w - unf * n[2]};
  const bool inflow = unb < R(0);
  const A3<R>& t = inflow ? tf : ti;
  const R rb = inflow ? pf.rho : pi.rho;
  const R pbase = inflow ? pf.p : pi.p;
  const R ab = inflow ? af : ai;
  Prim<R> b;
  b.rho = smax(rb + (pb - pbase) / smax(ab * ab, AO_K(1.0e-8)), AO_K(1.0e-8));
  b.u = unb * n[0] + t[0];
  b.v = unb * n[1] + t[1];
  b.w = unb * n[2] + t[2];
  b.p = pb;
  return prim_to_cons(b, g);
}

template <class R>
A5<R> boundary_flux(int bc, const Cons<R>& in, const Cons<R>& ff, const A3<R>& n, int scheme, R g, R s,
                    bool char_bc, bool as_fix, R as_cut, R as_fmin, R as_w, R* lam,
                    bool* fallback) {  // boundary_conditions.cpp:81-123
  if (fallback) *fallback = false;
  if (bc == kSlipWall) {
    const Prim<R> p = cons_to_prim(in, g);
    const R un = p.u * n[0] + p.v * n[1] + p.w * n[2];
    const R a = speed_of_sound(p, g);
    *lam = precond_spectral_radius(un, a, s);
    return {R(0), p.p * n[0], p.p * n[1], p.p * n[2], R(0)};
  }
  const Cons<R> b = char_bc ? precond_farfield_state(in, ff, n, g, s) : ff;
  return face_flux(in, b, n, g, scheme, lam, s, as_fix, as_cut, as_fmin, as_w, fallback);
}

// ---------------------------------------------------------------------------------------------
// Residual assembly: serial face loop, increasing face index (residual_assembly.cpp:73-182).
// residual is AoS [cell][5], lambda [cell]; both accumulated from zero.
// ---------------------------------------------------------------------------------------------
struct AssemblyDiag {
  int first_order_fallback = 0;
  int diffusive_fallback = 0;
  int first_failure_face = -1;
};

// The frozen limiter's storage resolution (no reference counterpart): a captured ratio r in [0, 1] is kept
// as the nearest multiple of 2^-15, ties to even (NaN -> 0), so it can be held as a 16-bit integer.
inline double quantize_ratio(double r) {
  const double c = (r < 0.0) ? 0.0 : (1.0 < r) ? 1.0 : r;
  return (c == c ? std::nearbyint(c * 32768.0) : 0.0) * (1.0 / 32768.0);
}

// Limiter ratios of every interior face at a state (see reconstruct); boundary faces keep 0.  quantized:
// held at the frozen limiter's 2^-15 resolution (its definition); false gives the exact ratios.
template <class R>
void limiter_ratios(const MeshView<R>& m, const std::vector<Prim<R>>& prim, const std::vector<R>& grad, int lim,
                    std::vector<double>* out, bool quantized = true) {
  out->assign(8 * static_cast<std::size_t>(m.num_faces), 0.0);
  for (int f = 0; f < m.num_faces; ++f) {
    if (m.neighbor[f] < 0) continue;
    Prim<R> L, Rt;
    reconstruct(m, prim, grad, f, lim, &L, &Rt, nullptr, out->data());
  }
  if (quantized)
    for (double& r : *out) r = quantize_ratio(r);
}

// One face of the assembly loop (residual_assembly.cpp:97-166): its flux F and wave speed w (before the
// area), whether the reconstruction fell back to first order (*fo) and whether the flux used its
// diffusive fallback (*fbk).
template <class R>
void face_contribution(const MeshView<R>& m, const AssemblyCfg<R>& c, const std::vector<Prim<R>>& prim,
                       const std::vector<R>& grad, int f, A5<R>* Fout, R* wout, bool* fo, bool* fbk) {
  const R g = c.gamma;
  const int o = m.owner[f];
  const int nb = m.neighbor[f];
  const A3<R> n = normal_of(m, f);
  A5<R> F{};
  R w = R(0);
  *fo = false;
  *fbk = false;
  if (nb >= 0) {
    Prim<R> L = prim[o], Rt = prim[nb];
    if (c.second_order) {
      reconstruct(m, prim, grad, f, c.limiter, &L, &Rt, c.frozen);
      if (!valid_prim(L) || !valid_prim(Rt)) {
        *fo = true;
        L = prim[o];
        Rt = prim[nb];
      }
    }
    const Cons<R> ul = prim_to_cons(L, g);
    const Cons<R> ur = prim_to_cons(Rt, g);
    const R fb = face_beta(ul, ur, g, c.precond);
    const R s = acoustic_scale(c, ul, ur, fb);
    F = face_flux(ul, ur, n, g, c.scheme, &w, s, c.as_fix, c.as_cutoff, c.as_fmin, c.as_ramp, fbk);
  } else {
    const Cons<R> in = prim_to_cons(prim[o], g);
    const int bc = m.bc[f];
    Cons<R> ex = c.farfield;
    R fb;
    if (bc == kSlipWall) {
      fb = face_beta(in, in, g, c.precond);
      ex = in;
    } else {
      fb = face_beta(in, c.farfield, g, c.precond);
    }
    const R s = acoustic_scale(c, in, ex, fb);
    F = boundary_flux(bc, in, c.farfield, n, c.scheme, g, s, c.precond_farfield_bc, c.as_fix, c.as_cutoff,
                      c.as_fmin, c.as_ramp, &w, fbk);
  }
  *Fout = F;
  *wout = w;
}

template <class R>
void assemble_residual(const MeshView<R>& m, const AssemblyCfg<R>& c, const std::vector<Prim<R>>& prim,
                       const std::vector<R>& grad, std::vector<R>* res, std::vector<R>* lam,
                       AssemblyDiag* diag, std::vector<R>* abs_sum = nullptr) {
  // abs_sum (test scale only, not a reference output): per cell and component, sum over faces of |F A|.
  if (abs_sum) abs_sum->assign(static_cast<std::size_t>(m.num_cells) * 5, R(0));
  res->assign(static_cast<std::size_t>(m.num_cells) * 5, R(0));
  lam->assign(static_cast<std::size_t>(m.num_cells), R(0));
  if (diag) *diag = AssemblyDiag{};
  for (int f = 0; f < m.num_faces; ++f) {
    const int o = m.owner[f];
    const int nb = m.neighbor[f];
    const R A = m.area[f];
    A5<R> F;
    R w;
    bool fo, fbk;
    face_contribution(m, c, prim, grad, f, &F, &w, &fo, &fbk);
    if (diag && fo) {
      if (diag->first_failure_face < 0) diag->first_failure_face = f;
      ++diag->first_order_fallback;
    }
    if (diag && fbk) ++diag->diffusive_fallback;
    const R wave = w * A;
    (*lam)[o] += wave;
    for (int k = 0; k < 5; ++k) {
      const R fa = F[k] * A;
      (*res)[static_cast<std::size_t>(5 * o + k)] += fa;
      if (nb >= 0) (*res)[static_cast<std::size_t>(5 * nb + k)] -= fa;
      if (abs_sum) {
        (*abs_sum)[static_cast<std::size_t>(5 * o + k)] += ao_abs(fa);
        if (nb >= 0) (*abs_sum)[static_cast<std::size_t>(5 * nb + k)] += ao_abs(fa);
      }
    }
    if (nb >= 0) (*lam)[nb] += wave;
  }
}

// ---- the same evaluation on several host threads (OpenMP; serial without -fopenmp) ----------------
// Per-face values in parallel, then per cell the sum over its faces in increasing face index: the serial
// loops' accumulation order (reconstruction.cpp:85-107, residual_assembly.cpp:97-181), so the results are
// the serial ones bit for bit.
struct CellFaces {
  std::vector<int> off, ent;  // per cell, its faces in increasing index: face << 1 | (cell is the neighbour)
};

template <class R> CellFaces cell_faces(const MeshView<R>& m) {
  CellFaces cf;
  cf.off.assign(static_cast<std::size_t>(m.num_cells) + 1, 0);
  for (int f = 0; f < m.num_faces; ++f) {
    ++cf.off[m.owner[f] + 1];
    if (m.neighbor[f] >= 0) ++cf.off[m.neighbor[f] + 1];
  }
  for (int c = 0; c < m.num_cells; ++c) cf.off[c + 1] += cf.off[c];
  cf.ent.resize(cf.off[m.num_cells]);
  std::vector<int> fill(cf.off.begin(), cf.off.end() - 1);
  for (int f = 0; f < m.num_faces; ++f) {
    cf.ent[fill[m.owner[f]]++] = f << 1;
    if (m.neighbor[f] >= 0) cf.ent[fill[m.neighbor[f]]++] = (f << 1) | 1;
  }
  return cf;
}

// Gradients of the second-order reconstruction: Green-Gauss face values in parallel, then the ordered
// per-cell sums; least squares runs the serial routine.  qf is scratch.
template <class R>
void cell_gradients_threaded(const MeshView<R>& v, const CellFaces& cf, const std::vector<Prim<R>>& p,
                             std::vector<R>* qf, std::vector<R>* gr) {
  if (v.gradient == 1) {
    cell_gradients(v, p, gr);
    return;
  }
  const int N = v.num_cells, F = v.num_faces;
  qf->resize(static_cast<std::size_t>(F) * 5);
  gr->assign(static_cast<std::size_t>(N) * 10, R(0));
#pragma omp parallel for schedule(static)
  for (int f = 0; f < F; ++f)
    for (int k = 0; k < 5; ++k) {
      const R qo = prim_var(p[v.owner[f]], k);
      R q = qo;
      if (v.neighbor[f] >= 0) q = AO_K(0.5) * (qo + prim_var(p[v.neighbor[f]], k));
      (*qf)[static_cast<std::size_t>(5 * f + k)] = q;
    }
#pragma omp parallel for schedule(static)
  for (int c = 0; c < N; ++c) {
    const R iv = AO_K(1.0) / smax(v.volume[c], AO_K(1.0e-12));
    for (int k = 0; k < 5; ++k)
      for (int dd = 0; dd < 2; ++dd) {
        R acc = R(0);
        for (int e = cf.off[c]; e < cf.off[c + 1]; ++e) {
          const int f = cf.ent[e] >> 1;
          const R nA = v.normal[3 * f + dd] * v.area[f];
          const R t = (*qf)[static_cast<std::size_t>(5 * f + k)] * nA;
          if (cf.ent[e] & 1) acc -= t; else acc += t;
        }
        (*gr)[static_cast<std::size_t>((c * 5 + k) * 2 + dd)] = acc * iv;
      }
  }
}

// assemble_residual with the face loop in parallel and the per-cell sums in face order.  fa is scratch.
template <class R>
void assemble_residual_threaded(const MeshView<R>& v, const AssemblyCfg<R>& a, const CellFaces& cf,
                                const std::vector<Prim<R>>& p, const std::vector<R>& gr, std::vector<R>* fa,
                                std::vector<R>* r, std::vector<R>* l, AssemblyDiag* d) {
  const int N = v.num_cells, F = v.num_faces;
  fa->resize(static_cast<std::size_t>(F) * 6);
  r->resize(static_cast<std::size_t>(N) * 5);
  l->resize(N);
  int n_fo = 0, n_diff = 0, first = F;
#pragma omp parallel for schedule(static) reduction(+ : n_fo, n_diff) reduction(min : first)
  for (int f = 0; f < F; ++f) {
    A5<R> Fl;
    R w;
    bool fo, fbk;
    face_contribution(v, a, p, gr, f, &Fl, &w, &fo, &fbk);
    const R A = v.area[f];
    for (int k = 0; k < 5; ++k) (*fa)[static_cast<std::size_t>(6 * f + k)] = Fl[k] * A;
    (*fa)[static_cast<std::size_t>(6 * f + 5)] = w * A;
    if (fo) {
      ++n_fo;
      first = std::min(first, f);
    }
    if (fbk) ++n_diff;
  }
#pragma omp parallel for schedule(static)
  for (int c = 0; c < N; ++c) {
    R acc[5] = {R(0), R(0), R(0), R(0), R(0)};
    R w = R(0);
    for (int e = cf.off[c]; e < cf.off[c + 1]; ++e) {
      const std::size_t f = static_cast<std::size_t>(cf.ent[e] >> 1);
      for (int k = 0; k < 5; ++k) {
        if (cf.ent[e] & 1) acc[k] -= (*fa)[6 * f + k]; else acc[k] += (*fa)[6 * f + k];
      }
      w += (*fa)[6 * f + 5];
    }
    for (int k = 0; k < 5; ++k) (*r)[static_cast<std::size_t>(5 * c + k)] = acc[k];
    (*l)[c] = w;
  }
  if (d) {
    d->first_order_fallback = n_fo;
    d->diffusive_fallback = n_diff;
    d->first_failure_face = first < F ? first : -1;
  }
}

// ---------------------------------------------------------------------------------------------
// Time integration (src/cfd_core/solvers/euler/time_integrator.cpp:31-104)
// ---------------------------------------------------------------------------------------------
template <class R> std::vector<R> dt_over_volume(const std::vector<R>& lam, R cfl, bool lts) {  // :31-59
  std::vector<R> dt(lam.size(), R(0));
  if (lam.empty()) return dt;
  if (lts) {
    for (std::size_t i = 0; i < lam.size(); ++i) dt[i] = cfl / smax(lam[i], AO_K(1.0e-8));
    return dt;
  }
  R gdt = std::numeric_limits<R>::max();
  for (R x : lam) gdt = smin(gdt, cfl / smax(x, AO_K(1.0e-8)));
  if (!ao_isfinite(gdt)) gdt = R(0);
  std::fill(dt.begin(), dt.end(), gdt);
  return dt;
}

template <class R>
void rk3_stage_update(int stage, const std::vector<R>& un, const std::vector<R>& us, const std::vector<R>& rp,
                      const std::vector<R>& dt, R g, std::vector<R>* next) {  // :61-104
  const std::size_t n = dt.size();
  next->assign(n * 5, R(0));
  for (std::size_t i = 0; i < n; ++i) {
    Cons<R> s{us[5 * i + 0], us[5 * i + 1], us[5 * i + 2], us[5 * i + 3], us[5 * i + 4]};
    const R d = dt[i];
    s.rho -= d * rp[5 * i + 0];
    s.rhou -= d * rp[5 * i + 1];
    s.rhov -= d * rp[5 * i + 2];
    s.rhow -= d * rp[5 * i + 3];
    s.rhoE -= d * rp[5 * i + 4];
    if (stage == 1) {
      const R* b = &un[5 * i];
      s.rho = AO_K(0.75) * b[0] + AO_K(0.25) * s.rho;
      s.rhou = AO_K(0.75) * b[1] + AO_K(0.25) * s.rhou;
      s.rhov = AO_K(0.75) * b[2] + AO_K(0.25) * s.rhov;
      s.rhow = AO_K(0.75) * b[3] + AO_K(0.25) * s.rhow;
      s.rhoE = AO_K(0.75) * b[4] + AO_K(0.25) * s.rhoE;
    } else if (stage >= 2) {
      const R* b = &un[5 * i];
      const R t1 = AO_K(1.0) / AO_K(3.0);
      const R t2 = AO_K(2.0) / AO_K(3.0);
      s.rho = t1 * b[0] + t2 * s.rho;
      s.rhou = t1 * b[1] + t2 * s.rhou;
      s.rhov = t1 * b[2] + t2 * s.rhov;
      s.rhow = t1 * b[3] + t2 * s.rhow;
      s.rhoE = t1 * b[4] + t2 * s.rhoE;
    }
    enforce_physical_state(&s, g);
    R* o = &(*next)[5 * i];
    o[0] = s.rho;
    o[1] = s.rhou;
    o[2] = s.rhov;
    o[3] = s.rhow;
    o[4] = s.rhoE;
  }
}

// ---------------------------------------------------------------------------------------------
// Residual norms (src/cfd_core/solvers/euler_solver.cpp:309-339); rho_l2 is the density-only L2
// over component 0 (no reference function: it is the same sum restricted to k = 0, over N).
// ---------------------------------------------------------------------------------------------
template <class R> struct Norms { R l1 = 0, l2 = 0, linf = 0, rho_l2 = 0; };

template <class R> Norms<R> residual_norms(const MeshView<R>& m, const std::vector<R>& res) {
  double l1 = 0.0, l2s = 0.0, linf = 0.0, r2 = 0.0;
  for (int c = 0; c < m.num_cells; ++c) {
    const R iv = AO_K(1.0) / smax(m.volume[c], AO_K(1.0e-12));
    for (int k = 0; k < 5; ++k) {
      const R v = res[static_cast<std::size_t>(5 * c + k)] * iv;
      const double mag = ao_abs(static_cast<double>(v));
      l1 += mag;
      l2s += mag * mag;
      linf = std::max(linf, mag);
      if (k == 0) r2 += mag * mag;
    }
  }
  Norms<R> n;
  n.l1 = static_cast<R>(l1);
  n.l2 = static_cast<R>(ao_sqrt(l2s / std::max(1, m.num_cells * 5)));
  n.linf = static_cast<R>(linf);
  n.rho_l2 = static_cast<R>(ao_sqrt(r2 / std::max(1, m.num_cells)));
  return n;
}

// ---------------------------------------------------------------------------------------------
// Forces (src/cfd_core/post/forces.cpp:177-294) over the contiguous "wall" patch range.
// ---------------------------------------------------------------------------------------------
template <class R> struct FreeRef {  // forces.hpp FreestreamReference
  R rho_inf = AO_K(1.225), p_inf = AO_K(101325.0), aoa_deg = R(0), speed_inf = AO_K(1.0), chord = AO_K(1.0),
    x_ref = AO_K(0.25), y_ref = R(0);
};
template <class R> struct Coeffs { R fx = 0, fy = 0, lift = 0, drag = 0, moment = 0, cl = 0, cd = 0, cm = 0; };

template <class R> R pi_const() { return kc<R>(3.14159265358979323846f, 3.14159265358979323846); }

template <class R>
Coeffs<R> integrate_forces(const MeshView<R>& m, const std::vector<R>& p, const FreeRef<R>& ref,
                           bool subtract) {  // forces.cpp:208-294
  R fx_abs = 0, fy_abs = 0, fx_g = 0, fy_g = 0;
  for (int i = 0; i < m.wall_count; ++i) {  // compute_pressure_force_diagnostics :220-246
    const int f = m.wall_start + i;
    const R A = m.area[f];
    const R nA0 = -m.normal[3 * f + 0] * A;
    const R nA1 = -m.normal[3 * f + 1] * A;
    const R pp = p[m.owner[f]];
    const R gp = pp - ref.p_inf;
    fx_abs -= pp * nA0;
    fy_abs -= pp * nA1;
    fx_g -= gp * nA0;
    fy_g -= gp * nA1;
  }
  const R q = AO_K(0.5) * ref.rho_inf * ref.speed_inf * ref.speed_inf + AO_K(1.0e-12);
  const R alpha = ref.aoa_deg * (pi_const<R>() / AO_K(180.0));
  const R ca = std::cos(alpha);
  const R sa = std::sin(alpha);
  const R chord = smax(ref.chord, AO_K(1.0e-8));
  const R sref = chord * AO_K(1.0);
  Coeffs<R> c;
  c.fx = subtract ? fx_g : fx_abs;
  c.fy = subtract ? fy_g : fy_abs;
  for (int i = 0; i < m.wall_count; ++i) {
    const int f = m.wall_start + i;
    const R pp = p[m.owner[f]];
    const R pr = subtract ? (pp - ref.p_inf) : pp;
    const R A = m.area[f];
    const R nA0 = -m.normal[3 * f + 0] * A;
    const R nA1 = -m.normal[3 * f + 1] * A;
    const R dfx = -pr * nA0;
    const R dfy = -pr * nA1;
    const R x = m.face_center[3 * f + 0] - ref.x_ref;
    const R y = m.face_center[3 * f + 1] - ref.y_ref;
    c.moment += x * dfy - y * dfx;
  }
  c.drag = c.fx * ca + c.fy * sa;
  c.lift = -c.fx * sa + c.fy * ca;
  c.cd = c.drag / (q * sref);
  c.cl = c.lift / (q * sref);
  c.cm = c.moment / (q * sref * chord);
  return c;
}

// ---------------------------------------------------------------------------------------------
// Convergence monitor (src/cfd_core/solvers/euler/convergence_monitor.cpp:7-127)
// ---------------------------------------------------------------------------------------------
template <class R> struct MonitorCfg {
  R ratio_target = AO_K(1.0e-3);
  R force_tol = AO_K(2.0e-5);
  int window = 6;
  int min_iterations = 40;
  R drift_tol = AO_K(2.0e-5);
};

template <class R> class Monitor {
 public:
  explicit Monitor(MonitorCfg<R> c) : c_(c) {
    if (c_.drift_tol <= R(0)) c_.drift_tol = c_.force_tol;
  }
  R update_residual(R l2) {  // :13-19
    const R b = smax(l2, AO_K(1.0e-20));
    if (init_ < R(0)) init_ = b;
    return b / smax(init_, AO_K(1.0e-20));
  }
  void update_forces(const Coeffs<R>& f) {  // :21-37
    const int req = std::max(c_.window, 1);
    win_.push_back(f);
    while (static_cast<int>(win_.size()) > 2 * req) win_.pop_front();
  }
  bool forces_stable() const {  // :39-67
    const int req = std::max(c_.window, 1);
    if (static_cast<int>(win_.size()) < req) return false;
    const int b = static_cast<int>(win_.size()) - req;
    R cl0 = win_[b].cl, cl1 = win_[b].cl, cd0 = win_[b].cd, cd1 = win_[b].cd, cm0 = win_[b].cm, cm1 = win_[b].cm;
    for (int i = b; i < static_cast<int>(win_.size()); ++i) {
      cl0 = smin(cl0, win_[i].cl);
      cl1 = smax(cl1, win_[i].cl);
      cd0 = smin(cd0, win_[i].cd);
      cd1 = smax(cd1, win_[i].cd);
      cm0 = smin(cm0, win_[i].cm);
      cm1 = smax(cm1, win_[i].cm);
    }
    return (cl1 - cl0) <= c_.force_tol && (cd1 - cd0) <= c_.force_tol && (cm1 - cm0) <= c_.force_tol;
  }
  bool drift_stable() const {  // :69-113
    const int req = std::max(c_.window, 1);
    const int need = 2 * req;
    if (static_cast<int>(win_.size()) < need) return false;
    const int pb = static_cast<int>(win_.size()) - need;
    const int cb = static_cast<int>(win_.size()) - req;
    R a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0;
    for (int i = pb; i < cb; ++i) {
      a0 += win_[i].cl;
      a1 += win_[i].cd;
      a2 += win_[i].cm;
    }
    for (int i = cb; i < static_cast<int>(win_.size()); ++i) {
      b0 += win_[i].cl;
      b1 += win_[i].cd;
      b2 += win_[i].cm;
    }
    const R inv = AO_K(1.0) / static_cast<R>(req);
    const R d0 = ao_abs(b0 * inv - a0 * inv);
    const R d1 = ao_abs(b1 * inv - a1 * inv);
    const R d2 = ao_abs(b2 * inv - a2 * inv);
    return d0 <= c_.drift_tol && d1 <= c_.drift_tol && d2 <= c_.drift_tol;
  }
  bool converged(int iter, R ratio) const {  // :119-127
    if (iter + 1 < c_.min_iterations) return false;
    if (ratio > c_.ratio_target) return false;
    return forces_stable() && drift_stable();
  }

 private:
  MonitorCfg<R> c_;
  R init_ = R(-1);
  std::deque<Coeffs<R>> win_;
};

// ---------------------------------------------------------------------------------------------
// Mesh preparation used by the driver (src/cfd_core/solvers/euler/mesh_geometry.cpp:49-80).
// ---------------------------------------------------------------------------------------------
template <class R> int normalize_face_normals(int nf, R* normal, R tol) {
  int invalid = 0;
  for (int f = 0; f < nf; ++f) {
    const R nx = normal[3 * f + 0], ny = normal[3 * f + 1], nz = normal[3 * f + 2];
    const R nrm = ao_sqrt(nx * nx + ny * ny + nz * nz);
    if (!ao_isfinite(nrm) || nrm < AO_K(1.0e-12)) {
      ++invalid;
      normal[3 * f + 0] = AO_K(1.0);
      normal[3 * f + 1] = R(0);
      normal[3 * f + 2] = R(0);
      continue;
    }
    const R dev = ao_abs(nrm - AO_K(1.0));
    if (dev > tol) {
      const R inv = AO_K(1.0) / nrm;
      normal[3 * f + 0] = nx * inv;
      normal[3 * f + 1] = ny * inv;
      normal[3 * f + 2] = nz * inv;
    }
  }
  return invalid;
}

}  // namespace aero_oracle
