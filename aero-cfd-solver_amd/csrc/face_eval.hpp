// face_eval.hpp -- one face of the residual (residual_assembly.cpp:97-166), templated on the scalar.
//
// The residual kernels instantiate it on double; the tangent and face-Jacobian kernels on Dual, so the
// derivative code is the residual code.
#pragma once

#include "euler_device.hpp"

namespace aero_dev {

// Unlimited reconstruction increment g . d of one primitive variable (reconstruction.cpp:194-213).
template <class T>
__device__ __forceinline__ T recon_incr(T gx, T gy, T dx, T dy) {
  T a = 0.0;
  a += gx * dx;
  a += gy * dy;
  return a;
}

// Sources of the increments of variable k (0..3 = rho, u, v, p) on the owner (da) and neighbour (db)
// sides: from the cells' gradients and centre-to-face vectors, or given directly.
// NOZERO: gx dx + gy dy without the leading 0.0 +, for increments that only enter the limiters: the sum
// differs from recon_incr's at most in the sign of a zero, and both limiters return +0 for a zero
// increment (a * b <= 0).
template <class T, bool NOZERO = false>
struct GradIncr {
  const T* go;
  const T* gn;
  T dox, doy, dnx, dny;
  __device__ __forceinline__ void operator()(int k, T* da, T* db) const {
    if constexpr (NOZERO) {
      *da = go[2 * k] * dox + go[2 * k + 1] * doy;
      *db = gn[2 * k] * dnx + gn[2 * k + 1] * dny;
    } else {
      *da = recon_incr(go[2 * k], go[2 * k + 1], dox, doy);
      *db = recon_incr(gn[2 * k], gn[2 * k + 1], dnx, dny);
    }
  }
};
template <class T>
struct GivenIncr {
  const T* ao;
  const T* an;
  __device__ __forceinline__ void operator()(int k, T* da, T* db) const {
    *da = ao[k];
    *db = an[k];
  }
};

// MUSCL reconstruction of primitive variable k (reconstruction.cpp:194-213).  FAST: the limiters' exact
// shortcuts, *ok cleared where their bounds fail.
template <int LIMITER, bool FAST = false, class T, class Src>
__device__ __forceinline__ void recon_var(T qo, T qn, int k, const Src& src, T* ql, T* qr, bool* ok = nullptr) {
  const T dc = qn - qo;
  T da, db;
  src(k, &da, &db);
  *ql = qo + limiter<LIMITER, FAST>(da, dc, ok);
  *qr = qn + limiter<LIMITER, FAST>(db, -dc, ok);
}

// Frozen limiter (beyond the reference): the limited increment is ratio * a with the ratio
// phi(a, b) / a captured once (0 where a == 0), a smooth function of the state.
template <class T, class Src>
__device__ __forceinline__ void recon_var_frozen(T qo, T qn, int k, const Src& src, double ro, double rn, T* ql,
                                                 T* qr) {
  T da, db;
  src(k, &da, &db);
  *ql = qo + ro * da;
  *qr = qn + rn * db;
}

// The capture: ratio phi(a, b) / a of both sides of one variable, held at 2^-15 (quantize_ratio).
template <int LIMITER>
__device__ __forceinline__ void limiter_ratio(double qo, double qn, double gox, double goy, double gnx, double gny,
                                              double dox, double doy, double dnx, double dny, double* ro, double* rn) {
  const double dc = qn - qo;
  double da = 0.0;
  da += gox * dox;
  da += goy * doy;
  double db = 0.0;
  db += gnx * dnx;
  db += gny * dny;
  *ro = quantize_ratio(da != 0.0 ? limiter<LIMITER>(da, dc) / da : 0.0);
  *rn = quantize_ratio(db != 0.0 ? limiter<LIMITER>(db, -dc) / db : 0.0);
}

// The face flux from the two face states' primitives (residual_assembly.cpp:122-137): conservative maps,
// face beta, acoustic scale and the Riemann flux.  The stored face Jacobians of the frozen adjoint
// differentiate this function of (L, R).
//
// FAST (PLAIN, T = double): the quotients by gamma - 1 and by each side's density go through quot, and the
// wave speeds through pspeeds<FAST>, first.  With 1 < gamma <= 30 (plain_options), conservative
// densities in [kFloor, 2^300] and energies at most 2^280 (checked below; p2c makes them positive and at
// least kFloor / (gamma - 1)), every operand of those shortcuts lies where they are exact: the
// pressures (gamma - 1) (E - ke) are at most 29 E, so gamma p <= 2^290; the kinetic energies are at
// most E, so velocities and normal velocities stay finite; sound speeds lie within [2^-165, 2^165].  An
// energy above the bound can only come from a pressure or velocity above it, whatever p2c_q returned.
// *ok reports whether a face is within the bounds; the flux of one that is not (non-finite or extreme
// states) is meaningless, and the caller evaluates that face again without FAST.
template <int SCHEME, bool PLAIN, bool FAST = false, class T>
__device__ __forceinline__ FluxT<T> flux_of_states(const P4T<T>& L, const P4T<T>& Rt, T nx, T ny,
                                                   const FaceOptsT<T>& o, bool* ok = nullptr) {
  if constexpr (FAST) {
    static_assert(PLAIN && std::is_same_v<T, double>, "FAST face fluxes are PLAIN and in double");
    constexpr double kEnergyHi = 0x1p280;
    const QuotDen dg = quot_den(o.g - 1.0);
    const C4 ql = p2c_q(L, o.g, dg);
    const C4 qr = p2c_q(Rt, o.g, dg);
    *ok = (ql.r <= kQuotHi) & (qr.r <= kQuotHi) & (ql.E <= kEnergyHi) & (qr.E <= kEnergyHi);
    return face_flux<SCHEME, PLAIN, true>(ql, qr, nx, ny, T(1.0), o);
  }
  const C4T<T> ul = p2c(L, o.g);
  const C4T<T> ur = p2c(Rt, o.g);
  const T fb = face_beta<PLAIN>(ul, ur, o);
  const T sc = acoustic_scale<PLAIN>(ul, ur, fb, o);
  return face_flux<SCHEME, PLAIN>(ul, ur, nx, ny, sc, o);
}

// Interior face with the first-order fallback (residual_assembly.cpp:107-137), the reconstruction
// increments taken from src.  *reverted reports the fallback to cell values.
// FROZEN: fr holds the face's frozen limiter ratios (owner side rho, u, v, p, then neighbour side).
template <int SCHEME, int LIMITER, bool SECOND, bool FROZEN, bool PLAIN, bool FAST = false, class T, class Src>
__device__ __forceinline__ FluxT<T> interior_face_src(const C4T<T>& uo, const C4T<T>& un, const Src& src, T nx, T ny,
                                                      const FaceOptsT<T>& o, bool* reverted, const double* fr,
                                                      bool* ok = nullptr) {
  const double g = o.g;
  bool rok = true;  // FAST: the bounds of the cell densities and of the limiters
  P4T<T> L, Rt;
  if constexpr (FAST) {
    L = c2p_cell_q(uo, g, &rok);
    Rt = c2p_cell_q(un, g, &rok);
  } else {
    L = c2p(uo, g);
    Rt = c2p(un, g);
  }
  *reverted = false;
  if (SECOND) {
    P4T<T> l2, r2;
    if (FROZEN) {
      recon_var_frozen(L.r, Rt.r, 0, src, fr[0], fr[4], &l2.r, &r2.r);
      recon_var_frozen(L.u, Rt.u, 1, src, fr[1], fr[5], &l2.u, &r2.u);
      recon_var_frozen(L.v, Rt.v, 2, src, fr[2], fr[6], &l2.v, &r2.v);
      recon_var_frozen(L.p, Rt.p, 3, src, fr[3], fr[7], &l2.p, &r2.p);
    } else {
      recon_var<LIMITER, FAST>(L.r, Rt.r, 0, src, &l2.r, &r2.r, &rok);
      recon_var<LIMITER, FAST>(L.u, Rt.u, 1, src, &l2.u, &r2.u, &rok);
      recon_var<LIMITER, FAST>(L.v, Rt.v, 2, src, &l2.v, &r2.v, &rok);
      recon_var<LIMITER, FAST>(L.p, Rt.p, 3, src, &l2.p, &r2.p, &rok);
    }
    // FAST venkat: a NaN here needs a NaN cell value, which makes dc NaN and fails that limiter's bound
    // (rok), so on the faces kept the floors are maxNum (std::max differs from it only on NaN).
    constexpr bool kMaxNum = FAST && !FROZEN && LIMITER == 1;
    l2.r = fmx<kMaxNum>(l2.r, T(kFloor));
    r2.r = fmx<kMaxNum>(r2.r, T(kFloor));
    l2.p = fmx<kMaxNum>(l2.p, T(kFloor));
    r2.p = fmx<kMaxNum>(r2.p, T(kFloor));
    if (valid_prim(l2) && valid_prim(r2)) {
      L = l2;
      Rt = r2;
    } else {
      *reverted = true;
    }
  }
  const FluxT<T> F = flux_of_states<SCHEME, PLAIN, FAST>(L, Rt, nx, ny, o, ok);
  if constexpr (FAST) *ok = *ok & rok;
  return F;
}

// Interior face (residual_assembly.cpp:107-137).  go / gn are the cells' primitive gradients (rho_x,
// rho_y, u_x, u_y, v_x, v_y, p_x, p_y); dO / dN the vectors from the cell centres to the face centre.
// FAST: see flux_of_states and limiter (*ok).
template <int SCHEME, int LIMITER, bool SECOND, bool FROZEN = false, bool PLAIN = false, bool FAST = false, class T>
__device__ __forceinline__ FluxT<T> interior_face(const C4T<T>& uo, const C4T<T>& un, const T* go, const T* gn,
                                                  T nx, T ny, T dox, T doy, T dnx, T dny, const FaceOptsT<T>& o,
                                                  bool* reverted, const double* fr = nullptr, bool* ok = nullptr) {
  const GradIncr<T, FAST && !FROZEN> src{go, gn, dox, doy, dnx, dny};
  return interior_face_src<SCHEME, LIMITER, SECOND, FROZEN, PLAIN, FAST>(uo, un, src, nx, ny, o, reverted, fr, ok);
}

// The same face from its reconstruction increments ao / an (owner / neighbour g . d of rho, u, v, p):
// the flux depends on the cell gradients only through these eight values, which is the form the stored
// face Jacobians take.
template <int SCHEME, int LIMITER, bool SECOND, bool FROZEN = false, class T>
__device__ __forceinline__ FluxT<T> interior_face_incr(const C4T<T>& uo, const C4T<T>& un, const T* ao, const T* an,
                                                       T nx, T ny, const FaceOptsT<T>& o, bool* reverted,
                                                       const double* fr = nullptr) {
  const GivenIncr<T> src{ao, an};
  return interior_face_src<SCHEME, LIMITER, SECOND, FROZEN, false>(uo, un, src, nx, ny, o, reverted, fr);
}

// Boundary face: slip wall or farfield on the owner's first-order state (residual_assembly.cpp:138-166,
// boundary_conditions.cpp:81-123).
// The same from the owner's primitive state W = c2p(uo) (the frozen adjoint's stored Jacobians
// differentiate it in W).
template <int SCHEME, bool PLAIN = false, class T>
__device__ __forceinline__ FluxT<T> boundary_face_prim(const P4T<T>& W, bool wall, T nx, T ny, const FaceOptsT<T>& o) {
  const double g = o.g;
  const C4T<T> in = p2c(W, g);
  // The exterior state of beta and the acoustic scale: the interior state at a wall, the freestream at a
  // farfield (selected per component, which keeps both in registers).
  C4T<T> ext;
  ext.r = wall ? in.r : o.ff.r;
  ext.ru = wall ? in.ru : o.ff.ru;
  ext.rv = wall ? in.rv : o.ff.rv;
  ext.E = wall ? in.E : o.ff.E;
  const T fb = face_beta<PLAIN>(in, ext, o);
  const T sc = acoustic_scale<PLAIN>(in, ext, fb, o);
  FluxT<T> F;
  if (wall) {  // boundary_conditions.cpp:97-111
    const P4T<T> p = c2p(in, g);
    const T un = p.u * nx + p.v * ny;
    const T a = sound(p, g);
    F.lam = pspeeds(un, a, sc).sr;
    F.f0 = 0.0;
    F.f1 = p.p * nx;
    F.f2 = p.p * ny;
    F.f3 = 0.0;
    F.fallback = 0;
  } else {  // boundary_conditions.cpp:113-122
    C4T<T> bs = ext;
    if (o.char_bc) bs = char_farfield_state(in, nx, ny, sc, o);
    F = face_flux<SCHEME, PLAIN>(in, bs, nx, ny, sc, o);
  }
  return F;
}

template <int SCHEME, bool PLAIN = false, class T>
__device__ __forceinline__ FluxT<T> boundary_face(const C4T<T>& uo, bool wall, T nx, T ny, const FaceOptsT<T>& o) {
  return boundary_face_prim<SCHEME, PLAIN>(c2p(uo, o.g), wall, nx, ny, o);
}

}  // namespace aero_dev
