// aero_adjoint.hip -- gfx950 kernels of the discrete tangent and adjoint of the residual.
//
// The reference has no derivative code; these kernels differentiate the residual the other kernels
// compute (face_eval.hpp, instantiated on the dual-number type), at a fixed state.
//
//   tangent:  dR = R_U dU + R_p dp           k_tan_gradient -> k_tan_face -> k_tan_gather
//   adjoint:  out = R_U^T psi                 k_adj_gather -> k_adj_grad, over the face Jacobians
//             J_f = d(F_f A_f)/d(U_o, U_n, a_o, a_n) (a: the reconstruction increments) that
//             k_face_jacobian stores once per state, each side in its cell's face entry; inside the
//             pseudo-time iteration k_adj_grad_update also applies the stage update
//
// All sums run per cell over its faces in increasing reference face index (no atomics), so results are
// deterministic.  Face-indexed arrays are SoA over the F faces in reference order; cell arrays are SoA
// over the N cells in the device (internal) order.
#include <hip/hip_runtime.h>

#include "aero_internal.hpp"
#include "face_eval.hpp"

namespace aero_dev {

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ Dual dual_at(const double* v, const double* d, int i) {
  return Dual(v[i], d ? d[i] : 0.0);
}

__device__ __forceinline__ C4T<Dual> load_dual_u(const double* U, const double* dU, int ld, int c) {
  C4T<Dual> s;
  s.r = dual_at(U, dU, c);
  s.ru = dual_at(U + ld, dU ? dU + ld : nullptr, c);
  s.rv = dual_at(U + 2 * ld, dU ? dU + 2 * ld : nullptr, c);
  s.E = dual_at(U + 3 * ld, dU ? dU + 3 * ld : nullptr, c);
  return s;
}

__device__ __forceinline__ C4 load_c4(const double* U, int ld, int c) {
  C4 s;
  s.r = U[c];
  s.ru = U[ld + c];
  s.rv = U[2 * ld + c];
  s.E = U[3 * ld + c];
  return s;
}

// ---- tangent of the gradient ---------------------------------------------------------------------
// Same arithmetic as the residual's gradient kernels (Green-Gauss: reconstruction.cpp:73-119; least
// squares: the weighted normal equations), per cell over its faces in face order, on dual numbers.
template <int GK>
__global__ __launch_bounds__(kThreads) void k_tan_gradient(AdjMesh am, const double* __restrict__ U,
                                                           const double* __restrict__ dU, GeoTan gt, double g,
                                                           double* __restrict__ G, double* __restrict__ dG) {
  const int c = blockIdx.x * kThreads + threadIdx.x;
  if (c >= am.N) return;
  const int ld = am.ld;
  Dual acc[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) acc[j] = Dual(0.0);
  const int e1 = am.cf_off[c + 1];
  if (GK == 1) {
    Dual a(0.0), bb(0.0), cc(0.0);
    for (int e = am.cf_off[c]; e < e1; ++e) {
      const int f = am.cf_ent[e] >> 1;
      const int o = am.fo[f], n = am.fn[f];
      if (n < 0) continue;
      const P4T<Dual> wo = c2p(load_dual_u(U, dU, ld, o), g);
      const P4T<Dual> wn = c2p(load_dual_u(U, dU, ld, n), g);
      const Dual dx = dual_at(am.ccx, gt.dccx, n) - dual_at(am.ccx, gt.dccx, o);
      const Dual dy = dual_at(am.ccy, gt.dccy, n) - dual_at(am.ccy, gt.dccy, o);
      const Dual w = 1.0 / (dx * dx + dy * dy);
      const Dual wdx = w * dx, wdy = w * dy;
      a = a + wdx * dx;
      bb = bb + wdx * dy;
      cc = cc + wdy * dy;
      const Dual d0 = wn.r - wo.r, d1 = wn.u - wo.u, d2 = wn.v - wo.v, d3 = wn.p - wo.p;
      acc[0] = acc[0] + wdx * d0; acc[1] = acc[1] + wdy * d0;
      acc[2] = acc[2] + wdx * d1; acc[3] = acc[3] + wdy * d1;
      acc[4] = acc[4] + wdx * d2; acc[5] = acc[5] + wdy * d2;
      acc[6] = acc[6] + wdx * d3; acc[7] = acc[7] + wdy * d3;
    }
    const Dual det = a * cc - bb * bb;
    const Dual tr = a + cc;
    Dual out[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) out[j] = Dual(0.0);
    if (det > 1.0e-12 * (tr * tr)) {
      const Dual inv = 1.0 / det;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        out[2 * k] = (cc * acc[2 * k] - bb * acc[2 * k + 1]) * inv;
        out[2 * k + 1] = (a * acc[2 * k + 1] - bb * acc[2 * k]) * inv;
      }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      G[j * ld + c] = out[j].v;
      dG[j * ld + c] = out[j].d;
    }
    return;
  }
  for (int e = am.cf_off[c]; e < e1; ++e) {
    const int ent = am.cf_ent[e];
    const int f = ent >> 1;
    const int o = am.fo[f], n = am.fn[f];
    const P4T<Dual> wo = c2p(load_dual_u(U, dU, ld, o), g);
    Dual q0 = wo.r, q1 = wo.u, q2 = wo.v, q3 = wo.p;
    if (n >= 0) {
      const P4T<Dual> wn = c2p(load_dual_u(U, dU, ld, n), g);
      q0 = 0.5 * (wo.r + wn.r);
      q1 = 0.5 * (wo.u + wn.u);
      q2 = 0.5 * (wo.v + wn.v);
      q3 = 0.5 * (wo.p + wn.p);
    }
    const Dual A = dual_at(am.fA, gt.dA, f);
    const Dual x = dual_at(am.fnx, gt.dnx, f) * A;
    const Dual y = dual_at(am.fny, gt.dny, f) * A;
    if (ent & 1) {
      acc[0] = acc[0] - q0 * x; acc[1] = acc[1] - q0 * y;
      acc[2] = acc[2] - q1 * x; acc[3] = acc[3] - q1 * y;
      acc[4] = acc[4] - q2 * x; acc[5] = acc[5] - q2 * y;
      acc[6] = acc[6] - q3 * x; acc[7] = acc[7] - q3 * y;
    } else {
      acc[0] = acc[0] + q0 * x; acc[1] = acc[1] + q0 * y;
      acc[2] = acc[2] + q1 * x; acc[3] = acc[3] + q1 * y;
      acc[4] = acc[4] + q2 * x; acc[5] = acc[5] + q2 * y;
      acc[6] = acc[6] + q3 * x; acc[7] = acc[7] + q3 * y;
    }
  }
  const Dual iv = 1.0 / dmax(dual_at(am.vol, gt.dvol, c), Dual(1.0e-12));
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const Dual v = acc[j] * iv;
    G[j * ld + c] = v.v;
    dG[j * ld + c] = v.d;
  }
}

// ---- tangent of the face fluxes ------------------------------------------------------------------
// fa[k][F] (k = 0..3 values F_k A, 4..7 their derivatives).
template <int SCHEME, int LIMITER, bool SECOND, bool FROZEN>
__global__ __launch_bounds__(kThreads) void k_tan_face(AdjMesh am, const double* __restrict__ U,
                                                       const double* __restrict__ dU, const double* __restrict__ G,
                                                       const double* __restrict__ dG, GeoTan gt, FaceOptsT<Dual> o,
                                                       double* __restrict__ fa) {
  const int f = blockIdx.x * kThreads + threadIdx.x;
  if (f >= am.F) return;
  const int ld = am.ld;
  const int co = am.fo[f], cn = am.fn[f];
  const Dual nx = dual_at(am.fnx, gt.dnx, f);
  const Dual ny = dual_at(am.fny, gt.dny, f);
  const Dual A = dual_at(am.fA, gt.dA, f);
  FluxT<Dual> Fl;
  if (cn >= 0) {
    Dual go[8], gn[8];
    Dual dox(0.0), doy(0.0), dnx(0.0), dny(0.0);
    if (SECOND) {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        go[j] = dual_at(G + j * ld, dG ? dG + j * ld : nullptr, co);
        gn[j] = dual_at(G + j * ld, dG ? dG + j * ld : nullptr, cn);
      }
      const Dual fx = dual_at(am.ffx, gt.dfx, f), fy = dual_at(am.ffy, gt.dfy, f);
      dox = fx - dual_at(am.ccx, gt.dccx, co);
      doy = fy - dual_at(am.ccy, gt.dccy, co);
      dnx = fx - dual_at(am.ccx, gt.dccx, cn);
      dny = fy - dual_at(am.ccy, gt.dccy, cn);
    }
    double fr[8];
    if (FROZEN) {
#pragma unroll
      for (int j = 0; j < 8; ++j) fr[j] = o.frozen[j * am.F + f];
    }
    bool reverted;
    Fl = interior_face<SCHEME, LIMITER, SECOND, FROZEN>(load_dual_u(U, dU, ld, co), load_dual_u(U, dU, ld, cn), go,
                                                        gn, nx, ny, dox, doy, dnx, dny, o, &reverted, fr);
  } else {
    Fl = boundary_face<SCHEME>(load_dual_u(U, dU, ld, co), am.fbc[f] == 2, nx, ny, o);
  }
  const Dual f0 = Fl.f0 * A, f1 = Fl.f1 * A, f2 = Fl.f2 * A, f3 = Fl.f3 * A;
  const int F = am.F;
  fa[0 * F + f] = f0.v;
  fa[1 * F + f] = f1.v;
  fa[2 * F + f] = f2.v;
  fa[3 * F + f] = f3.v;
  fa[4 * F + f] = f0.d;
  fa[5 * F + f] = f1.d;
  fa[6 * F + f] = f2.d;
  fa[7 * F + f] = f3.d;
}

// Per cell: R = sum_f +-F A in face order (residual_assembly.cpp:168-180), values and derivatives.
__global__ __launch_bounds__(kThreads) void k_tan_gather(AdjMesh am, const double* __restrict__ fa,
                                                         double* __restrict__ R, double* __restrict__ dR) {
  const int c = blockIdx.x * kThreads + threadIdx.x;
  if (c >= am.N) return;
  const int F = am.F, ld = am.ld;
  double r[4] = {0.0, 0.0, 0.0, 0.0}, d[4] = {0.0, 0.0, 0.0, 0.0};
  const int e1 = am.cf_off[c + 1];
  for (int e = am.cf_off[c]; e < e1; ++e) {
    const int ent = am.cf_ent[e];
    const int f = ent >> 1;
    if (ent & 1) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        r[k] -= fa[k * F + f];
        d[k] -= fa[(4 + k) * F + f];
      }
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        r[k] += fa[k * F + f];
        d[k] += fa[(4 + k) * F + f];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (R) R[k * ld + c] = r[k];
    dR[k * ld + c] = d[k];
  }
}

// ---- face Jacobians ----------------------------------------------------------------------------------
// Thread (f, dir): one column of d(F_f A_f) / d(U_o[0..3], U_n[0..3], a_o[0..3], a_n[0..3]), where a_o / a_n
// are the owner's and neighbour's reconstruction increments of (rho, u, v, p) (interior_face_incr: the flux
// depends on the gradients only through them).  Columns of the owner's variables go to the face's
// owner-side slot, the neighbour's to its neighbour-side slot (kJacCols layout).  Boundary faces depend on
// U_o only; their increment columns are zero.
template <int SCHEME, int LIMITER, bool SECOND, bool FROZEN>
__global__ __launch_bounds__(kThreads) void k_face_jacobian(AdjMesh am, const double* __restrict__ U,
                                                            const double* __restrict__ G, FaceOpts o0,
                                                            double* __restrict__ J) {
  const long long tid = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
  const int F = am.F;
  if (tid >= 16LL * F) return;
  const int f = static_cast<int>(tid >> 4);
  const int dir = static_cast<int>(tid & 15);
  const int side = (dir >> 2) & 1;  // 0..3 and 8..11: owner; 4..7 and 12..15: neighbour
  const int slot = am.f_slot[side * F + f];
  if (slot < 0) return;  // the missing neighbour side of a boundary face
  const int ld = am.ld;
  const int co = am.fo[f], cn = am.fn[f];
  FaceOptsT<Dual> o;
  o.ff.r = o0.ff.r; o.ff.ru = o0.ff.ru; o.ff.rv = o0.ff.rv; o.ff.E = o0.ff.E;
  o.g = o0.g;
  o.low_mach_fix = o0.low_mach_fix;
  o.low_mach_cutoff = o0.low_mach_cutoff;
  o.stab_floor = o0.stab_floor;
  o.char_bc = o0.char_bc;
  o.precond = o0.precond;
  o.mach_ref = o0.mach_ref;
  o.mach_min = o0.mach_min;
  o.beta_min = o0.beta_min;
  o.beta_max = o0.beta_max;
  o.as_fix = o0.as_fix;
  o.as_cut = o0.as_cut;
  o.as_fmin = o0.as_fmin;
  o.as_w = o0.as_w;
  o.frozen = o0.frozen;
  const double A = am.fA[f];
  const Dual nx(am.fnx[f]), ny(am.fny[f]);
  double col[4] = {0.0, 0.0, 0.0, 0.0};
  if (cn >= 0) {
    C4T<Dual> uo = load_dual_u(U, nullptr, ld, co), un = load_dual_u(U, nullptr, ld, cn);
    Dual ao[4], an[4];
    if (SECOND) {
      const double fx = am.ffx[f], fy = am.ffy[f];
      const double dox = fx - am.ccx[co], doy = fy - am.ccy[co];
      const double dnx = fx - am.ccx[cn], dny = fy - am.ccy[cn];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        ao[k] = Dual(recon_incr(G[(2 * k) * ld + co], G[(2 * k + 1) * ld + co], dox, doy));
        an[k] = Dual(recon_incr(G[(2 * k) * ld + cn], G[(2 * k + 1) * ld + cn], dnx, dny));
      }
    }
    bool active = true;
    if (dir < 4) {
      (dir == 0 ? uo.r : dir == 1 ? uo.ru : dir == 2 ? uo.rv : uo.E).d = 1.0;
    } else if (dir < 8) {
      (dir == 4 ? un.r : dir == 5 ? un.ru : dir == 6 ? un.rv : un.E).d = 1.0;
    } else if (dir < 12) {
      if (SECOND) ao[dir - 8].d = 1.0; else active = false;
    } else {
      if (SECOND) an[dir - 12].d = 1.0; else active = false;
    }
    if (active) {
      double fr[8];
      if (FROZEN) {
#pragma unroll
        for (int j = 0; j < 8; ++j) fr[j] = o0.frozen[j * F + f];
      }
      bool reverted;
      const FluxT<Dual> Fl =
          interior_face_incr<SCHEME, LIMITER, SECOND, FROZEN>(uo, un, ao, an, nx, ny, o, &reverted, fr);
      col[0] = Fl.f0.d * A;
      col[1] = Fl.f1.d * A;
      col[2] = Fl.f2.d * A;
      col[3] = Fl.f3.d * A;
    }
  } else if (dir < 4) {
    C4T<Dual> uo = load_dual_u(U, nullptr, ld, co);
    (dir == 0 ? uo.r : dir == 1 ? uo.ru : dir == 2 ? uo.rv : uo.E).d = 1.0;
    const FluxT<Dual> Fl = boundary_face<SCHEME>(uo, am.fbc[f] == 2, nx, ny, o);
    col[0] = Fl.f0.d * A;
    col[1] = Fl.f1.d * A;
    col[2] = Fl.f2.d * A;
    col[3] = Fl.f3.d * A;
  }
  const int q0 = (dir >= 8 ? 16 : 0) + (dir & 3) * 4;
  double* t = J + static_cast<size_t>(slot / kEntryTile) * (kJacCols * kEntryTile) + (slot % kEntryTile);
#pragma unroll
  for (int k = 0; k < 4; ++k) t[(q0 + k) * kEntryTile] = col[k];
}

// Factored face Jacobians (the frozen limiter, or first order).  There the flux of face f depends on a
// side's conserved state U and increments a only through the side's face state W = floors(c2p(U) + r a),
// or c2p(U) where the face reverts to first order (r: the frozen ratios; no increments at first order).
// So d(F A)/dU = (d(F A)/dW) M C and d(F A)/da = (d(F A)/dW) M diag(r) with C = dc2p/dU of the cell, the
// same for all its faces, and M the 0/1 mask of the rho and p floors of the reconstruction (identity on a
// reverted face).  Thread (f, dir), dir < 8: side dir >> 2, primitive var dir & 3; it stores column
// var * 4 + k = d(F_k A)/dW_var times M_var, and column 16 + var = r_eff (r of this side and variable,
// 0 on a reverted face, at first order and on boundary faces).  The gather multiplies by r_eff and, once
// per cell, by C^T (k_adj_grad / adj_update): 20 of the 32 columns are read, 160 B per face side.
template <int SCHEME, bool SECOND>
__global__ __launch_bounds__(kThreads) void k_face_jacobian_w(AdjMesh am, const double* __restrict__ U,
                                                              const double* __restrict__ G, FaceOpts o0,
                                                              double* __restrict__ J) {
  const long long tid = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
  const int F = am.F;
  if (tid >= 8LL * F) return;
  const int f = static_cast<int>(tid >> 3);
  const int dir = static_cast<int>(tid & 7);
  const int side = dir >> 2;
  const int var = dir & 3;
  const int slot = am.f_slot[side * F + f];
  if (slot < 0) return;  // the missing neighbour side of a boundary face
  const int ld = am.ld;
  const double g = o0.g;
  const int co = am.fo[f], cn = am.fn[f];
  FaceOptsT<Dual> o;
  o.ff.r = o0.ff.r; o.ff.ru = o0.ff.ru; o.ff.rv = o0.ff.rv; o.ff.E = o0.ff.E;
  o.g = g;
  o.low_mach_fix = o0.low_mach_fix;
  o.low_mach_cutoff = o0.low_mach_cutoff;
  o.stab_floor = o0.stab_floor;
  o.char_bc = o0.char_bc;
  o.precond = o0.precond;
  o.mach_ref = o0.mach_ref;
  o.mach_min = o0.mach_min;
  o.beta_min = o0.beta_min;
  o.beta_max = o0.beta_max;
  o.as_fix = o0.as_fix;
  o.as_cut = o0.as_cut;
  o.as_fmin = o0.as_fmin;
  o.as_w = o0.as_w;
  o.frozen = o0.frozen;
  const double A = am.fA[f];
  const Dual nx(am.fnx[f]), ny(am.fny[f]);
  double col[4] = {0.0, 0.0, 0.0, 0.0};
  double reff = 0.0;
  const auto seeded = [&](const P4& w, bool seed) {  // w as duals, d/dW_var = 1 where seeded
    P4T<Dual> d;
    d.r = Dual(w.r, seed && var == 0 ? 1.0 : 0.0);
    d.u = Dual(w.u, seed && var == 1 ? 1.0 : 0.0);
    d.v = Dual(w.v, seed && var == 2 ? 1.0 : 0.0);
    d.p = Dual(w.p, seed && var == 3 ? 1.0 : 0.0);
    return d;
  };
  if (cn >= 0) {
    // the face states of interior_face_src on doubles: reconstruction, floors, validity fallback
    P4 L = c2p(load_c4(U, ld, co), g), Rt = c2p(load_c4(U, ld, cn), g);
    bool mask = true;  // this side's variable passes its floor (always for u, v and on a reverted face)
    if (SECOND) {
      const double fx = am.ffx[f], fy = am.ffy[f];
      const double dox = fx - am.ccx[co], doy = fy - am.ccy[co];
      const double dnx = fx - am.ccx[cn], dny = fy - am.ccy[cn];
      const double* fr = o0.frozen;
      P4 l2, r2;
      double* pl[4] = {&l2.r, &l2.u, &l2.v, &l2.p};
      double* pr[4] = {&r2.r, &r2.u, &r2.v, &r2.p};
      const double ql0[4] = {L.r, L.u, L.v, L.p}, qr0[4] = {Rt.r, Rt.u, Rt.v, Rt.p};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double ao = recon_incr(G[(2 * k) * ld + co], G[(2 * k + 1) * ld + co], dox, doy);
        const double an = recon_incr(G[(2 * k) * ld + cn], G[(2 * k + 1) * ld + cn], dnx, dny);
        *pl[k] = ql0[k] + fr[k * F + f] * ao;
        *pr[k] = qr0[k] + fr[(4 + k) * F + f] * an;
      }
      const P4& q = side ? r2 : l2;
      mask = !((var == 0 && q.r < kFloor) || (var == 3 && q.p < kFloor));
      l2.r = dmax(l2.r, kFloor);
      r2.r = dmax(r2.r, kFloor);
      l2.p = dmax(l2.p, kFloor);
      r2.p = dmax(r2.p, kFloor);
      if (valid_prim(l2) && valid_prim(r2)) {
        L = l2;
        Rt = r2;
        reff = fr[(4 * side + var) * F + f];
      } else {
        mask = true;
      }
    }
    if (mask) {
      const FluxT<Dual> Fl =
          flux_of_states<SCHEME, false>(seeded(L, side == 0), seeded(Rt, side == 1), nx, ny, o);
      col[0] = Fl.f0.d * A;
      col[1] = Fl.f1.d * A;
      col[2] = Fl.f2.d * A;
      col[3] = Fl.f3.d * A;
    }
  } else {
    const FluxT<Dual> Fl = boundary_face_prim<SCHEME>(seeded(c2p(load_c4(U, ld, co), g), true), am.fbc[f] == 2, nx,
                                                      ny, o);
    col[0] = Fl.f0.d * A;
    col[1] = Fl.f1.d * A;
    col[2] = Fl.f2.d * A;
    col[3] = Fl.f3.d * A;
  }
  double* t = J + static_cast<size_t>(slot / kEntryTile) * (kJacCols * kEntryTile) + (slot % kEntryTile);
#pragma unroll
  for (int k = 0; k < 4; ++k) t[(var * 4 + k) * kEntryTile] = col[k];
  t[(16 + var) * kEntryTile] = reff;
}

// ---- adjoint product out = R_U^T psi ---------------------------------------------------------------
// Face f adds F_f A_f to R_o and subtracts it from R_n, so its weight is w = psi_o - psi_n (psi_o on
// boundary faces).  Per cell, over its face entries (increasing face index), its side of J_f^T w: the
// conserved-state columns give the direct part of R_U^T psi, the increment columns b_d = (dF/da_d)^T w
// give the gradient adjoint Gbar[d] += b_d (face centre - cell centre).
//
// The per-cell kernels come in two launch shapes with the same sums.  One workgroup covers one entry
// group of kEntryTile cells, either with one thread per cell (serial over the cell's entries), or, WIDE,
// with kWide threads per cell (kWideThreads per workgroup: entry lane `sub` handles entries sub,
// sub + kWide, ...) so that a launch of few groups still fills the chip.  wide_accumulate adds the
// entries' values into LDS one entry lane at a time, in increasing entry order, so every per-cell sum is
// that of the serial loop, bit for bit.  The launchers take WIDE below kWideBelow groups.
constexpr int kWide = 4;
constexpr int kWideThreads = kWide * kEntryTile;
constexpr int kWideBelow = 1024;

__device__ __forceinline__ void adj_entry_weight(const double* __restrict__ psi, int ld, const double* pc, int nb,
                                                 double* w) {
  const int other = nb >> 1;  // -1 on boundary faces
  if (nb & 1) {               // this cell is the face's neighbour: w = psi_other - psi_c
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = psi[k * ld + other] - pc[k];
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = pc[k] - (other >= 0 ? psi[k * ld + other] : 0.0);
  }
}

// The thread's cell and entry lane in entry group `grp` (sub = 0 in the one-thread-per-cell shape).
struct WideCell {
  int lane, sub, c, ne, t0, tiles;
  bool live;
};

template <bool WIDE>
__device__ __forceinline__ WideCell wide_cell(const AdjMesh& am, int grp) {
  WideCell w;
  w.lane = WIDE ? threadIdx.x % kEntryTile : threadIdx.x;
  w.sub = WIDE ? threadIdx.x / kEntryTile : 0;
  w.c = grp * kEntryTile + w.lane;
  w.live = w.c < am.N;
  w.ne = w.live ? am.cf_off[w.c + 1] - am.cf_off[w.c] : 0;
  w.t0 = am.e_tile[grp];
  w.tiles = am.e_tile[grp + 1] - w.t0;  // uniform over the workgroup
  return w;
}

// acc[v * kEntryTile + lane] = sum over the cell's entries i (in increasing order) of the values that
// entry(i, val) reports (entry returns false for an entry that adds nothing).  All threads of the
// workgroup must call it; acc is ready for every thread on return.
template <int NV, class Fn>
__device__ __forceinline__ void wide_accumulate(double* acc, const WideCell& w, Fn&& entry) {
  for (int v = threadIdx.x; v < NV * kEntryTile; v += kWideThreads) acc[v] = 0.0;
  __syncthreads();
  for (int r = 0; r < w.tiles; r += kWide) {
    const int i = r + w.sub;
    double val[NV];
    bool add = false;
    if (i < w.ne) add = entry(i, val);
    for (int j = 0; j < kWide; ++j) {
      if (w.sub == j && add) {
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v * kEntryTile + w.lane] += val[v];
      }
      __syncthreads();
    }
  }
}

// One face entry of the gather for NS solves sharing the face Jacobians: val[12 s + d] (d < 4) the direct
// part, val[12 s + 4 + j] the gradient adjoint of solve s.  FACT (factored Jacobians, k_face_jacobian_w):
// the direct part is in the primitive variables (C^T is applied per cell afterwards), and the increment
// columns are the direct ones times r_eff.
template <int NS, bool FACT>
__device__ __forceinline__ bool adj_gather_entry(const AdjMesh& am, const double* __restrict__ J,
                                                 const double* const* psi, const double (*pc)[4], const WideCell& w,
                                                 int i, double* val) {
  const int s = (w.t0 + i) * kEntryTile + w.lane;
  const int nb = am.e_nb[s];
  double wt[NS][4];
#pragma unroll
  for (int q = 0; q < NS; ++q) adj_entry_weight(psi[q], am.ld, pc[q], nb, wt[q]);
  const double* Jt = J + static_cast<size_t>(w.t0 + i) * (kJacCols * kEntryTile) + w.lane;
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    double t[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) t[q] = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double j = Jt[(d * 4 + k) * kEntryTile];
#pragma unroll
      for (int q = 0; q < NS; ++q) t[q] += j * wt[q][k];
    }
#pragma unroll
    for (int q = 0; q < NS; ++q) val[12 * q + d] = t[q];
  }
  const double dx = am.e_dx[s], dy = am.e_dy[s];
  if (FACT) {
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      const double r = Jt[(16 + d) * kEntryTile];
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        const double t = r * val[12 * q + d];
        val[12 * q + 4 + 2 * d] = t * dx;
        val[12 * q + 4 + 2 * d + 1] = t * dy;
      }
    }
    return true;
  }
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    double t[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) t[q] = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double j = Jt[(16 + d * 4 + k) * kEntryTile];
#pragma unroll
      for (int q = 0; q < NS; ++q) t[q] += j * wt[q][k];
    }
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      val[12 * q + 4 + 2 * d] = t[q] * dx;
      val[12 * q + 4 + 2 * d + 1] = t[q] * dy;
    }
  }
  return true;
}

// The gather of entry group grp for NS solves: direct[q] = the direct part of R_U^T psi[q], Gbar[q] its
// gradient adjoint.  acc (WIDE): 12 NS kEntryTile doubles of LDS.
template <bool WIDE, int NS, bool FACT>
__device__ __forceinline__ void adj_gather_group(const AdjMesh& am, const double* __restrict__ J,
                                                 const double* const* psi, double* const* direct,
                                                 double* const* Gbar, int grp, double* acc) {
  const WideCell w = wide_cell<WIDE>(am, grp);
  const int ld = am.ld;
  double pc[NS][4];
#pragma unroll
  for (int q = 0; q < NS; ++q)
#pragma unroll
    for (int k = 0; k < 4; ++k) pc[q][k] = w.live ? psi[q][k * ld + w.c] : 0.0;
  if (WIDE) {
    wide_accumulate<12 * NS>(acc, w, [&](int i, double* val) {
      return adj_gather_entry<NS, FACT>(am, J, psi, pc, w, i, val);
    });
    if (!w.live) return;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
#pragma unroll
      for (int j = 0; j < 12; ++j) {
        const int v = 12 * q + j;
        if (v % kWide != w.sub) continue;  // the entry lanes share the stores
        const double x = acc[v * kEntryTile + w.lane];
        if (j < 4)
          direct[q][j * ld + w.c] = x;
        else
          Gbar[q][(j - 4) * ld + w.c] = x;
      }
    }
  } else {
    if (!w.live) return;
    double sum[12 * NS];
#pragma unroll
    for (int v = 0; v < 12 * NS; ++v) sum[v] = 0.0;
    for (int i = 0; i < w.ne; ++i) {
      double val[12 * NS];
      adj_gather_entry<NS, FACT>(am, J, psi, pc, w, i, val);
#pragma unroll
      for (int v = 0; v < 12 * NS; ++v) sum[v] += val[v];
    }
#pragma unroll
    for (int q = 0; q < NS; ++q) {
#pragma unroll
      for (int j = 0; j < 4; ++j) direct[q][j * ld + w.c] = sum[12 * q + j];
#pragma unroll
      for (int j = 0; j < 8; ++j) Gbar[q][j * ld + w.c] = sum[12 * q + 4 + j];
    }
  }
}

template <bool WIDE, bool FACT>
__global__ __launch_bounds__(WIDE ? kWideThreads : kEntryTile) void k_adj_gather(AdjMesh am,
                                                                                 const double* __restrict__ J,
                                                                                 const double* __restrict__ psi,
                                                                                 double* __restrict__ out,
                                                                                 double* __restrict__ Gbar) {
  __shared__ double acc[WIDE ? 12 * kEntryTile : 1];
  const double* p[1] = {psi};
  double* o[1] = {out};
  double* g[1] = {Gbar};
  adj_gather_group<WIDE, 1, FACT>(am, J, p, o, g, blockIdx.x, acc);
}

// The gradient-adjoint term of one face entry of cell c, val[k] for primitive k (returns false where the
// face adds nothing):
//   Green-Gauss: G_c[k] = iv_c sum_f +-q_f nA_f, q_f = (W_o + W_n) / 2 (W_o on boundary faces), so both
//     cells of an interior face get (1/2) nA . (iv_o Gbar_o[k] - iv_n Gbar_n[k]); a boundary face gives
//     its owner nA . iv_o Gbar_o[k].
//   Least squares: G_c[k] = H_c sum_f (w d)_f (W_n - W_o)_k, so s = (w d)_f . (H_o Gbar_o[k] +
//     H_n Gbar_n[k]) goes to the neighbour with + and to the owner with -.
template <int GK>
__device__ __forceinline__ bool adj_wbar_entry(const AdjMesh& am, const double* __restrict__ Gbar, int c, int s,
                                               double* val) {
  const int ld = am.ld;
  const int nb = am.e_nb[s];
  const int side = nb & 1;
  const int other = nb >> 1;
  const int fo = side ? other : c;
  const int fn = side ? c : other;
  if (GK == 1) {
    if (fn < 0) return false;
    const double wx = am.e_gx[s], wy = am.e_gy[s];
    const double hxo = am.H[fo], hyo = am.H[ld + fo], hzo = am.H[2 * ld + fo];
    const double hxn = am.H[fn], hyn = am.H[ld + fn], hzn = am.H[2 * ld + fn];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double gxo = Gbar[(2 * k) * ld + fo], gyo = Gbar[(2 * k + 1) * ld + fo];
      const double gxn = Gbar[(2 * k) * ld + fn], gyn = Gbar[(2 * k + 1) * ld + fn];
      const double ax = hxo * gxo + hyo * gyo + hxn * gxn + hyn * gyn;
      const double ay = hyo * gxo + hzo * gyo + hyn * gxn + hzn * gyn;
      const double sv = wx * ax + wy * ay;
      val[k] = side ? sv : -sv;
    }
  } else {
    const double nax = am.e_gx[s], nay = am.e_gy[s];
    const double ivo = 1.0 / dmax(am.vol[fo], 1.0e-12);
    if (fn >= 0) {
      const double ivn = 1.0 / dmax(am.vol[fn], 1.0e-12);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double tx = ivo * Gbar[(2 * k) * ld + fo] - ivn * Gbar[(2 * k) * ld + fn];
        const double ty = ivo * Gbar[(2 * k + 1) * ld + fo] - ivn * Gbar[(2 * k + 1) * ld + fn];
        val[k] = 0.5 * (nax * tx + nay * ty);
      }
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        val[k] = nax * (ivo * Gbar[(2 * k) * ld + fo]) + nay * (ivo * Gbar[(2 * k + 1) * ld + fo]);
    }
  }
  return true;
}

// Wbar of cell c from the gradient adjoint of the cell and its face neighbours (one thread per cell).
template <int GK>
__device__ __forceinline__ void adj_wbar(const AdjMesh& am, const double* __restrict__ Gbar, int c, double* wb) {
  const int lane = c % kEntryTile;
  const int t0 = am.e_tile[c / kEntryTile];
  const int ne = am.cf_off[c + 1] - am.cf_off[c];
#pragma unroll
  for (int k = 0; k < 4; ++k) wb[k] = 0.0;
  for (int i = 0; i < ne; ++i) {
    double val[4];
    if (adj_wbar_entry<GK>(am, Gbar, c, (t0 + i) * kEntryTile + lane, val)) {
#pragma unroll
      for (int k = 0; k < 4; ++k) wb[k] += val[k];
    }
  }
}

// Per cell: Wbar (adj_wbar), then out += (dW/dU)^T Wbar; FACT: out holds the gather's primitive-variable
// part s, and out = (dW/dU)^T (Wbar + s).
template <int GK, bool FACT>
__global__ __launch_bounds__(kThreads) void k_adj_grad(AdjMesh am, const double* __restrict__ U, double g,
                                                       const double* __restrict__ Gbar, double* __restrict__ out) {
  const int c = blockIdx.x * kThreads + threadIdx.x;
  if (c >= am.N) return;
  const int ld = am.ld;
  double wb[4];
  adj_wbar<GK>(am, Gbar, c, wb);
  if (FACT) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      wb[k] += out[k * ld + c];
      out[k * ld + c] = 0.0;
    }
  }
  // (dW/dU)^T Wbar by four dual evaluations of the primitive map (euler_flux.cpp:115-127).
  const C4 u0 = load_c4(U, ld, c);
  double add[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    C4T<Dual> s;
    s.r = Dual(u0.r, j == 0 ? 1.0 : 0.0);
    s.ru = Dual(u0.ru, j == 1 ? 1.0 : 0.0);
    s.rv = Dual(u0.rv, j == 2 ? 1.0 : 0.0);
    s.E = Dual(u0.E, j == 3 ? 1.0 : 0.0);
    const P4T<Dual> w = c2p(s, g);
    add[j] = w.r.d * wb[0] + w.u.d * wb[1] + w.v.d * wb[2] + w.p.d * wb[3];
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) out[j * ld + c] += add[j];
}

// ---- objective gradient ----------------------------------------------------------------------------
// dJ/dU of J = sum over wall faces of coef_f * p(owner) (the gauge pressure forces of forces.cpp:208-294
// are linear in the owner pressures).  One thread walks the wall patch in order, so cells with several
// wall faces sum them deterministically.  `out` is zeroed by the caller.
__global__ void k_objective_gradient(AdjMesh am, const double* __restrict__ U, double g,
                                     const double* __restrict__ coef, double* __restrict__ out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const int ld = am.ld;
  for (int i = 0; i < am.wall_count; ++i) {
    const int f = am.wall_start + i;
    const int o = am.fo[f];
    const C4 u0 = load_c4(U, ld, o);
    for (int j = 0; j < 4; ++j) {
      C4T<Dual> s;
      s.r = Dual(u0.r, j == 0 ? 1.0 : 0.0);
      s.ru = Dual(u0.ru, j == 1 ? 1.0 : 0.0);
      s.rv = Dual(u0.rv, j == 2 ? 1.0 : 0.0);
      s.E = Dual(u0.E, j == 3 ? 1.0 : 0.0);
      const P4T<Dual> w = c2p(s, g);
      out[j * ld + o] += coef[i] * w.p.d;
    }
  }
}

// ---- pseudo-time iteration of the linearised problem --------------------------------------------
// The primal stage is U* = U_s - D P R(U_s), D = cfl / max(lambda, 1e-8) per cell and P the low-Mach
// energy-row preconditioner (preconditioning.cpp:186-199).  Linearised at a fixed state it is
// dU* = dU_s - D P (R_U dU_s + R_p); the adjoint iterates the transposed map,
// phi* = phi_s - (R_U^T P^T D phi_s - rhs), whose fixed point gives psi = P^T D phi with
// R_U^T psi = rhs.  Both use the RK stage combinations of time_integrator.cpp:61-104, so they converge
// at the rate of the linearised primal.
__device__ __forceinline__ void precond_row(const C4& us, const FaceOpts& o, double* b2m1_inv_u, double* b2m1_inv_v,
                                            double* b2) {
  const P4 wp = c2p(us, o.g);
  const double bt = cell_beta(wp, o);
  *b2 = bt * bt;
  const double rho = dmax(wp.r, 1.0e-8);
  const double inv = 1.0 / rho;
  *b2m1_inv_u = (*b2 - 1.0) * inv * wp.u;
  *b2m1_inv_v = (*b2 - 1.0) * inv * wp.v;
}

// d(rho, u, v, p)/dU_j of the primitive map at u0 (euler_flux.cpp:115-127), by four dual evaluations:
// jac[j] = (drho, du, dv, dp) / dU_j.
__device__ __forceinline__ void c2p_jacobian(const C4& u0, double g, double (*jac)[4]) {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    C4T<Dual> sdu;
    sdu.r = Dual(u0.r, j == 0 ? 1.0 : 0.0);
    sdu.ru = Dual(u0.ru, j == 1 ? 1.0 : 0.0);
    sdu.rv = Dual(u0.rv, j == 2 ? 1.0 : 0.0);
    sdu.E = Dual(u0.E, j == 3 ? 1.0 : 0.0);
    const P4T<Dual> w = c2p(sdu, g);
    jac[j][0] = w.r.d;
    jac[j][1] = w.u.d;
    jac[j][2] = w.v.d;
    jac[j][3] = w.p.d;
  }
}

// r = direct + (dW/dU)^T Wbar, the update x* = x_s - (r - rhs) with the RK stage combination, and
// psi = P^T D x* for the next stage; returns the squared residual |r - rhs|^2 of the cell.
// FACT: direct holds the gather's primitive-variable part s, and r = (dW/dU)^T (Wbar + s).
template <bool FACT>
__device__ __forceinline__ double adj_update(int ld, const C4& u0, const FaceOpts& o, const double (*jac)[4],
                                             const double* wb, const double* __restrict__ direct, int stage, int rk3,
                                             const double* __restrict__ x0, const double* __restrict__ xs,
                                             const double* __restrict__ rhs, const double* __restrict__ lam,
                                             double cfl, double* __restrict__ xout, double* __restrict__ psi, int c) {
  double sq = 0.0;
  double r[4];
  if (FACT) {
    double v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = wb[k] + direct[k * ld + c];
#pragma unroll
    for (int j = 0; j < 4; ++j) r[j] = jac[j][0] * v[0] + jac[j][1] * v[1] + jac[j][2] * v[2] + jac[j][3] * v[3];
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const double add = jac[j][0] * wb[0] + jac[j][1] * wb[1] + jac[j][2] * wb[2] + jac[j][3] * wb[3];
      r[j] = direct[j * ld + c] + add;
    }
  }
  double x[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double res = r[k] - (rhs ? rhs[k * ld + c] : 0.0);
    double v = xs[k * ld + c] - res;
    if (rk3 && stage == 1) v = 0.75 * x0[k * ld + c] + 0.25 * v;
    if (rk3 && stage == 2) v = (1.0 / 3.0) * x0[k * ld + c] + (2.0 / 3.0) * v;
    x[k] = v;
    xout[k * ld + c] = v;
    sq += res * res;
  }
  const double dt = cfl / dmax(lam[c], 1.0e-8);
  double y[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) y[k] = dt * x[k];
  if (o.precond) {
    double au, av, b2;
    precond_row(u0, o, &au, &av, &b2);
    psi[c] = y[0];
    psi[ld + c] = y[1] + au * y[3];
    psi[2 * ld + c] = y[2] + av * y[3];
    psi[3 * ld + c] = b2 * y[3];
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) psi[k * ld + c] = y[k];
  }
  return sq;
}

// Fixed-order sum of one value per cell of the workgroup's entry group (threads below kEntryTile) into
// *dst (thread 0 writes; dst may be null): the tree of a kEntryTile-thread workgroup.  All threads call it.
__device__ __forceinline__ void group_sum_store(double v, double* red, double* dst) {
  if (threadIdx.x < kEntryTile) red[threadIdx.x] = v;
  __syncthreads();
  for (int w = kEntryTile / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0 && dst) *dst = red[0];
}

// Wbar of NS solves for the thread's cell, wb[q] (set where w.sub == 0 and the cell is live).  All threads
// of the workgroup call it.  acc (WIDE): 4 NS kEntryTile doubles of LDS.
template <bool WIDE, int GK, int NS>
__device__ __forceinline__ void adj_wbar_cells(const AdjMesh& am, const double* const* Gbar, const WideCell& w,
                                               double* acc, double (*wb)[4]) {
  if (WIDE) {
    wide_accumulate<4 * NS>(acc, w, [&](int i, double* val) {
      const int s = (w.t0 + i) * kEntryTile + w.lane;
      bool any = false;  // the same for every solve: it depends on the face only
#pragma unroll
      for (int q = 0; q < NS; ++q) any = adj_wbar_entry<GK>(am, Gbar[q], w.c, s, val + 4 * q) || any;
      return any;
    });
    if (w.sub == 0 && w.live) {
#pragma unroll
      for (int q = 0; q < NS; ++q)
#pragma unroll
        for (int k = 0; k < 4; ++k) wb[q][k] = acc[(4 * q + k) * kEntryTile + w.lane];
    }
  } else {
#pragma unroll
    for (int q = 0; q < NS; ++q)
#pragma unroll
      for (int k = 0; k < 4; ++k) wb[q][k] = 0.0;
    if (!w.live) return;
    for (int i = 0; i < w.ne; ++i) {
      const int s = (w.t0 + i) * kEntryTile + w.lane;
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        double val[4];
        if (adj_wbar_entry<GK>(am, Gbar[q], w.c, s, val)) {
#pragma unroll
          for (int k = 0; k < 4; ++k) wb[q][k] += val[k];
        }
      }
    }
  }
}

// One stage of the transposed pseudo-time iteration, per cell, after k_adj_gather: the gradient-adjoint
// gather and (dW/dU)^T of k_adj_grad (same arithmetic), r = R_U^T psi, the update
// x* = x_s - (r - rhs) with the RK stage combination, the squared norm of r - rhs into the block partial,
// and psi = P^T D x* for the next stage (k_adj_scale's arithmetic).
template <bool WIDE, int GK, bool FACT>
__global__ __launch_bounds__(WIDE ? kWideThreads : kEntryTile) void k_adj_grad_update(
    AdjMesh am, const double* __restrict__ U, FaceOpts o, const double* __restrict__ Gbar,
    const double* __restrict__ direct, int stage, int rk3, const double* __restrict__ x0,
    const double* __restrict__ xs, const double* __restrict__ rhs, const double* __restrict__ lam, double cfl,
    double* __restrict__ xout, double* __restrict__ partials, double* __restrict__ psi) {
  __shared__ double acc[WIDE ? 4 * kEntryTile : 1];
  __shared__ double red[kEntryTile];
  const WideCell w = wide_cell<WIDE>(am, blockIdx.x);
  const double* g[1] = {Gbar};
  double wb[1][4];
  adj_wbar_cells<WIDE, GK, 1>(am, g, w, acc, wb);
  double sq = 0.0;
  if (w.sub == 0 && w.live) {
    const C4 u0 = load_c4(U, am.ld, w.c);
    double jac[4][4];
    c2p_jacobian(u0, o.g, jac);
    sq = adj_update<FACT>(am.ld, u0, o, jac, wb[0], direct, stage, rk3, x0, xs, rhs, lam, cfl, xout, psi, w.c);
  }
  group_sum_store(sq, red, partials ? partials + blockIdx.x : nullptr);
}

// ---- batched adjoint solves: one workgroup per (solve a, solve b or -1, entry group) -----------------
// Two solves of one linearisation (same face Jacobians, state, options and CFL; e.g. the CL and CD
// adjoints of a case) share a work item: one pass over the face Jacobians and one primitive-map Jacobian
// per cell serve both.  Each solve runs the same per-cell arithmetic as a single solve, so every solve
// of a batch reproduces its single solve bit for bit; solves that have stopped are neither computed nor
// written.
// FM: the solves' face-Jacobian form, 0 all unfactored, 1 all factored, 2 each solve's own (AdjBatchSolve
// fact, taken per work item; both code paths then share the kernel's registers).
template <bool WIDE, int FM>
__global__ __launch_bounds__(WIDE ? kWideThreads : kEntryTile) void k_adj_gather_batch(
    const AdjBatchSolve* __restrict__ solves, const AdjBatchCtl* __restrict__ ctl, const int4* __restrict__ work) {
  __shared__ double acc[WIDE ? 24 * kEntryTile : 1];
  const int4 wk = work[blockIdx.x];
  if (wk.x < 0) return;
  const bool aa = ctl[wk.x].active != 0;
  const bool ab = wk.y >= 0 && ctl[wk.y].active != 0;
  if (!aa && !ab) return;  // uniform over the workgroup
  const AdjBatchSolve& sa = solves[aa ? wk.x : wk.y];
  if (aa && ab) {  // paired solves share one linearisation, so one Jacobian form
    const AdjBatchSolve& sb = solves[wk.y];
    const double* p[2] = {sa.psi, sb.psi};
    double* o[2] = {sa.direct, sb.direct};
    double* g[2] = {sa.Gbar, sb.Gbar};
    if (FM == 1 || (FM == 2 && sa.fact))
      adj_gather_group<WIDE, 2, FM != 0>(sa.am, sa.J, p, o, g, wk.z, acc);
    else
      adj_gather_group<WIDE, 2, false>(sa.am, sa.J, p, o, g, wk.z, acc);
  } else {
    const double* p[1] = {sa.psi};
    double* o[1] = {sa.direct};
    double* g[1] = {sa.Gbar};
    if (FM == 1 || (FM == 2 && sa.fact))
      adj_gather_group<WIDE, 1, FM != 0>(sa.am, sa.J, p, o, g, wk.z, acc);
    else
      adj_gather_group<WIDE, 1, false>(sa.am, sa.J, p, o, g, wk.z, acc);
  }
}

template <bool WIDE, int GK, int FM>
__global__ __launch_bounds__(WIDE ? kWideThreads : kEntryTile) void k_adj_grad_update_batch(
    const AdjBatchSolve* __restrict__ solves, const AdjBatchCtl* __restrict__ ctl, const int4* __restrict__ work,
    int stage, int rk3, int ix0, int ixs, int ixo, int want_partials) {
  __shared__ double acc[WIDE ? 8 * kEntryTile : 1];
  __shared__ double red[kEntryTile];
  const int4 wk = work[blockIdx.x];
  if (wk.x < 0) return;
  const bool aa = ctl[wk.x].active != 0;
  const bool ab = wk.y >= 0 && ctl[wk.y].active != 0;
  if (!aa && !ab) return;  // uniform over the workgroup
  const AdjBatchSolve& s0 = solves[aa ? wk.x : wk.y];
  const AdjMesh& am = s0.am;
  const WideCell w = wide_cell<WIDE>(am, wk.z);
  // wb[0]: the first running solve's Wbar, wb[1] the second's when both run
  double wb[2][4];
  if (aa && ab) {
    const double* g[2] = {solves[wk.x].Gbar, solves[wk.y].Gbar};
    adj_wbar_cells<WIDE, GK, 2>(am, g, w, acc, wb);
  } else {
    const double* g[1] = {s0.Gbar};
    adj_wbar_cells<WIDE, GK, 1>(am, g, w, acc, wb);
  }
  double sqa = 0.0, sqb = 0.0;
  if (w.sub == 0 && w.live) {
    const int c = w.c;
    const C4 u0 = load_c4(s0.U, am.ld, c);
    double jac[4][4];
    c2p_jacobian(u0, s0.o.g, jac);
    if (aa) {
      const AdjBatchSolve& s = solves[wk.x];
      sqa = (FM == 1 || (FM == 2 && s.fact))
                ? adj_update<FM != 0>(am.ld, u0, s.o, jac, wb[0], s.direct, stage, rk3, s.x[ix0], s.x[ixs], s.rhs,
                                      s.lam, s.cfl, s.x[ixo], s.psi, c)
                   : adj_update<false>(am.ld, u0, s.o, jac, wb[0], s.direct, stage, rk3, s.x[ix0], s.x[ixs], s.rhs,
                                       s.lam, s.cfl, s.x[ixo], s.psi, c);
    }
    if (ab) {
      const AdjBatchSolve& s = solves[wk.y];
      const double* wq = wb[aa ? 1 : 0];
      sqb = (FM == 1 || (FM == 2 && s.fact))
                ? adj_update<FM != 0>(am.ld, u0, s.o, jac, wq, s.direct, stage, rk3, s.x[ix0], s.x[ixs], s.rhs,
                                      s.lam, s.cfl, s.x[ixo], s.psi, c)
                   : adj_update<false>(am.ld, u0, s.o, jac, wq, s.direct, stage, rk3, s.x[ix0], s.x[ixs], s.rhs,
                                       s.lam, s.cfl, s.x[ixo], s.psi, c);
    }
  }
  if (want_partials) {
    if (aa) group_sum_store(sqa, red, solves[wk.x].partials + wk.z);
    if (aa && ab) __syncthreads();
    if (ab) group_sum_store(sqb, red, solves[wk.y].partials + wk.z);
  }
}

// Per solve (one workgroup each): the squared norm of the iteration's stage-0 linear residual from the
// block partials (the fixed order of k_partials_sum), the iteration count, and the stopping tests:
// residual <= tol * first residual, a non-finite residual, or max_iters.
__global__ __launch_bounds__(kThreads) void k_adj_batch_control(const AdjBatchSolve* __restrict__ solves,
                                                                AdjBatchCtl* __restrict__ ctl, double tol,
                                                                int max_iters) {
  __shared__ double red[kThreads];
  AdjBatchCtl& k = ctl[blockIdx.x];
  if (!k.active) return;
  const AdjBatchSolve& s = solves[blockIdx.x];
  double v = 0.0;
  for (int i = threadIdx.x; i < s.nparts; i += kThreads) v += s.partials[i];
  double sum = 0.0;
  static_assert(kThreads == kEntryTile, "the control sum uses the entry-group tree");
  group_sum_store(v, red, &sum);
  if (threadIdx.x != 0) return;
  const double res = sqrt(sum);
  k.iter += 1;
  if (k.iter == 1) k.res0 = res;
  k.res = res;
  if (!(res == res) || res > 1.7976931348623157e308) {
    k.failed = 1;
    k.active = 0;
  } else if (res <= tol * k.res0) {
    k.converged = 1;
    k.active = 0;
  } else if (k.iter >= max_iters) {
    k.active = 0;
  }
}

__global__ __launch_bounds__(kThreads) void k_adj_scale(AdjMesh am, const double* __restrict__ lam, double cfl,
                                                        const double* __restrict__ U, FaceOpts o,
                                                        const double* __restrict__ phi, double* __restrict__ psi) {
  const int c = blockIdx.x * kThreads + threadIdx.x;
  if (c >= am.N) return;
  const int ld = am.ld;
  const double dt = cfl / dmax(lam[c], 1.0e-8);
  double y[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) y[k] = dt * phi[k * ld + c];
  if (o.precond) {
    double au, av, b2;
    precond_row(load_c4(U, ld, c), o, &au, &av, &b2);
    psi[c] = y[0];
    psi[ld + c] = y[1] + au * y[3];
    psi[2 * ld + c] = y[2] + av * y[3];
    psi[3 * ld + c] = b2 * y[3];
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) psi[k * ld + c] = y[k];
  }
}

__global__ __launch_bounds__(kThreads) void k_lin_update(AdjMesh am, int stage, int rk3, int forward,
                                                         const double* __restrict__ x0, const double* __restrict__ xs,
                                                         const double* __restrict__ r, const double* __restrict__ rhs,
                                                         const double* __restrict__ lam, double cfl,
                                                         const double* __restrict__ U, FaceOpts o,
                                                         double* __restrict__ xout, double* __restrict__ partials) {
  __shared__ double red[kThreads];
  const int c = blockIdx.x * kThreads + threadIdx.x;
  double sq = 0.0;
  if (c < am.N) {
    const int ld = am.ld;
    double res[4], inc[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) res[k] = r[k * ld + c] - (rhs ? rhs[k * ld + c] : 0.0);
    if (forward) {
      const double dt = cfl / dmax(lam[c], 1.0e-8);
      double p3 = res[3];
      if (o.precond) {
        double au, av, b2;
        precond_row(load_c4(U, ld, c), o, &au, &av, &b2);
        p3 = b2 * res[3] + (au * res[1] + av * res[2]);
      }
      inc[0] = dt * res[0];
      inc[1] = dt * res[1];
      inc[2] = dt * res[2];
      inc[3] = dt * p3;
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) inc[k] = res[k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double x = xs[k * ld + c] - inc[k];
      if (rk3 && stage == 1) x = 0.75 * x0[k * ld + c] + 0.25 * x;
      if (rk3 && stage == 2) x = (1.0 / 3.0) * x0[k * ld + c] + (2.0 / 3.0) * x;
      xout[k * ld + c] = x;
      sq += res[k] * res[k];
    }
  }
  red[threadIdx.x] = sq;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0 && partials) partials[blockIdx.x] = red[0];
}

// Fixed-order sum of block partials (one workgroup).
__global__ __launch_bounds__(kThreads) void k_partials_sum(int n, const double* __restrict__ partials,
                                                           double* __restrict__ out) {
  __shared__ double red[kThreads];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += kThreads) s += partials[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// The same sum stored at hist[*counter] (counter incremented, entries beyond cap dropped): the iteration
// index lives on the device, so a captured iteration graph is replayable.
__global__ __launch_bounds__(kThreads) void k_partials_sum_counted(int n, const double* __restrict__ partials,
                                                                   double* __restrict__ hist, int* __restrict__ counter,
                                                                   int cap) {
  __shared__ double red[kThreads];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += kThreads) s += partials[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int i = *counter;
    if (i < cap) hist[i] = red[0];
    *counter = i + 1;
  }
}

__global__ __launch_bounds__(kThreads) void k_fill(long long n, double v, double* __restrict__ y) {
  const long long k = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
  if (k < n) y[k] = v;
}

// ---- Krylov vector kernels (flat vectors of n doubles) -------------------------------------------
// Block partials of <w, V_i> for i < nv, V_i = V + i * n; then a fixed-order sum per i.
__global__ __launch_bounds__(kThreads) void k_multidot(long long n, const double* __restrict__ V, int nv,
                                                       const double* __restrict__ w, double* __restrict__ partials) {
  __shared__ double red[kThreads];
  const long long per = (n + gridDim.x - 1) / gridDim.x;
  const long long lo = blockIdx.x * per;
  const long long hi = lo + per < n ? lo + per : n;
  for (int i = 0; i < nv; ++i) {
    const double* v = V + static_cast<long long>(i) * n;
    double s = 0.0;
    for (long long k = lo + threadIdx.x; k < hi; k += kThreads) s += v[k] * w[k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int t = kThreads / 2; t > 0; t >>= 1) {
      if (threadIdx.x < t) red[threadIdx.x] += red[threadIdx.x + t];
      __syncthreads();
    }
    if (threadIdx.x == 0) partials[static_cast<long long>(i) * gridDim.x + blockIdx.x] = red[0];
    __syncthreads();
  }
}

__global__ __launch_bounds__(kThreads) void k_multidot_sum(int nparts, int nv, const double* __restrict__ partials,
                                                           double* __restrict__ out) {
  __shared__ double red[kThreads];
  for (int i = 0; i < nv; ++i) {
    double s = 0.0;
    for (int k = threadIdx.x; k < nparts; k += kThreads) s += partials[static_cast<long long>(i) * nparts + k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int t = kThreads / 2; t > 0; t >>= 1) {
      if (threadIdx.x < t) red[threadIdx.x] += red[threadIdx.x + t];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[i] = red[0];
    __syncthreads();
  }
}

// y = a * y + sum_i c_i V_i  (c in device memory).
__global__ __launch_bounds__(kThreads) void k_combine(long long n, double a, const double* __restrict__ V, int nv,
                                                      const double* __restrict__ c, double* __restrict__ y) {
  const long long k = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
  if (k >= n) return;
  double s = a == 0.0 ? 0.0 : a * y[k];
  for (int i = 0; i < nv; ++i) s += c[i] * V[static_cast<long long>(i) * n + k];
  y[k] = s;
}

// y = a * x + b * y.
__global__ __launch_bounds__(kThreads) void k_axpby(long long n, double a, const double* __restrict__ x, double b,
                                                    double* __restrict__ y) {
  const long long k = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
  if (k >= n) return;
  y[k] = a * x[k] + (b == 0.0 ? 0.0 : b * y[k]);
}

}  // namespace
}  // namespace aero_dev

// ------------------------------------------------------------------------------------------------
// Launchers
// ------------------------------------------------------------------------------------------------
namespace aero_host {

using namespace aero_dev;

namespace {
inline unsigned blocks_for(long long n) { return static_cast<unsigned>((n + kThreads - 1) / kThreads); }
// Workgroups of the wide per-cell kernels: one per entry group of kEntryTile cells.
inline unsigned entry_groups(int n) { return static_cast<unsigned>((n + kEntryTile - 1) / kEntryTile); }
}  // namespace

hipError_t launch_tan_gradient(const AdjMesh& am, const double* U, const double* dU, const GeoTan& gt, double g,
                               double* G, double* dG, hipStream_t s) {
  if (am.grad == 1)
    hipLaunchKernelGGL(k_tan_gradient<1>, dim3(blocks_for(am.N)), dim3(kThreads), 0, s, am, U, dU, gt, g, G, dG);
  else
    hipLaunchKernelGGL(k_tan_gradient<0>, dim3(blocks_for(am.N)), dim3(kThreads), 0, s, am, U, dU, gt, g, G, dG);
  return hipGetLastError();
}

template <int SCHEME, int LIMITER, bool SECOND, bool FROZEN = false>
static void tan_face_t(const AdjMesh& am, const double* U, const double* dU, const double* G, const double* dG,
                       const GeoTan& gt, const FaceOptsT<Dual>& o, double* fa, hipStream_t s) {
  hipLaunchKernelGGL((k_tan_face<SCHEME, LIMITER, SECOND, FROZEN>), dim3(blocks_for(am.F)), dim3(kThreads), 0, s, am,
                     U, dU, G, dG, gt, o, fa);
}

hipError_t launch_tan_face(const AdjMesh& am, int scheme, int limiter, bool second, const double* U, const double* dU,
                           const double* G, const double* dG, const GeoTan& gt, const FaceOptsT<Dual>& o, double* fa,
                           hipStream_t s) {
  if (second && o.frozen) {  // the frozen ratios replace the limiter
    if (scheme)
      tan_face_t<1, 0, true, true>(am, U, dU, G, dG, gt, o, fa, s);
    else
      tan_face_t<0, 0, true, true>(am, U, dU, G, dG, gt, o, fa, s);
    return hipGetLastError();
  }
  const int key = (scheme ? 4 : 0) | (limiter ? 2 : 0) | (second ? 1 : 0);
  switch (key) {
    case 0: tan_face_t<0, 0, false>(am, U, dU, G, dG, gt, o, fa, s); break;
    case 1: tan_face_t<0, 0, true>(am, U, dU, G, dG, gt, o, fa, s); break;
    case 2: tan_face_t<0, 1, false>(am, U, dU, G, dG, gt, o, fa, s); break;
    case 3: tan_face_t<0, 1, true>(am, U, dU, G, dG, gt, o, fa, s); break;
    case 4: tan_face_t<1, 0, false>(am, U, dU, G, dG, gt, o, fa, s); break;
    case 5: tan_face_t<1, 0, true>(am, U, dU, G, dG, gt, o, fa, s); break;
    case 6: tan_face_t<1, 1, false>(am, U, dU, G, dG, gt, o, fa, s); break;
    default: tan_face_t<1, 1, true>(am, U, dU, G, dG, gt, o, fa, s); break;
  }
  return hipGetLastError();
}

hipError_t launch_tan_gather(const AdjMesh& am, const double* fa, double* R, double* dR, hipStream_t s) {
  hipLaunchKernelGGL(k_tan_gather, dim3(blocks_for(am.N)), dim3(kThreads), 0, s, am, fa, R, dR);
  return hipGetLastError();
}

template <int SCHEME, int LIMITER, bool SECOND, bool FROZEN = false>
static void face_jac_t(const AdjMesh& am, const double* U, const double* G, const FaceOpts& o, double* J,
                       hipStream_t s) {
  hipLaunchKernelGGL((k_face_jacobian<SCHEME, LIMITER, SECOND, FROZEN>), dim3(blocks_for(16LL * am.F)), dim3(kThreads),
                     0, s, am, U, G, o, J);
}

hipError_t launch_face_jacobian(const AdjMesh& am, int scheme, int limiter, bool second, const double* U,
                                const double* G, const FaceOpts& o, double* J, hipStream_t s) {
  if (jacobian_factored(second, o.frozen != nullptr)) {  // frozen ratios or first order: d(F A)/dW per side
    const dim3 grid(blocks_for(8LL * am.F)), block(kThreads);
    if (scheme && second)
      hipLaunchKernelGGL((k_face_jacobian_w<1, true>), grid, block, 0, s, am, U, G, o, J);
    else if (scheme)
      hipLaunchKernelGGL((k_face_jacobian_w<1, false>), grid, block, 0, s, am, U, G, o, J);
    else if (second)
      hipLaunchKernelGGL((k_face_jacobian_w<0, true>), grid, block, 0, s, am, U, G, o, J);
    else
      hipLaunchKernelGGL((k_face_jacobian_w<0, false>), grid, block, 0, s, am, U, G, o, J);
    return hipGetLastError();
  }
  const int key = (scheme ? 4 : 0) | (limiter ? 2 : 0) | (second ? 1 : 0);
  switch (key) {
    case 0: face_jac_t<0, 0, false>(am, U, G, o, J, s); break;
    case 1: face_jac_t<0, 0, true>(am, U, G, o, J, s); break;
    case 2: face_jac_t<0, 1, false>(am, U, G, o, J, s); break;
    case 3: face_jac_t<0, 1, true>(am, U, G, o, J, s); break;
    case 4: face_jac_t<1, 0, false>(am, U, G, o, J, s); break;
    case 5: face_jac_t<1, 0, true>(am, U, G, o, J, s); break;
    case 6: face_jac_t<1, 1, false>(am, U, G, o, J, s); break;
    default: face_jac_t<1, 1, true>(am, U, G, o, J, s); break;
  }
  return hipGetLastError();
}

template <bool FACT>
static void adj_product_t(const AdjMesh& am, const double* J, const double* U, const double* psi, double g,
                          double* Gbar, double* out, hipStream_t s) {
  if (entry_groups(am.N) < kWideBelow)
    hipLaunchKernelGGL((k_adj_gather<true, FACT>), dim3(entry_groups(am.N)), dim3(kWideThreads), 0, s, am, J, psi, out,
                       Gbar);
  else
    hipLaunchKernelGGL((k_adj_gather<false, FACT>), dim3(entry_groups(am.N)), dim3(kEntryTile), 0, s, am, J, psi, out,
                       Gbar);
  if (am.grad == 1)
    hipLaunchKernelGGL((k_adj_grad<1, FACT>), dim3(blocks_for(am.N)), dim3(kThreads), 0, s, am, U, g, Gbar, out);
  else
    hipLaunchKernelGGL((k_adj_grad<0, FACT>), dim3(blocks_for(am.N)), dim3(kThreads), 0, s, am, U, g, Gbar, out);
}

hipError_t launch_adj_product(const AdjMesh& am, const double* J, bool fact, const double* U, const double* psi,
                              double g, double* Gbar, double* out, hipStream_t s) {
  if (fact)
    adj_product_t<true>(am, J, U, psi, g, Gbar, out, s);
  else
    adj_product_t<false>(am, J, U, psi, g, Gbar, out, s);
  return hipGetLastError();
}

hipError_t launch_objective_gradient(const AdjMesh& am, const double* U, double g, const double* coef, double* out,
                                     hipStream_t s) {
  hipLaunchKernelGGL(k_objective_gradient, dim3(1), dim3(64), 0, s, am, U, g, coef, out);
  return hipGetLastError();
}

template <bool WIDE, bool FACT>
static void adj_stage_t(const AdjMesh& am, const double* J, const double* U, const FaceOpts& o, int stage, int rk3,
                        const double* x0, const double* xs, const double* rhs, const double* lam, double cfl,
                        double* Gbar, double* direct, double* xout, double* partials, double* psi, hipStream_t s) {
  const dim3 grid(entry_groups(am.N)), block(WIDE ? kWideThreads : kEntryTile);
  hipLaunchKernelGGL((k_adj_gather<WIDE, FACT>), grid, block, 0, s, am, J, psi, direct, Gbar);
  if (am.grad == 1)
    hipLaunchKernelGGL((k_adj_grad_update<WIDE, 1, FACT>), grid, block, 0, s, am, U, o, Gbar, direct, stage, rk3, x0,
                       xs, rhs, lam, cfl, xout, partials, psi);
  else
    hipLaunchKernelGGL((k_adj_grad_update<WIDE, 0, FACT>), grid, block, 0, s, am, U, o, Gbar, direct, stage, rk3, x0,
                       xs, rhs, lam, cfl, xout, partials, psi);
}

hipError_t launch_adj_stage(const AdjMesh& am, const double* J, bool fact, const double* U, const FaceOpts& o,
                            int stage, int rk3, const double* x0, const double* xs, const double* rhs,
                            const double* lam, double cfl, double* Gbar, double* direct, double* xout,
                            double* partials, double* psi, hipStream_t s) {
  const bool wide = entry_groups(am.N) < kWideBelow;
  if (wide && fact)
    adj_stage_t<true, true>(am, J, U, o, stage, rk3, x0, xs, rhs, lam, cfl, Gbar, direct, xout, partials, psi, s);
  else if (wide)
    adj_stage_t<true, false>(am, J, U, o, stage, rk3, x0, xs, rhs, lam, cfl, Gbar, direct, xout, partials, psi, s);
  else if (fact)
    adj_stage_t<false, true>(am, J, U, o, stage, rk3, x0, xs, rhs, lam, cfl, Gbar, direct, xout, partials, psi, s);
  else
    adj_stage_t<false, false>(am, J, U, o, stage, rk3, x0, xs, rhs, lam, cfl, Gbar, direct, xout, partials, psi, s);
  return hipGetLastError();
}

template <bool WIDE, int FM>
static void adj_batch_stage_t(int grad, const AdjBatchSolve* solves, const AdjBatchCtl* ctl, const int4* work,
                              int nwork, int stage, int rk3, int ix0, int ixs, int ixo, int want_partials,
                              hipStream_t s) {
  const dim3 grid(nwork), block(WIDE ? kWideThreads : kEntryTile);
  hipLaunchKernelGGL((k_adj_gather_batch<WIDE, FM>), grid, block, 0, s, solves, ctl, work);
  if (grad == 1)
    hipLaunchKernelGGL((k_adj_grad_update_batch<WIDE, 1, FM>), grid, block, 0, s, solves, ctl, work, stage, rk3, ix0,
                       ixs, ixo, want_partials);
  else
    hipLaunchKernelGGL((k_adj_grad_update_batch<WIDE, 0, FM>), grid, block, 0, s, solves, ctl, work, stage, rk3, ix0,
                       ixs, ixo, want_partials);
}

template <int FM>
static void adj_batch_stage_fm(int grad, const AdjBatchSolve* solves, const AdjBatchCtl* ctl, const int4* work,
                               int nwork, int stage, int rk3, int ix0, int ixs, int ixo, int want_partials,
                               hipStream_t s) {
  if (nwork < kWideBelow)
    adj_batch_stage_t<true, FM>(grad, solves, ctl, work, nwork, stage, rk3, ix0, ixs, ixo, want_partials, s);
  else
    adj_batch_stage_t<false, FM>(grad, solves, ctl, work, nwork, stage, rk3, ix0, ixs, ixo, want_partials, s);
}

// Work items are (solve pair, entry group): the four-lanes-per-cell shape below kWideBelow of them.
// fact_mode: 0 every solve's Jacobians unfactored, 1 every solve's factored, 2 mixed.
hipError_t launch_adj_batch_stage(int grad, int fact_mode, const AdjBatchSolve* solves, const AdjBatchCtl* ctl,
                                  const int4* work, int nwork, int stage, int rk3, int ix0, int ixs, int ixo,
                                  int want_partials, hipStream_t s) {
  if (fact_mode == 1)
    adj_batch_stage_fm<1>(grad, solves, ctl, work, nwork, stage, rk3, ix0, ixs, ixo, want_partials, s);
  else if (fact_mode == 0)
    adj_batch_stage_fm<0>(grad, solves, ctl, work, nwork, stage, rk3, ix0, ixs, ixo, want_partials, s);
  else
    adj_batch_stage_fm<2>(grad, solves, ctl, work, nwork, stage, rk3, ix0, ixs, ixo, want_partials, s);
  return hipGetLastError();
}

hipError_t launch_adj_batch_control(const AdjBatchSolve* solves, AdjBatchCtl* ctl, int nsolves, double tol,
                                    int max_iters, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_batch_control, dim3(nsolves), dim3(kThreads), 0, s, solves, ctl, tol, max_iters);
  return hipGetLastError();
}

hipError_t launch_adj_scale(const AdjMesh& am, const double* lam, double cfl, const double* U, const FaceOpts& o,
                            const double* phi, double* psi, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_scale, dim3(blocks_for(am.N)), dim3(kThreads), 0, s, am, lam, cfl, U, o, phi, psi);
  return hipGetLastError();
}

hipError_t launch_lin_update(const AdjMesh& am, int stage, int rk3, int forward, const double* x0, const double* xs,
                             const double* r, const double* rhs, const double* lam, double cfl, const double* U,
                             const FaceOpts& o, double* xout, double* partials, hipStream_t s) {
  hipLaunchKernelGGL(k_lin_update, dim3(blocks_for(am.N)), dim3(kThreads), 0, s, am, stage, rk3, forward, x0, xs, r,
                     rhs, lam, cfl, U, o, xout, partials);
  return hipGetLastError();
}

constexpr int kDotBlocks = 512;

hipError_t launch_multidot(long long n, const double* V, int nv, const double* w, double* partials, double* out,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_multidot, dim3(kDotBlocks), dim3(kThreads), 0, s, n, V, nv, w, partials);
  hipLaunchKernelGGL(k_multidot_sum, dim3(1), dim3(kThreads), 0, s, kDotBlocks, nv, partials, out);
  return hipGetLastError();
}

hipError_t launch_combine(long long n, double a, const double* V, int nv, const double* c, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_combine, dim3(blocks_for(n)), dim3(kThreads), 0, s, n, a, V, nv, c, y);
  return hipGetLastError();
}

hipError_t launch_axpby(long long n, double a, const double* x, double b, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_axpby, dim3(blocks_for(n)), dim3(kThreads), 0, s, n, a, x, b, y);
  return hipGetLastError();
}

hipError_t launch_partials_sum_counted(int nparts, const double* partials, double* hist, int* counter, int cap,
                                      hipStream_t s) {
  hipLaunchKernelGGL(k_partials_sum_counted, dim3(1), dim3(kThreads), 0, s, nparts, partials, hist, counter, cap);
  return hipGetLastError();
}

hipError_t launch_fill(long long n, double v, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(n)), dim3(kThreads), 0, s, n, v, y);
  return hipGetLastError();
}

hipError_t launch_partials_sum(int nparts, const double* partials, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_partials_sum, dim3(1), dim3(kThreads), 0, s, nparts, partials, out);
  return hipGetLastError();
}

}  // namespace aero_host
