// aero_kernels.hip -- gfx950 FP64 kernels of the Euler residual path and their launchers.
//
// Layout (device, internal cell order):
//   state U[4][N] SoA (rho, rhou, rhov, rhoE), gradients G[8][N] (rho_x, rho_y, u_x, u_y, v_x, v_y,
//   p_x, p_y), residual R[4][N], spectral radius lam[N].
//   Cells are grouped in blocks of BC consecutive cells; one workgroup owns one block.  Every face
//   touching a block's cells is one "slot" of that block (faces on a block boundary are slots of both
//   blocks), stored in slot-major SoA streams so that phase 1 reads them coalesced.
// Each kernel runs in two phases:
//   1. one thread per slot computes the face quantity (Green-Gauss face value, or the full
//      reconstruction + Riemann flux) once per block and stages it in LDS;
//   2. one thread per cell gathers its faces' staged values in increasing reference face index, which
//      is the summation order of the reference's serial face loop (reconstruction.cpp:85-107,
//      residual_assembly.cpp:97-181).  No floating-point atomics; results are bitwise reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>

#include "aero_internal.hpp"
#include "face_eval.hpp"
#include "euler_device.hpp"

namespace aero_dev {

// Gathers at a 32-bit byte offset from a wave-uniform base.  Indexing an array with a per-lane int costs
// a sign extension and a 64-bit add per load; with the base in scalar registers and the offset formed
// once per index, every load of that index takes the SGPR-base form of global_load and costs no vector
// instruction.  Offsets stay below 2^32: meshes have fewer than 2^28 cells and face slots (checked at
// mesh creation).
template <class T> __device__ __forceinline__ T at_off(const T* base, unsigned off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}
template <class T> __device__ __forceinline__ void st_off(T* base, unsigned off, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + off) = v;
}
__device__ __forceinline__ unsigned off8(int i) { return static_cast<unsigned>(i) << 3; }
__device__ __forceinline__ unsigned off4(int i) { return static_cast<unsigned>(i) << 2; }
__device__ __forceinline__ unsigned off2(int i) { return static_cast<unsigned>(i) << 1; }

__device__ __forceinline__ C4 load_u(const double* __restrict__ U, int ld, int c) {
  const unsigned o = off8(c);
  C4 s;
  s.r = at_off(U, o);
  s.ru = at_off(U + ld, o);
  s.rv = at_off(U + 2 * ld, o);
  s.E = at_off(U + 3 * ld, o);
  return s;
}

// Workgroup -> block.  Workgroups are dispatched round-robin over the 8 XCDs; with `xcd_remap` the
// blocks handled by one XCD form a contiguous run, so neighbouring blocks share that XCD's L2.
__device__ __forceinline__ int block_of(const DevMesh& m) {
  const int i = blockIdx.x;
  if (!m.xcd_remap) return i;
  const int per = (m.nblk + 7) / 8;
  return (i & 7) * per + (i >> 3);
}

// Weighted least-squares gradient, per block (no reference function; the reference roadmap lists it,
// docs/project_outline.md:49).  Per interior face, with d = x_neighbour - x_owner, w = 1 / (d.d) and
// dq = q(neighbour) - q(owner), both cells add (w dx) dx, (w dx) dy, (w dy) dy, (w dx) dq, (w dy) dq in
// increasing reference face index; the 2x2 normal equations are solved by Cramer's rule, and a cell
// whose det = a c - b b is not above 1e-12 (a + c)^2 gets a zero gradient.  Boundary slots stage zeros,
// which leave the sums unchanged.
template <int BC>
__device__ __forceinline__ void lsq_gradient_body(const DevMesh& m, int b, const double* __restrict__ U,
                                                  double* __restrict__ G, double g, double* lds) {
  const int s0 = m.slot_off[b];
  const int ns = m.slot_off[b + 1] - s0;
  const int smax = m.smax;
  double* swx = lds;
  double* swy = lds + smax;
  double* sdx = lds + 2 * smax;
  double* sdy = lds + 3 * smax;
  double* q0 = lds + 4 * smax;
  double* q1 = lds + 5 * smax;
  double* q2 = lds + 6 * smax;
  double* q3 = lds + 7 * smax;
  const int ld = m.ld;
  for (int i = threadIdx.x; i < ns; i += BC) {
    const int s = s0 + i;
    const int o = m.s_o[s];
    const int n = slot_nb(m.s_nm[s]);
    double wdx = 0.0, wdy = 0.0, dx = 0.0, dy = 0.0, d0 = 0.0, d1 = 0.0, d2 = 0.0, d3 = 0.0;
    if (n >= 0) {
      const P4 wo = c2p(load_u(U, ld, o), g);
      const P4 wn = c2p(load_u(U, ld, n), g);
      dx = m.ccx[n] - m.ccx[o];
      dy = m.ccy[n] - m.ccy[o];
      const double w = 1.0 / (dx * dx + dy * dy);
      wdx = w * dx;
      wdy = w * dy;
      d0 = wn.r - wo.r;
      d1 = wn.u - wo.u;
      d2 = wn.v - wo.v;
      d3 = wn.p - wo.p;
    }
    swx[i] = wdx;
    swy[i] = wdy;
    sdx[i] = dx;
    sdy[i] = dy;
    q0[i] = d0;
    q1[i] = d1;
    q2[i] = d2;
    q3[i] = d3;
  }
  __syncthreads();
  const int c = b * BC + threadIdx.x;
  if (c >= m.N) return;
  double a = 0.0, bb = 0.0, cc = 0.0;
  double rx0 = 0.0, ry0 = 0.0, rx1 = 0.0, ry1 = 0.0, rx2 = 0.0, ry2 = 0.0, rx3 = 0.0, ry3 = 0.0;
  const int e1 = m.c_off[c + 1];
  for (int e = m.c_off[c]; e < e1; ++e) {
    const int l = m.c_ent[e] >> 1;
    const double wdx = swx[l], wdy = swy[l], dx = sdx[l], dy = sdy[l];
    a += wdx * dx;
    bb += wdx * dy;
    cc += wdy * dy;
    const double v0 = q0[l], v1 = q1[l], v2 = q2[l], v3 = q3[l];
    rx0 += wdx * v0; ry0 += wdy * v0;
    rx1 += wdx * v1; ry1 += wdy * v1;
    rx2 += wdx * v2; ry2 += wdy * v2;
    rx3 += wdx * v3; ry3 += wdy * v3;
  }
  const double det = a * cc - bb * bb;
  const double tr = a + cc;
  double gx0 = 0.0, gy0 = 0.0, gx1 = 0.0, gy1 = 0.0, gx2 = 0.0, gy2 = 0.0, gx3 = 0.0, gy3 = 0.0;
  if (det > 1.0e-12 * (tr * tr)) {
    const double inv = 1.0 / det;
    gx0 = (cc * rx0 - bb * ry0) * inv; gy0 = (a * ry0 - bb * rx0) * inv;
    gx1 = (cc * rx1 - bb * ry1) * inv; gy1 = (a * ry1 - bb * rx1) * inv;
    gx2 = (cc * rx2 - bb * ry2) * inv; gy2 = (a * ry2 - bb * rx2) * inv;
    gx3 = (cc * rx3 - bb * ry3) * inv; gy3 = (a * ry3 - bb * rx3) * inv;
  }
  G[0 * ld + c] = gx0;
  G[1 * ld + c] = gy0;
  G[2 * ld + c] = gx1;
  G[3 * ld + c] = gy1;
  G[4 * ld + c] = gx2;
  G[5 * ld + c] = gy2;
  G[6 * ld + c] = gx3;
  G[7 * ld + c] = gy3;
}

// Green-Gauss gradient (reconstruction.cpp:73-119), per block.
template <int BC>
__device__ __forceinline__ void gg_gradient_body(const DevMesh& m, int b, const double* __restrict__ U,
                                                 double* __restrict__ G, double g, double* lds) {
  const int s0 = m.slot_off[b];
  const int ns = m.slot_off[b + 1] - s0;
  const int smax = m.smax;
  // Phase 1 stages the four face values only (4 doubles per slot: more workgroups fit the LDS); phase 2
  // forms n A from the slot stream again, the same product.
  double* q0 = lds;
  double* q1 = lds + smax;
  double* q2 = lds + 2 * smax;
  double* q3 = lds + 3 * smax;
  const int ld = m.ld;
  for (int i = threadIdx.x; i < ns; i += BC) {
    const unsigned s4 = off4(s0 + i);
    const int o = at_off(m.s_o, s4);
    const int n = slot_nb(at_off(m.s_nm, s4));
    const P4 wo = c2p(load_u(U, ld, o), g);
    double a0 = wo.r, a1 = wo.u, a2 = wo.v, a3 = wo.p;
    if (n >= 0) {
      const P4 wn = c2p(load_u(U, ld, n), g);
      a0 = 0.5 * (wo.r + wn.r);
      a1 = 0.5 * (wo.u + wn.u);
      a2 = 0.5 * (wo.v + wn.v);
      a3 = 0.5 * (wo.p + wn.p);
    }
    q0[i] = a0;
    q1[i] = a1;
    q2[i] = a2;
    q3[i] = a3;
  }
  __syncthreads();
  const int c = b * BC + threadIdx.x;
  if (c >= m.N) return;
  double gx0 = 0.0, gy0 = 0.0, gx1 = 0.0, gy1 = 0.0, gx2 = 0.0, gy2 = 0.0, gx3 = 0.0, gy3 = 0.0;
  const int e1 = at_off(m.c_off, off4(c + 1));
  for (int e = at_off(m.c_off, off4(c)); e < e1; ++e) {
    const int ent = at_off(m.c_ent, off2(e));
    const int l = ent >> 1;
    const unsigned sl = off8(s0 + l);
    const double A = at_off(m.s_A, sl);
    const double x = at_off(m.s_nx, sl) * A, y = at_off(m.s_ny, sl) * A;
    const double v0 = q0[l], v1 = q1[l], v2 = q2[l], v3 = q3[l];
    if (ent & 1) {
      gx0 -= v0 * x; gy0 -= v0 * y;
      gx1 -= v1 * x; gy1 -= v1 * y;
      gx2 -= v2 * x; gy2 -= v2 * y;
      gx3 -= v3 * x; gy3 -= v3 * y;
    } else {
      gx0 += v0 * x; gy0 += v0 * y;
      gx1 += v1 * x; gy1 += v1 * y;
      gx2 += v2 * x; gy2 += v2 * y;
      gx3 += v3 * x; gy3 += v3 * y;
    }
  }
  const unsigned c8 = off8(c);
  const double iv = 1.0 / dmax(at_off(m.vol, c8), 1.0e-12);
  st_off(G, c8, gx0 * iv);
  st_off(G + ld, c8, gy0 * iv);
  st_off(G + 2 * ld, c8, gx1 * iv);
  st_off(G + 3 * ld, c8, gy1 * iv);
  st_off(G + 4 * ld, c8, gx2 * iv);
  st_off(G + 5 * ld, c8, gy2 * iv);
  st_off(G + 6 * ld, c8, gx3 * iv);
  st_off(G + 7 * ld, c8, gy3 * iv);
}

template <int BC, int GK>
__device__ __forceinline__ void gradient_body(const DevMesh& m, int b, const double* __restrict__ U,
                                              double* __restrict__ G, double g, double* lds) {
  if (GK == 1)
    lsq_gradient_body<BC>(m, b, U, G, g, lds);
  else
    gg_gradient_body<BC>(m, b, U, G, g, lds);
}

// LDS doubles per face slot of the gradient kernels.
constexpr int grad_lds_per_slot(int gk) { return gk == 1 ? 8 : 4; }

template <int BC, int GK>
__global__ __launch_bounds__(BC) void k_gradient(DevMesh m, const double* __restrict__ U, double* __restrict__ G,
                                                  double g, int* __restrict__ diag) {
  extern __shared__ double lds[];
  // Optional reset of the residual counters that the flux kernel launched after this one accumulates
  // (first-order fallback, diffusive fallback, first failing face, non-finite flag).
  if (diag && blockIdx.x == 0 && threadIdx.x == 0) {
    diag[0] = 0;
    diag[1] = 0;
    diag[2] = INT_MAX;
    diag[3] = 0;
  }
  const int b = block_of(m);
  if (b >= m.nblk) return;
  gradient_body<BC, GK>(m, b, U, G, g, lds);
}

// Wavefront sum of an int (64 lanes).
__device__ __forceinline__ int wave_sum_i(int v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ int wave_min_i(int v) {
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(v, off, 64);
    v = o < v ? o : v;
  }
  return v;
}

// One face slot of flux_body's phase 1: the face flux times the area staged in LDS slot i, the fallback
// counters updated.  FAST (PLAIN): interior faces only, with the exact shortcuts of flux_of_states; for a
// boundary slot or one outside the shortcuts' bounds (*ok false) nothing is staged or counted.
template <int SCHEME, int LIMITER, bool SECOND, bool FROZEN, bool PLAIN, bool FAST>
__device__ __forceinline__ void flux_slot(const DevMesh& m, int s, int i, const FaceOpts& o,
                                          const double* __restrict__ U, const double* __restrict__ G, double* lds,
                                          int* n_fo, int* n_diff, int* first_fail, bool* ok) {
  const int smax = m.smax;
  const int ld = m.ld;
  const unsigned s4 = off4(s), s8 = off8(s);
  const int nm = at_off(m.s_nm, s4);
  const int cn = slot_nb(nm);
  if constexpr (FAST) {
    if (cn < 0) {
      *ok = false;
      return;
    }
  }
  const int co = at_off(m.s_o, s4);
  const int meta = slot_meta(nm);
  const double nx = at_off(m.s_nx, s8);
  const double ny = at_off(m.s_ny, s8);
  const double A = at_off(m.s_A, s8);
  Flux F;
  bool reverted = false;
  if (cn >= 0) {
    double go[8], gn[8];
    double dox = 0.0, doy = 0.0, dnx = 0.0, dny = 0.0;
    if (SECOND) {
      const unsigned oo = off8(co), on = off8(cn);
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        go[j] = at_off(G + j * ld, oo);
        gn[j] = at_off(G + j * ld, on);
      }
      const double fx = at_off(m.s_fx, s8), fy = at_off(m.s_fy, s8);
      dox = fx - at_off(m.ccx, oo);
      doy = fy - at_off(m.ccy, oo);
      dnx = fx - at_off(m.ccx, on);
      dny = fy - at_off(m.ccy, on);
    }
    double fr[8];
    if (FROZEN && SECOND) {  // one 16-byte load per slot: eight 16-bit ratios in units of 2^-15
      const uint4 q4 = reinterpret_cast<const uint4*>(o.frozen_slots)[s];
      const unsigned int w[4] = {q4.x, q4.y, q4.z, q4.w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        fr[2 * j] = static_cast<double>(w[j] & 0xffffu) * kRatioStep;
        fr[2 * j + 1] = static_cast<double>(w[j] >> 16) * kRatioStep;
      }
    }
    F = interior_face<SCHEME, LIMITER, SECOND, FROZEN, PLAIN, FAST>(load_u(U, ld, co), load_u(U, ld, cn), go, gn,
                                                                    nx, ny, dox, doy, dnx, dny, o, &reverted, fr,
                                                                    ok);
  } else if constexpr (!FAST) {
    F = boundary_face<SCHEME, PLAIN>(load_u(U, ld, co), (meta & kMetaBcMask) == 2, nx, ny, o);
  }
  if constexpr (FAST) {
    if (!*ok) return;
  }
  if (reverted && (meta & kMetaPrimary)) {
    ++*n_fo;
    *first_fail = min(*first_fail, m.s_face[s]);
  }
  if (F.fallback && (meta & kMetaPrimary)) ++*n_diff;
  lds[i] = F.f0 * A;
  lds[smax + i] = F.f1 * A;
  lds[2 * smax + i] = F.f2 * A;
  lds[3 * smax + i] = F.f3 * A;
  lds[4 * smax + i] = F.lam * A;
}

// Face flux + accumulation (residual_assembly.cpp:73-182), with the RK stage tail fused in MODE 1.
// MODE 0: write R and lam.   MODE 1: LTS stage update into Unext (euler_solver.cpp:893-899,
// preconditioning.cpp:51-76, time_integrator.cpp:31-104); optional R / lam write and norm partials.
// PLAIN: the options have the preconditioner, the low-Mach fix and the all-speed fix off (plain_options).
template <int BC, int SCHEME, int LIMITER, bool SECOND, int MODE, bool FROZEN = false, bool PLAIN = false>
__device__ __forceinline__ void flux_body(const DevMesh& m, int b, const FaceOpts& o, const double* __restrict__ U,
                                          const double* __restrict__ G, double* __restrict__ R,
                                          double* __restrict__ lam, int* __restrict__ diag, const StageArgs& st,
                                          double* lds) {
  const int s0 = m.slot_off[b];
  const int ns = m.slot_off[b + 1] - s0;
  const int smax = m.smax;
  double* f0 = lds;
  double* f1 = lds + smax;
  double* f2 = lds + 2 * smax;
  double* f3 = lds + 3 * smax;
  double* fw = lds + 4 * smax;
  const int ld = m.ld;
  const double g = o.g;
  int n_fo = 0, n_diff = 0, first_fail = 0x7fffffff;
  // PLAIN: the interior slots are evaluated with the exact shortcuts of the FAST face flux first; a
  // boundary slot, or one outside the shortcuts' operand bounds, sets its bit in redo (bit k: this
  // thread's k-th slot; a block has at most 64 KiB / 40 B = 1638 slots, checked at mesh creation, so at
  // most 13 per thread) and is evaluated after the loop without them, which keeps the boundary and
  // unbounded paths out of the loop's registers.
  unsigned redo = 0;
  for (int i = threadIdx.x, k = 0; i < ns; i += BC, ++k) {
    bool ok = true;
    flux_slot<SCHEME, LIMITER, SECOND, FROZEN, PLAIN, PLAIN>(m, s0 + i, i, o, U, G, lds, &n_fo, &n_diff, &first_fail,
                                                             &ok);
    if (!ok) redo |= 1u << k;
  }
  if (PLAIN && redo) {
    for (int i = threadIdx.x, k = 0; i < ns; i += BC, ++k) {
      if ((redo >> k) & 1u)
        flux_slot<SCHEME, LIMITER, SECOND, FROZEN, PLAIN, false>(m, s0 + i, i, o, U, G, lds, &n_fo, &n_diff,
                                                                 &first_fail, nullptr);
    }
  }
  // Fallback counters: one atomic per wavefront that saw any.
  {
    const int a = wave_sum_i(n_fo);
    const int d = wave_sum_i(n_diff);
    const int ff = wave_min_i(first_fail);
    if ((threadIdx.x & 63) == 0) {
      if (a) atomicAdd(&diag[0], a);
      if (d) atomicAdd(&diag[1], d);
      if (ff != 0x7fffffff) atomicMin(&diag[2], ff);
    }
  }
  __syncthreads();
  const int c = b * BC + threadIdx.x;
  double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0, w = 0.0;
  const bool active = c < m.N;
  const unsigned c8 = off8(c);
  if (active) {
    const int e1 = at_off(m.c_off, off4(c + 1));
    for (int e = at_off(m.c_off, off4(c)); e < e1; ++e) {
      const int ent = at_off(m.c_ent, off2(e));
      const int l = ent >> 1;
      if (ent & 1) {
        r0 -= f0[l];
        r1 -= f1[l];
        r2 -= f2[l];
        r3 -= f3[l];
      } else {
        r0 += f0[l];
        r1 += f1[l];
        r2 += f2[l];
        r3 += f3[l];
      }
      w += fw[l];
    }
  }
  if (MODE == 0) {
    if (active) {
      st_off(R, c8, r0);
      st_off(R + ld, c8, r1);
      st_off(R + 2 * ld, c8, r2);
      st_off(R + 3 * ld, c8, r3);
      st_off(lam, c8, w);
    }
    return;
  } else {
    double nl1 = 0.0, nl2 = 0.0, nlinf = 0.0, nr2 = 0.0;
    int bad = 0;
    if (active) {
      if (st.write_res) {
        st_off(R, c8, r0);
        st_off(R + ld, c8, r1);
        st_off(R + 2 * ld, c8, r2);
        st_off(R + 3 * ld, c8, r3);
        st_off(lam, c8, w);
      }
      const C4 us = load_u(U, ld, c);
      double p3 = r3;
      if (!PLAIN && o.precond) {  // preconditioning.cpp:186-199, on the stage state's primitive
        const P4 wp = c2p(us, g);
        const double bt = cell_beta(wp, o);
        const double b2 = bt * bt;
        const double rho = dmax(wp.r, 1.0e-8);
        const double inv = 1.0 / rho;
        const double udr = wp.u * r1 + wp.v * r2;
        p3 = b2 * r3 + (b2 - 1.0) * inv * udr;
      }
      const double dt = st.cfl / dmax(w, 1.0e-8);  // time_integrator.cpp:268-273
      C4 s;
      s.r = us.r - dt * r0;
      s.ru = us.ru - dt * r1;
      s.rv = us.rv - dt * r2;
      s.E = us.E - dt * p3;
      if (st.stage >= 1) {
        const C4 un = load_u(st.Un, ld, c);
        if (st.stage == 1) {
          s.r = 0.75 * un.r + 0.25 * s.r;
          s.ru = 0.75 * un.ru + 0.25 * s.ru;
          s.rv = 0.75 * un.rv + 0.25 * s.rv;
          s.E = 0.75 * un.E + 0.25 * s.E;
        } else {
          const double t1 = 1.0 / 3.0, t2 = 2.0 / 3.0;
          s.r = t1 * un.r + t2 * s.r;
          s.ru = t1 * un.ru + t2 * s.ru;
          s.rv = t1 * un.rv + t2 * s.rv;
          s.E = t1 * un.E + t2 * s.E;
        }
      }
      enforce_physical(&s, g);
      st_off(st.Unext, c8, s.r);
      st_off(st.Unext + ld, c8, s.ru);
      st_off(st.Unext + 2 * ld, c8, s.rv);
      st_off(st.Unext + 3 * ld, c8, s.E);
      if (st.last_stage) {
        if (!(fin(s.r) && fin(s.ru) && fin(s.rv) && fin(s.E))) {
          bad = 1;
        } else {
          const P4 q = c2p(s, g);
          if (!(fin(q.r) && fin(q.u) && fin(q.v) && fin(q.p))) bad = 1;
        }
        // residual norms of the un-preconditioned residual (euler_solver.cpp:315-333)
        const double iv = 1.0 / dmax(at_off(m.vol, c8), 1.0e-12);
        const double v0 = fabs(r0 * iv), v1 = fabs(r1 * iv), v2 = fabs(r2 * iv), v3 = fabs(r3 * iv);
        nl1 = v0 + v1 + v2 + v3;
        nl2 = v0 * v0 + v1 * v1 + v2 * v2 + v3 * v3;
        nlinf = dmax(dmax(v0, v1), dmax(v2, v3));
        nr2 = v0 * v0;
      }
    }
    if (st.last_stage) {
      // Block reduction in a fixed order: wavefront butterflies, then wavefronts in index order.
      for (int off = 32; off > 0; off >>= 1) {
        nl1 += __shfl_xor(nl1, off, 64);
        nl2 += __shfl_xor(nl2, off, 64);
        nlinf = dmax(nlinf, __shfl_xor(nlinf, off, 64));
        nr2 += __shfl_xor(nr2, off, 64);
        bad |= __shfl_xor(bad, off, 64);
      }
      __syncthreads();
      constexpr int NW = BC / 64;
      double* red = lds;  // reuse: slot data no longer needed
      const int wv = threadIdx.x >> 6;
      if ((threadIdx.x & 63) == 0) {
        red[4 * wv + 0] = nl1;
        red[4 * wv + 1] = nl2;
        red[4 * wv + 2] = nlinf;
        red[4 * wv + 3] = nr2;
        if (bad) atomicOr(&diag[3], 1);
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        for (int k = 0; k < NW; ++k) {
          a0 += red[4 * k + 0];
          a1 += red[4 * k + 1];
          a2 = dmax(a2, red[4 * k + 2]);
          a3 += red[4 * k + 3];
        }
        st.partials[4 * b + 0] = a0;
        st.partials[4 * b + 1] = a1;
        st.partials[4 * b + 2] = a2;
        st.partials[4 * b + 3] = a3;
      }
    }
  }
}

// Waves per SIMD the flux kernels must fit (register budget 512 / 4 = 128 VGPRs).  At the compiler's
// own choice the HLLC + venkat kernel takes 138 VGPRs (3 waves); forcing 4 spills 40 bytes per lane and
// still measured 5% faster on the 1M-cell mixed mesh (102.7 us vs 107.9 us per launch).
#ifndef AERO_FLUX_WAVES_PER_EU
#define AERO_FLUX_WAVES_PER_EU 4
#endif

template <int BC, int SCHEME, int LIMITER, bool SECOND, int MODE, bool FROZEN = false, bool PLAIN = false>
__global__ __launch_bounds__(BC, AERO_FLUX_WAVES_PER_EU) void k_flux(DevMesh m, FaceOpts o, const double* __restrict__ U,
                                              const double* __restrict__ G, double* __restrict__ R,
                                              double* __restrict__ lam, int* __restrict__ diag, StageArgs st) {
  extern __shared__ double lds[];
  const int b = block_of(m);
  if (b >= m.nblk) return;
  flux_body<BC, SCHEME, LIMITER, SECOND, MODE, FROZEN, PLAIN>(m, b, o, U, G, R, lam, diag, st, lds);
}

// The residual on the face-coloured layout: one thread per face of colour range [begin, end); the face
// flux of flux_body's phase 1, then F A added to the owner's and subtracted from the neighbour's R (and
// lam A to both) in place: no two faces of a colour share a cell.
template <int SCHEME, int LIMITER, bool SECOND, bool FROZEN>
__global__ __launch_bounds__(256) void k_flux_colour(DevMesh m, ColourFaces cf, int begin, int end, FaceOpts o,
                                                     const double* __restrict__ U, const double* __restrict__ G,
                                                     double* __restrict__ R, double* __restrict__ lam,
                                                     int* __restrict__ diag) {
  const int i = begin + blockIdx.x * 256 + threadIdx.x;
  const int ld = m.ld;
  int n_fo = 0, n_diff = 0, first_fail = 0x7fffffff;
  if (i < end) {
    const int co = cf.o[i], cn = cf.n[i];
    const double nx = cf.nx[i], ny = cf.ny[i], A = cf.A[i];
    Flux F;
    if (cn >= 0) {
      double go[8], gn[8];
      double dox = 0.0, doy = 0.0, dnx = 0.0, dny = 0.0;
      if (SECOND) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          go[j] = G[j * ld + co];
          gn[j] = G[j * ld + cn];
        }
        const double fx = cf.fx[i], fy = cf.fy[i];
        dox = fx - m.ccx[co];
        doy = fy - m.ccy[co];
        dnx = fx - m.ccx[cn];
        dny = fy - m.ccy[cn];
      }
      double fr[8];
      if (FROZEN && SECOND) {  // face-order ratios, already held at 2^-15
        const int f = cf.face[i];
#pragma unroll
        for (int j = 0; j < 8; ++j) fr[j] = o.frozen[static_cast<size_t>(j) * m.nfaces + f];
      }
      bool reverted;
      F = interior_face<SCHEME, LIMITER, SECOND, FROZEN>(load_u(U, ld, co), load_u(U, ld, cn), go, gn, nx, ny, dox, doy,
                                                         dnx, dny, o, &reverted, fr);
      if (reverted) {
        ++n_fo;
        first_fail = cf.face[i];
      }
    } else {
      F = boundary_face<SCHEME>(load_u(U, ld, co), (cf.meta[i] & kMetaBcMask) == 2, nx, ny, o);
    }
    if (F.fallback) ++n_diff;
    const double a0 = F.f0 * A, a1 = F.f1 * A, a2 = F.f2 * A, a3 = F.f3 * A, w = F.lam * A;
    R[co] += a0;
    R[ld + co] += a1;
    R[2 * ld + co] += a2;
    R[3 * ld + co] += a3;
    lam[co] += w;
    if (cn >= 0) {
      R[cn] -= a0;
      R[ld + cn] -= a1;
      R[2 * ld + cn] -= a2;
      R[3 * ld + cn] -= a3;
      lam[cn] += w;
    }
  }
  const int a = wave_sum_i(n_fo);
  const int d = wave_sum_i(n_diff);
  const int ff = wave_min_i(first_fail);
  if ((threadIdx.x & 63) == 0) {
    if (a) atomicAdd(&diag[0], a);
    if (d) atomicAdd(&diag[1], d);
    if (ff != 0x7fffffff) atomicMin(&diag[2], ff);
  }
}

__global__ __launch_bounds__(256) void k_zero(long long n, double* __restrict__ y) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) y[i] = 0.0;
}

// Frozen limiter capture (beyond the reference): per interior face (taken once, from the block that owns
// its owner cell), the ratios phi(a, b) / a of both sides and the four primitive variables at the
// state U and its gradients G, stored at ratios[j * F + face], j = owner rho, u, v, p, neighbour
// rho, u, v, p.
template <int BC, int LIMITER>
__device__ __forceinline__ void limiter_capture_body(const DevMesh& m, int b, const double* __restrict__ U,
                                                     const double* __restrict__ G, double g,
                                                     double* __restrict__ ratios) {
  const int s0 = m.slot_off[b];
  const int ns = m.slot_off[b + 1] - s0;
  const int ld = m.ld, F = m.nfaces;
  for (int i = threadIdx.x; i < ns; i += BC) {
    const int s = s0 + i;
    const int nm = m.s_nm[s];
    const int co = m.s_o[s], cn = slot_nb(nm);
    if (cn < 0 || !(slot_meta(nm) & kMetaPrimary)) continue;
    const P4 L = c2p(load_u(U, ld, co), g);
    const P4 Rt = c2p(load_u(U, ld, cn), g);
    const double fx = m.s_fx[s], fy = m.s_fy[s];
    const double dox = fx - m.ccx[co], doy = fy - m.ccy[co];
    const double dnx = fx - m.ccx[cn], dny = fy - m.ccy[cn];
    const double q[4][2] = {{L.r, Rt.r}, {L.u, Rt.u}, {L.v, Rt.v}, {L.p, Rt.p}};
    const int f = m.s_face[s];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double ro, rn;
      limiter_ratio<LIMITER>(q[k][0], q[k][1], G[(2 * k) * ld + co], G[(2 * k + 1) * ld + co], G[(2 * k) * ld + cn],
                             G[(2 * k + 1) * ld + cn], dox, doy, dnx, dny, &ro, &rn);
      ratios[k * F + f] = ro;
      ratios[(4 + k) * F + f] = rn;
    }
  }
}

template <int BC, int LIMITER>
__global__ __launch_bounds__(BC) void k_limiter_capture(DevMesh m, const double* __restrict__ U,
                                                        const double* __restrict__ G, double g,
                                                        double* __restrict__ ratios) {
  const int b = block_of(m);
  if (b >= m.nblk) return;
  limiter_capture_body<BC, LIMITER>(m, b, U, G, g, ratios);
}

// Frozen limiter ratios from face order [8][F] to slot order [S][8] (16-bit, units of 2^-15), once per
// capture or install: the flux kernels then read one coalesced 16-byte record per slot (the faces of a
// block are scattered in reference numbering).
__global__ __launch_bounds__(256) void k_ratio_slots(DevMesh m, double* __restrict__ fr,
                                                     unsigned short* __restrict__ out) {
  // Also brings ratios installed from outside (aero_hip_set_limiter_ratios) to their 2^-15 resolution in
  // the face-order copy, which the tangent and adjoint kernels read: the face's primary slot writes it.
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= m.nslots) return;
  const int f = m.s_face[s];
  const bool primary = (slot_meta(m.s_nm[s]) & kMetaPrimary) != 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const size_t at = static_cast<size_t>(j) * m.nfaces + f;
    const double q = quantize_ratio(fr[at]);
    out[static_cast<size_t>(s) * 8 + j] = static_cast<unsigned short>(q * kRatioUnits);
    if (primary) fr[at] = q;
  }
}

// Per-face acoustic scaling record of a state, the post-processing diagnostics of the reference driver:
// compute_acoustic_scaling_info (euler_solver.cpp:392-416) plus the dissipation row's speeds and
// all-speed scale (:1412-1437).  The exterior state is the interior one at a slip wall and the freestream
// at a farfield (:1270-1279); cell states are taken as stored, since every quantity converts to
// primitives with the floors.  out[col][face] in the reference's face numbering, columns kFs*.
__global__ __launch_bounds__(256) void k_face_scales(DevMesh m, FaceOpts o, const double* __restrict__ U,
                                                     double* __restrict__ out) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= m.nslots) return;
  const int nm = m.s_nm[s];
  const int meta = slot_meta(nm);
  if (!(meta & kMetaPrimary)) return;
  const int f = m.s_face[s];
  const int cn = slot_nb(nm);
  const bool wall = cn < 0 && (meta & kMetaBcMask) == 2;
  const double g = o.g;
  const double nx = m.s_nx[s], ny = m.s_ny[s];
  const C4 l = load_u(U, m.ld, m.s_o[s]);
  const C4 r = cn >= 0 ? load_u(U, m.ld, cn) : (wall ? l : o.ff);
  const P4 pl = c2p(l, g), pr = c2p(r, g);
  const double unl = pl.u * nx + pl.v * ny, unr = pr.u * nx + pr.v * ny;
  const double al = sound(pl, g), ar = sound(pr, g);
  const double fm = 0.5 * (local_mach(pl, g) + local_mach(pr, g));
  const double beta = face_beta(l, r, o);
  double base = o.precond ? dclamp(beta, 1.0e-4, 1.0) : 1.0;
  const double lms = o.low_mach_fix ? low_mach_scale(l, r, o) : 1.0;
  base = dclamp(dmin(base, lms), 1.0e-4, 1.0);
  const double stab = dclamp(smooth_max(fm, dclamp(o.stab_floor, 0.0, 1.0)), 1.0e-4, 1.0);
  const double sc = dmax(base, stab);
  const size_t F = static_cast<size_t>(m.nfaces);
  out[kFsMach * F + f] = fm;
  out[kFsUnAbs * F + f] = 0.5 * (fabs(unl) + fabs(unr));
  out[kFsSound * F + f] = 0.5 * (al + ar);
  out[kFsBeta * F + f] = beta;
  out[kFsLowMach * F + f] = lms;
  out[kFsAcousticBase * F + f] = base;
  out[kFsStab * F + f] = stab;
  out[kFsAcoustic * F + f] = sc;
  out[kFsAllSpeed * F + f] = all_speed_scale(pl, pr, nx, ny, o);
  out[kFsSmaxPhysical * F + f] = dmax(fabs(unl) + al, fabs(unr) + ar);
  out[kFsSmaxDissipation * F + f] = dmax(pspeeds(unl, al, sc).sr, pspeeds(unr, ar, sc).sr);
  // mass component of the slip-wall flux evaluated at the cell state (euler_solver.cpp:1312-1324)
  out[kFsWallMassFlux * F + f] = wall ? boundary_face<0>(l, true, nx, ny, o).f0 : 0.0;
}

// Per-cell preconditioner beta and local Mach (euler_solver.cpp:1229-1246), out[2][N] in internal order.
__global__ __launch_bounds__(256) void k_cell_scales(DevMesh m, FaceOpts o, const double* __restrict__ U,
                                                     double* __restrict__ out) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= m.N) return;
  const P4 p = c2p(load_u(U, m.ld, c), o.g);
  out[c] = o.precond ? cell_beta(p, o) : 1.0;
  out[static_cast<size_t>(m.N) + c] = local_mach(p, o.g);
}

// Layout conversions between the caller's device arrays (reference layout: [N][5] conserved in the
// reference cell numbering, w dropped / written as 0) and the internal SoA [4][N]; perm maps an internal
// cell to its reference index.
__global__ __launch_bounds__(256) void k_aos_to_internal(int N, const int* __restrict__ perm,
                                                         const double* __restrict__ aos5, double* __restrict__ soa) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= N) return;
  const double* u = aos5 + 5 * static_cast<size_t>(perm[c]);
  soa[c] = u[0];
  soa[static_cast<size_t>(N) + c] = u[1];
  soa[2 * static_cast<size_t>(N) + c] = u[2];
  soa[3 * static_cast<size_t>(N) + c] = u[4];
}

__global__ __launch_bounds__(256) void k_internal_to_aos(int N, const int* __restrict__ perm,
                                                         const double* __restrict__ soa, double* __restrict__ aos5,
                                                         const double* __restrict__ lam, double* __restrict__ lam_ref) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= N) return;
  double* o = aos5 + 5 * static_cast<size_t>(perm[c]);
  o[0] = soa[c];
  o[1] = soa[static_cast<size_t>(N) + c];
  o[2] = soa[2 * static_cast<size_t>(N) + c];
  o[3] = 0.0;
  o[4] = soa[3 * static_cast<size_t>(N) + c];
  if (lam_ref) lam_ref[perm[c]] = lam[c];
}

__global__ void k_min_dt(int N, const double* __restrict__ lam, double cfl, double* __restrict__ out) {
  __shared__ double red[256];
  double v = 1.79769313486231570815e+308;
  for (int c = threadIdx.x; c < N; c += blockDim.x) v = dmin(v, cfl / dmax(lam[c], 1.0e-8));
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = dmin(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = fin(red[0]) ? red[0] : 0.0;
}

__global__ void k_update_global(DevMesh m, FaceOpts o, const double* __restrict__ U, const double* __restrict__ R,
                                const double* __restrict__ dtp, int* __restrict__ diag, StageArgs st) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m.N) return;
  const int ld = m.ld;
  const double g = o.g;
  const C4 us = load_u(U, ld, c);
  const double r0 = R[c], r1 = R[ld + c], r2 = R[2 * ld + c], r3 = R[3 * ld + c];
  double p3 = r3;
  if (o.precond) {
    const P4 wp = c2p(us, g);
    const double bt = cell_beta(wp, o);
    const double b2 = bt * bt;
    const double rho = dmax(wp.r, 1.0e-8);
    const double inv = 1.0 / rho;
    const double udr = wp.u * r1 + wp.v * r2;
    p3 = b2 * r3 + (b2 - 1.0) * inv * udr;
  }
  const double dt = dtp[0];
  C4 s;
  s.r = us.r - dt * r0;
  s.ru = us.ru - dt * r1;
  s.rv = us.rv - dt * r2;
  s.E = us.E - dt * p3;
  if (st.stage >= 1) {
    const C4 un = load_u(st.Un, ld, c);
    if (st.stage == 1) {
      s.r = 0.75 * un.r + 0.25 * s.r;
      s.ru = 0.75 * un.ru + 0.25 * s.ru;
      s.rv = 0.75 * un.rv + 0.25 * s.rv;
      s.E = 0.75 * un.E + 0.25 * s.E;
    } else {
      const double t1 = 1.0 / 3.0, t2 = 2.0 / 3.0;
      s.r = t1 * un.r + t2 * s.r;
      s.ru = t1 * un.ru + t2 * s.ru;
      s.rv = t1 * un.rv + t2 * s.rv;
      s.E = t1 * un.E + t2 * s.E;
    }
  }
  enforce_physical(&s, g);
  st.Unext[c] = s.r;
  st.Unext[ld + c] = s.ru;
  st.Unext[2 * ld + c] = s.rv;
  st.Unext[3 * ld + c] = s.E;
  if (st.last_stage) {
    bool bad = !(fin(s.r) && fin(s.ru) && fin(s.rv) && fin(s.E));
    if (!bad) {
      const P4 q = c2p(s, g);
      bad = !(fin(q.r) && fin(q.u) && fin(q.v) && fin(q.p));
    }
    if (bad) atomicOr(&diag[3], 1);
  }
}

// Norm partials from a stored residual (used with the non-LTS path), one workgroup per block.
template <int BC>
__global__ __launch_bounds__(BC) void k_norm_partials(DevMesh m, const double* __restrict__ R, double* __restrict__ partials) {
  __shared__ double red[4 * (BC / 64)];
  const int b = blockIdx.x;
  const int c = b * BC + threadIdx.x;
  const int ld = m.ld;
  double nl1 = 0.0, nl2 = 0.0, nlinf = 0.0, nr2 = 0.0;
  if (c < m.N) {
    const double iv = 1.0 / dmax(m.vol[c], 1.0e-12);
    const double v0 = fabs(R[c] * iv), v1 = fabs(R[ld + c] * iv), v2 = fabs(R[2 * ld + c] * iv),
                 v3 = fabs(R[3 * ld + c] * iv);
    nl1 = v0 + v1 + v2 + v3;
    nl2 = v0 * v0 + v1 * v1 + v2 * v2 + v3 * v3;
    nlinf = dmax(dmax(v0, v1), dmax(v2, v3));
    nr2 = v0 * v0;
  }
  for (int off = 32; off > 0; off >>= 1) {
    nl1 += __shfl_xor(nl1, off, 64);
    nl2 += __shfl_xor(nl2, off, 64);
    nlinf = dmax(nlinf, __shfl_xor(nlinf, off, 64));
    nr2 += __shfl_xor(nr2, off, 64);
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[4 * wv + 0] = nl1;
    red[4 * wv + 1] = nl2;
    red[4 * wv + 2] = nlinf;
    red[4 * wv + 3] = nr2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    for (int k = 0; k < BC / 64; ++k) {
      a0 += red[4 * k + 0];
      a1 += red[4 * k + 1];
      a2 = dmax(a2, red[4 * k + 2]);
      a3 += red[4 * k + 3];
    }
    partials[4 * b + 0] = a0;
    partials[4 * b + 1] = a1;
    partials[4 * b + 2] = a2;
    partials[4 * b + 3] = a3;
  }
}

// Final fixed-order reduction of per-block norm partials into out[0..3] = (sum|v|, sum v^2, max|v|,
// sum v_rho^2), then wall pressure-force sums into out[4..8]:
//   sum p nA_x, sum p nA_y (absolute), sum (p - p_inf) nA_x, sum (p - p_inf) nA_y, sum moment term,
// with nA = -n A (forces.cpp:134-149, :220-246, :274-285).  One workgroup of 256 threads.
__device__ __forceinline__ void iter_reduce_body(int nblk, const double* __restrict__ partials, const WallArgs& wa,
                                                 const double* __restrict__ U, int ld, double g,
                                                 double* __restrict__ out, double* __restrict__ cp,
                                                 double (*red)[256]) {
  const int t = threadIdx.x;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  for (int b = t; b < nblk; b += 256) {
    a0 += partials[4 * b + 0];
    a1 += partials[4 * b + 1];
    a2 = dmax(a2, partials[4 * b + 2]);
    a3 += partials[4 * b + 3];
  }
  double f0 = 0.0, f1 = 0.0, f2 = 0.0, f3 = 0.0, f4 = 0.0;
  for (int i = t; i < wa.n; i += 256) {
    const int c = wa.owner[i];
    const C4 s = load_u(U, ld, c);
    const double p = c2p(s, g).p;
    const double A = wa.A[i];
    const double nA0 = -wa.nx[i] * A;
    const double nA1 = -wa.ny[i] * A;
    const double gp = p - wa.p_inf;
    f0 += p * nA0;
    f1 += p * nA1;
    f2 += gp * nA0;
    f3 += gp * nA1;
    const double dfx = -gp * nA0;
    const double dfy = -gp * nA1;
    const double x = wa.fx[i] - wa.x_ref;
    const double y = wa.fy[i] - wa.y_ref;
    f4 += x * dfy - y * dfx;
    if (cp) cp[i] = gp / wa.q_inf;
  }
  red[0][t] = a0;
  red[1][t] = a1;
  red[2][t] = a2;
  red[3][t] = a3;
  red[4][t] = f0;
  red[5][t] = f1;
  red[6][t] = f2;
  red[7][t] = f3;
  red[8][t] = f4;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      for (int k = 0; k < 9; ++k) {
        if (k == 2)
          red[k][t] = dmax(red[k][t], red[k][t + w]);
        else
          red[k][t] += red[k][t + w];
      }
    }
    __syncthreads();
  }
  if (t < 9) out[t] = red[t][0];
}

__global__ __launch_bounds__(256) void k_iter_reduce(int nblk, const double* __restrict__ partials, WallArgs wa,
                                                      const double* __restrict__ U, int ld, double g,
                                                      double* __restrict__ out, double* __restrict__ cp) {
  __shared__ double red[9][256];
  iter_reduce_body(nblk, partials, wa, U, ld, g, out, cp, red);
}

// ---- batched kernels: one workgroup per (case, block) work item --------------------------------
// A case's state-ring buffer for a selector sel (kRingParity / kRingParityNext: aero_internal.hpp).
__device__ __forceinline__ int ring_buf(int sel, int iter) {
  return sel >= 0 ? sel : (sel == kRingParity ? (iter & 1) : 1 - (iter & 1));
}

template <int BC, int GK>
__global__ __launch_bounds__(BC) void k_gradient_batch(const DevMesh* __restrict__ geoms,
                                                        const BatchCase* __restrict__ cases,
                                                        const CaseCtl* __restrict__ ctl, const int2* __restrict__ work,
                                                        double g, int cur) {
  extern __shared__ double lds[];
  const int2 w = work[blockIdx.x];
  if (w.x < 0 || !ctl[w.x].active) return;
  const BatchCase& cs = cases[w.x];
  gradient_body<BC, GK>(geoms[cs.geom], w.y, cs.U[ring_buf(cur, ctl[w.x].iter)], cs.G, g, lds);
}

template <int BC, int SCHEME, int LIMITER, bool SECOND, bool FROZEN = false, bool PLAIN = false>
__global__ __launch_bounds__(BC, AERO_FLUX_WAVES_PER_EU) void k_flux_batch(const DevMesh* __restrict__ geoms,
                                                    const BatchCase* __restrict__ cases,
                                                    const CaseCtl* __restrict__ ctl, const int2* __restrict__ work,
                                                    BatchStage bs) {
  extern __shared__ double lds[];
  const int2 w = work[blockIdx.x];
  if (w.x < 0) return;
  const CaseCtl& k = ctl[w.x];
  if (!k.active) return;
  const BatchCase& cs = cases[w.x];
  FaceOpts o = cs.fo;
  o.stab_floor = k.floor_cur;
  o.as_w = k.as_w;
  o.frozen = FROZEN ? cs.lim_ratio : nullptr;
  o.frozen_slots = FROZEN ? cs.lim_slot : nullptr;
  StageArgs st;
  st.stage = bs.stage;
  st.last_stage = bs.last;
  st.write_res = 0;
  st.cfl = k.cfl;
  st.Un = cs.U[ring_buf(bs.base, k.iter)];
  st.Unext = cs.U[ring_buf(bs.out, k.iter)];
  st.partials = cs.partials;
  flux_body<BC, SCHEME, LIMITER, SECOND, 1, FROZEN, PLAIN>(geoms[cs.geom], w.y, o, cs.U[ring_buf(bs.cur, k.iter)],
                                                           cs.G, nullptr, nullptr, cs.diag, st, lds);
}

template <int BC, int LIMITER>
__global__ __launch_bounds__(BC) void k_limiter_capture_batch(const DevMesh* __restrict__ geoms,
                                                              const BatchCase* __restrict__ cases,
                                                              const CaseCtl* __restrict__ ctl,
                                                              const int2* __restrict__ work, double g, int cur) {
  const int2 w = work[blockIdx.x];
  if (w.x < 0 || !ctl[w.x].active) return;
  const BatchCase& cs = cases[w.x];
  limiter_capture_body<BC, LIMITER>(geoms[cs.geom], w.y, cs.U[ring_buf(cur, ctl[w.x].iter)], cs.G, g, cs.lim_ratio);
}

// Per case: norm partials + wall force sums of the updated state into ctl[case].sums (same fixed-order
// reduction as the single-case path, so a case's history is identical to a solo run).
// The per-iteration control kernels act on the cases of a list (one case group of the batch).
__global__ __launch_bounds__(256) void k_batch_reduce(const BatchCase* __restrict__ cases, CaseCtl* __restrict__ ctl,
                                                       const int* __restrict__ list, double g, int cur) {
  __shared__ double red[9][256];
  const int c = list[blockIdx.x];
  if (!ctl[c].active) return;
  const BatchCase& cs = cases[c];
  iter_reduce_body(cs.nblk, cs.partials, cs.wall, cs.U[ring_buf(cur, ctl[c].iter)], cs.N, g, ctl[c].sums, nullptr,
                   red);
}

__global__ void k_batch_init_diag(const BatchCase* __restrict__ cases, const int* __restrict__ list, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int* d = cases[list[i]].diag;
  d[0] = 0;
  d[1] = 0;
  d[2] = 0x7fffffff;
  d[3] = 0;
}

// The driver's per-iteration scalar logic on the device, one thread per case: norms (euler_solver.cpp:309-339),
// residual ratio (convergence_monitor.cpp:13-19), adaptive CFL (euler_solver.cpp:951-958), then the CFL,
// stabilisation-floor and all-speed ramps of the next iteration (:69-114).  Convergence is the density
// residual drop rho_l2 <= rho_drop * rho_l2(iteration 0).
__global__ void k_batch_control(const BatchCase* __restrict__ cases, CaseCtl* __restrict__ ctl,
                                const int* __restrict__ list, int n, BatchParams p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = list[i];
  CaseCtl& k = ctl[c];
  if (!k.active) return;
  const BatchCase& cs = cases[c];
  if (cs.diag[3]) {
    k.failed = 1;
    k.active = 0;
    return;
  }
  const double N5 = static_cast<double>(cs.N * 5 > 1 ? cs.N * 5 : 1);
  k.l1 = k.sums[0];
  k.l2 = sqrt(k.sums[1] / N5);
  k.linf = k.sums[2];
  k.rho_l2 = sqrt(k.sums[3] / static_cast<double>(cs.N > 1 ? cs.N : 1));
  const double b2 = dmax(k.l2, 1.0e-20);
  if (k.l2_0 < 0.0) k.l2_0 = b2;
  k.ratio = b2 / dmax(k.l2_0, 1.0e-20);
  const double r2 = dmax(k.rho_l2, 1.0e-300);
  if (k.rho_l2_0 < 0.0) k.rho_l2_0 = r2;
  k.rho_ratio = r2 / k.rho_l2_0;
  if (k.prev_l2 > 0.0) {
    if (k.l2 > k.prev_l2 * 1.05)
      k.cfl_scale = dmax(k.cfl_scale * 0.65, 0.05);
    else if (k.l2 < k.prev_l2 * 0.90)
      k.cfl_scale = dmin(k.cfl_scale * 1.03, 1.0);
  }
  k.prev_l2 = k.l2;
  if (p.hist && k.iter < p.hist_cap) {
    double* h = p.hist + (static_cast<long long>(c) * p.hist_cap + k.iter) * 4;
    h[0] = k.l2;
    h[1] = k.rho_l2;
    h[2] = k.cfl;
    h[3] = k.linf;
  }
  k.iter += 1;
  if (!p.fixed_iters && k.rho_ratio <= p.rho_drop) k.converged = 1;
  // The next iteration's parameters are set even when the case stops here, so that a later run resumes
  // it exactly where it left off.
  if (k.converged || k.iter >= p.max_iters) k.active = 0;
  const int it = k.iter;
  double base;
  if (cs.cfl_ramp <= 0) {
    base = cs.cfl_max;
  } else {
    const double t = dmin(1.0, static_cast<double>(it + 1) / static_cast<double>(cs.cfl_ramp > 1 ? cs.cfl_ramp : 1));
    base = cs.cfl_start + t * (cs.cfl_max - cs.cfl_start);
  }
  k.cfl = base * k.cfl_scale;
  if (cs.floor_ramp <= 0) {
    k.floor_cur = cs.floor_target;
  } else {
    const double x = dclamp(static_cast<double>(it) / static_cast<double>(cs.floor_ramp > 1 ? cs.floor_ramp : 1), 0.0, 1.0);
    k.floor_cur = cs.floor_start + x * (cs.floor_target - cs.floor_start);
  }
  if (cs.as_len <= 0)
    k.as_w = it >= cs.as_on ? 1.0 : 0.0;
  else
    k.as_w = dclamp(static_cast<double>(it - cs.as_on) / static_cast<double>(cs.as_len > 1 ? cs.as_len : 1), 0.0, 1.0);
}

}  // namespace aero_dev

// ------------------------------------------------------------------------------------------------
// Launchers (host side of this translation unit)
// ------------------------------------------------------------------------------------------------
namespace aero_host {

using namespace aero_dev;

template <int BC>
static hipError_t launch_gradient_bc(const DevMesh& m, const double* U, double* G, double g, int* diag,
                                     hipStream_t s) {
  const size_t lds = static_cast<size_t>(m.smax) * grad_lds_per_slot(m.grad) * sizeof(double);
  if (m.grad == 1)
    hipLaunchKernelGGL((k_gradient<BC, 1>), dim3(m.grid), dim3(BC), lds, s, m, U, G, g, diag);
  else
    hipLaunchKernelGGL((k_gradient<BC, 0>), dim3(m.grid), dim3(BC), lds, s, m, U, G, g, diag);
  return hipGetLastError();
}

hipError_t launch_gradient(const DevMesh& m, const double* U, double* G, double g, hipStream_t s, int* diag) {
  switch (m.BC) {
    case 128: return launch_gradient_bc<128>(m, U, G, g, diag, s);
    case 256: return launch_gradient_bc<256>(m, U, G, g, diag, s);
    case 512: return launch_gradient_bc<512>(m, U, G, g, diag, s);
    default: return hipErrorInvalidValue;
  }
}

template <int BC, int SCHEME, int LIMITER, bool SECOND, int MODE, bool FROZEN = false>
static hipError_t launch_flux_t(const DevMesh& m, const FaceOpts& o, const double* U, const double* G, double* R,
                                double* lam, int* diag, const StageArgs& st, hipStream_t s) {
  // The fused stage tail reuses the slot staging area for 4 doubles per wavefront.
  const size_t lds = static_cast<size_t>(std::max(m.smax * 5, 4 * (BC / 64))) * sizeof(double);
  if (plain_options(o))
    hipLaunchKernelGGL((k_flux<BC, SCHEME, LIMITER, SECOND, MODE, FROZEN, true>), dim3(m.grid), dim3(BC), lds, s, m, o,
                       U, G, R, lam, diag, st);
  else
    hipLaunchKernelGGL((k_flux<BC, SCHEME, LIMITER, SECOND, MODE, FROZEN, false>), dim3(m.grid), dim3(BC), lds, s, m,
                       o, U, G, R, lam, diag, st);
  return hipGetLastError();
}

template <int BC, int MODE>
static hipError_t launch_flux_bc(const DevMesh& m, const FaceOpts& o, int scheme, int limiter, bool second,
                                 const double* U, const double* G, double* R, double* lam, int* diag,
                                 const StageArgs& st, hipStream_t s) {
  if (second && o.frozen) {  // the frozen ratios replace the limiter
    if (scheme == 0) return launch_flux_t<BC, 0, 0, true, MODE, true>(m, o, U, G, R, lam, diag, st, s);
    return launch_flux_t<BC, 1, 0, true, MODE, true>(m, o, U, G, R, lam, diag, st, s);
  }
#define AERO_FLUX_CASE(SC, LI, SE) \
  if (scheme == SC && limiter == LI && second == SE) return launch_flux_t<BC, SC, LI, SE, MODE>(m, o, U, G, R, lam, diag, st, s);
  AERO_FLUX_CASE(0, 0, false)
  AERO_FLUX_CASE(1, 0, false)
  AERO_FLUX_CASE(0, 0, true)
  AERO_FLUX_CASE(0, 1, true)
  AERO_FLUX_CASE(1, 0, true)
  AERO_FLUX_CASE(1, 1, true)
#undef AERO_FLUX_CASE
  if (!second) return launch_flux_t<BC, 0, 0, false, MODE>(m, o, U, G, R, lam, diag, st, s);  // limiter unused
  return hipErrorInvalidValue;
}

hipError_t launch_limiter_capture(const DevMesh& m, int limiter, const double* U, const double* G, double g,
                                  double* ratios, hipStream_t s) {
#define AERO_CAP(BCV)                                                                                          \
  if (m.BC == BCV) {                                                                                           \
    if (limiter == 1)                                                                                          \
      hipLaunchKernelGGL((k_limiter_capture<BCV, 1>), dim3(m.grid), dim3(BCV), 0, s, m, U, G, g, ratios);      \
    else                                                                                                       \
      hipLaunchKernelGGL((k_limiter_capture<BCV, 0>), dim3(m.grid), dim3(BCV), 0, s, m, U, G, g, ratios);      \
    return hipGetLastError();                                                                                  \
  }
  AERO_CAP(128)
  AERO_CAP(256)
  AERO_CAP(512)
#undef AERO_CAP
  return hipErrorInvalidValue;
}

hipError_t launch_flux(const DevMesh& m, const FaceOpts& o, int scheme, int limiter, bool second, int mode,
                       const double* U, const double* G, double* R, double* lam, int* diag, const StageArgs& st,
                       hipStream_t s) {
  if (!second && scheme == 1) {
    if (mode == 0) {
      switch (m.BC) {
        case 128: return launch_flux_t<128, 1, 0, false, 0>(m, o, U, G, R, lam, diag, st, s);
        case 256: return launch_flux_t<256, 1, 0, false, 0>(m, o, U, G, R, lam, diag, st, s);
        case 512: return launch_flux_t<512, 1, 0, false, 0>(m, o, U, G, R, lam, diag, st, s);
      }
    } else {
      switch (m.BC) {
        case 128: return launch_flux_t<128, 1, 0, false, 1>(m, o, U, G, R, lam, diag, st, s);
        case 256: return launch_flux_t<256, 1, 0, false, 1>(m, o, U, G, R, lam, diag, st, s);
        case 512: return launch_flux_t<512, 1, 0, false, 1>(m, o, U, G, R, lam, diag, st, s);
      }
    }
    return hipErrorInvalidValue;
  }
  switch (m.BC) {
    case 128: return mode == 0 ? launch_flux_bc<128, 0>(m, o, scheme, limiter, second, U, G, R, lam, diag, st, s)
                               : launch_flux_bc<128, 1>(m, o, scheme, limiter, second, U, G, R, lam, diag, st, s);
    case 256: return mode == 0 ? launch_flux_bc<256, 0>(m, o, scheme, limiter, second, U, G, R, lam, diag, st, s)
                               : launch_flux_bc<256, 1>(m, o, scheme, limiter, second, U, G, R, lam, diag, st, s);
    case 512: return mode == 0 ? launch_flux_bc<512, 0>(m, o, scheme, limiter, second, U, G, R, lam, diag, st, s)
                               : launch_flux_bc<512, 1>(m, o, scheme, limiter, second, U, G, R, lam, diag, st, s);
    default: return hipErrorInvalidValue;
  }
}

template <int SCHEME, int LIMITER, bool SECOND, bool FROZEN>
static hipError_t launch_colours_t(const DevMesh& m, const ColourFaces& cf, const int* off, int nc, const FaceOpts& o,
                                   const double* U, const double* G, double* R, double* lam, int* diag,
                                   hipStream_t s) {
  for (int c = 0; c < nc; ++c) {
    const int n = off[c + 1] - off[c];
    if (n <= 0) continue;
    hipLaunchKernelGGL((k_flux_colour<SCHEME, LIMITER, SECOND, FROZEN>), dim3((n + 255) / 256), dim3(256), 0, s, m, cf,
                       off[c], off[c + 1], o, U, G, R, lam, diag);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_flux_colour(const DevMesh& m, const ColourFaces& cf, const int* off, int nc, const FaceOpts& o,
                              int scheme, int limiter, bool second, const double* U, const double* G, double* R,
                              double* lam, int* diag, hipStream_t s) {
  const long long nr = 4LL * m.N;
  hipLaunchKernelGGL(k_zero, dim3(static_cast<unsigned>((nr + 255) / 256)), dim3(256), 0, s, nr, R);
  hipLaunchKernelGGL(k_zero, dim3(static_cast<unsigned>((m.N + 255) / 256)), dim3(256), 0, s,
                     static_cast<long long>(m.N), lam);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (second && o.frozen) {
    if (scheme == 0) return launch_colours_t<0, 0, true, true>(m, cf, off, nc, o, U, G, R, lam, diag, s);
    return launch_colours_t<1, 0, true, true>(m, cf, off, nc, o, U, G, R, lam, diag, s);
  }
  if (!second) {
    if (scheme == 0) return launch_colours_t<0, 0, false, false>(m, cf, off, nc, o, U, G, R, lam, diag, s);
    return launch_colours_t<1, 0, false, false>(m, cf, off, nc, o, U, G, R, lam, diag, s);
  }
  if (scheme == 0 && limiter == 0) return launch_colours_t<0, 0, true, false>(m, cf, off, nc, o, U, G, R, lam, diag, s);
  if (scheme == 0 && limiter == 1) return launch_colours_t<0, 1, true, false>(m, cf, off, nc, o, U, G, R, lam, diag, s);
  if (scheme == 1 && limiter == 0) return launch_colours_t<1, 0, true, false>(m, cf, off, nc, o, U, G, R, lam, diag, s);
  if (scheme == 1 && limiter == 1) return launch_colours_t<1, 1, true, false>(m, cf, off, nc, o, U, G, R, lam, diag, s);
  return hipErrorInvalidValue;
}

hipError_t launch_face_scales(const DevMesh& m, const FaceOpts& o, const double* U, double* faces, double* cells,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_face_scales, dim3((m.nslots + 255) / 256), dim3(256), 0, s, m, o, U, faces);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_cell_scales, dim3((m.N + 255) / 256), dim3(256), 0, s, m, o, U, cells);
  return hipGetLastError();
}

hipError_t launch_aos_to_internal(int N, const int* perm, const double* aos5, double* soa, hipStream_t s) {
  hipLaunchKernelGGL(k_aos_to_internal, dim3((N + 255) / 256), dim3(256), 0, s, N, perm, aos5, soa);
  return hipGetLastError();
}

hipError_t launch_internal_to_aos(int N, const int* perm, const double* soa, double* aos5, const double* lam,
                                  double* lam_ref, hipStream_t s) {
  hipLaunchKernelGGL(k_internal_to_aos, dim3((N + 255) / 256), dim3(256), 0, s, N, perm, soa, aos5, lam, lam_ref);
  return hipGetLastError();
}

hipError_t launch_ratio_slots(const DevMesh& m, double* ratios, unsigned short* slots, hipStream_t s) {
  hipLaunchKernelGGL(k_ratio_slots, dim3((m.nslots + 255) / 256), dim3(256), 0, s, m, ratios, slots);
  return hipGetLastError();
}

hipError_t launch_global_update(const DevMesh& m, const FaceOpts& o, const double* U, const double* R,
                                const double* lam, double cfl, double* dt_scratch, int* diag, const StageArgs& st,
                                hipStream_t s) {
  hipLaunchKernelGGL(k_min_dt, dim3(1), dim3(256), 0, s, m.N, lam, cfl, dt_scratch);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_update_global, dim3((m.N + 255) / 256), dim3(256), 0, s, m, o, U, R, dt_scratch, diag, st);
  return hipGetLastError();
}

hipError_t launch_norm_partials(const DevMesh& m, const double* R, double* partials, hipStream_t s) {
  switch (m.BC) {
    case 128: hipLaunchKernelGGL(k_norm_partials<128>, dim3(m.nblk), dim3(128), 0, s, m, R, partials); break;
    case 256: hipLaunchKernelGGL(k_norm_partials<256>, dim3(m.nblk), dim3(256), 0, s, m, R, partials); break;
    case 512: hipLaunchKernelGGL(k_norm_partials<512>, dim3(m.nblk), dim3(512), 0, s, m, R, partials); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int GK>
static hipError_t launch_batch_gradient_gk(int BC, const DevMesh* geoms, const BatchCase* cases, const CaseCtl* ctl,
                                           const int2* work, int nwork, int smax, double g, int cur, hipStream_t s) {
  const size_t lds = static_cast<size_t>(smax) * grad_lds_per_slot(GK) * sizeof(double);
  switch (BC) {
    case 128: hipLaunchKernelGGL((k_gradient_batch<128, GK>), dim3(nwork), dim3(128), lds, s, geoms, cases, ctl, work, g, cur); break;
    case 256: hipLaunchKernelGGL((k_gradient_batch<256, GK>), dim3(nwork), dim3(256), lds, s, geoms, cases, ctl, work, g, cur); break;
    case 512: hipLaunchKernelGGL((k_gradient_batch<512, GK>), dim3(nwork), dim3(512), lds, s, geoms, cases, ctl, work, g, cur); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_batch_gradient(int BC, int grad, const DevMesh* geoms, const BatchCase* cases, const CaseCtl* ctl,
                                 const int2* work, int nwork, int smax, double g, int cur, hipStream_t s) {
  if (grad == 1) return launch_batch_gradient_gk<1>(BC, geoms, cases, ctl, work, nwork, smax, g, cur, s);
  return launch_batch_gradient_gk<0>(BC, geoms, cases, ctl, work, nwork, smax, g, cur, s);
}

template <int BC, bool PLAIN>
static hipError_t launch_batch_flux_bc(int scheme, int limiter, bool second, bool frozen, const DevMesh* geoms,
                                       const BatchCase* cases, const CaseCtl* ctl, const int2* work, int nwork,
                                       size_t lds, const BatchStage& bs, hipStream_t s) {
  if (second && frozen) {  // the frozen ratios replace the limiter
    if (scheme == 0)
      hipLaunchKernelGGL((k_flux_batch<BC, 0, 0, true, true, PLAIN>), dim3(nwork), dim3(BC), lds, s, geoms, cases, ctl,
                         work, bs);
    else
      hipLaunchKernelGGL((k_flux_batch<BC, 1, 0, true, true, PLAIN>), dim3(nwork), dim3(BC), lds, s, geoms, cases, ctl,
                         work, bs);
    return hipGetLastError();
  }
#define AERO_BFLUX(SC, LI, SE)                                                                              \
  if (scheme == SC && (limiter == LI || !SE) && second == SE) {                                             \
    hipLaunchKernelGGL((k_flux_batch<BC, SC, LI, SE, false, PLAIN>), dim3(nwork), dim3(BC), lds, s, geoms, cases, ctl, \
                       work, bs);                                                                           \
    return hipGetLastError();                                                                               \
  }
  AERO_BFLUX(0, 0, false)
  AERO_BFLUX(1, 0, false)
  AERO_BFLUX(0, 0, true)
  AERO_BFLUX(0, 1, true)
  AERO_BFLUX(1, 0, true)
  AERO_BFLUX(1, 1, true)
#undef AERO_BFLUX
  return hipErrorInvalidValue;
}

hipError_t launch_batch_flux(int BC, int scheme, int limiter, bool second, bool frozen, bool plain,
                             const DevMesh* geoms, const BatchCase* cases, const CaseCtl* ctl, const int2* work,
                             int nwork, int smax, const BatchStage& bs, hipStream_t s) {
  const size_t lds = std::max<size_t>(static_cast<size_t>(smax) * 5, 4 * (BC / 64)) * sizeof(double);
#define AERO_BFLUX_BC(BCV)                                                                                         \
  case BCV:                                                                                                        \
    return plain ? launch_batch_flux_bc<BCV, true>(scheme, limiter, second, frozen, geoms, cases, ctl, work, nwork, \
                                                   lds, bs, s)                                                     \
                 : launch_batch_flux_bc<BCV, false>(scheme, limiter, second, frozen, geoms, cases, ctl, work, nwork, \
                                                    lds, bs, s);
  switch (BC) {
    AERO_BFLUX_BC(128)
    AERO_BFLUX_BC(256)
    AERO_BFLUX_BC(512)
    default: return hipErrorInvalidValue;
  }
#undef AERO_BFLUX_BC
}

hipError_t launch_batch_limiter_capture(int BC, int limiter, const DevMesh* geoms, const BatchCase* cases,
                                        const CaseCtl* ctl, const int2* work, int nwork, double g, int cur,
                                        hipStream_t s) {
#define AERO_BCAP(BCV)                                                                                             \
  if (BC == BCV) {                                                                                                 \
    if (limiter == 1)                                                                                              \
      hipLaunchKernelGGL((k_limiter_capture_batch<BCV, 1>), dim3(nwork), dim3(BCV), 0, s, geoms, cases, ctl, work, g, \
                         cur);                                                                                     \
    else                                                                                                           \
      hipLaunchKernelGGL((k_limiter_capture_batch<BCV, 0>), dim3(nwork), dim3(BCV), 0, s, geoms, cases, ctl, work, g, \
                         cur);                                                                                     \
    return hipGetLastError();                                                                                      \
  }
  AERO_BCAP(128)
  AERO_BCAP(256)
  AERO_BCAP(512)
#undef AERO_BCAP
  return hipErrorInvalidValue;
}

hipError_t launch_batch_reduce(const BatchCase* cases, CaseCtl* ctl, const int* list, int n, double g, int cur,
                               hipStream_t s) {
  hipLaunchKernelGGL(k_batch_reduce, dim3(n), dim3(256), 0, s, cases, ctl, list, g, cur);
  return hipGetLastError();
}

hipError_t launch_batch_control(const BatchCase* cases, CaseCtl* ctl, const int* list, int n, const BatchParams& p,
                                hipStream_t s) {
  hipLaunchKernelGGL(k_batch_control, dim3((n + 63) / 64), dim3(64), 0, s, cases, ctl, list, n, p);
  return hipGetLastError();
}

hipError_t launch_batch_init_diag(const BatchCase* cases, const int* list, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_init_diag, dim3((n + 63) / 64), dim3(64), 0, s, cases, list, n);
  return hipGetLastError();
}

hipError_t launch_iter_reduce(const DevMesh& m, const double* partials, const WallArgs& wa, const double* U, double g,
                              double* out, double* cp, hipStream_t s) {
  hipLaunchKernelGGL(k_iter_reduce, dim3(1), dim3(256), 0, s, m.nblk, partials, wa, U, m.ld, g, out, cp);
  return hipGetLastError();
}

}  // namespace aero_host
