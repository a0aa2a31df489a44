// kernels.hip — CDNA4 (gfx950) kernels of the FDTD time step.
//
// The step is HBM-bound stencil arithmetic (no contraction, so no MFMA).  Two kernels
// carry the field updates:
//   k_fast  — the bulk of the cell, where no uPML profile is active: each thread owns 4 (FP32)
//             or 2 (FP64) consecutive points of a padded row and moves them with 16-byte
//             loads and stores; the stride-1 neighbour of the curl is a shift of the same
//             vector plus one extra element.
//   k_step  — the uPML slabs (and anything else): one thread per point, full reference logic.
// Both visit logical blocks in XCD-contiguous runs (hardware deals physical blocks
// round-robin over the 8 XCDs), so the neighbouring rows and planes a block reads were
// fetched into the same XCD's L2 by the blocks just before it.
//
// Per cell and component they fuse, in the reference's order:
//   curl + uPML            step_curl             src/step_generic.cpp:65-249
//   current injection      fields_chunk::step_source  src/step.cpp:295-318
//   metal walls            zero_metal            src/boundaries.cpp:310-313
//   f_minus_p = D - P - p  update_eh             src/update_eh.cpp:114-138
//   E = chi1inv f_minus_p  step_update_EDHB      src/step_generic.cpp:682-701,774-782
//   Lorentz/Drude P        update_P              src/susceptibility.cpp:188-262
#include "engine.hpp"
#include "stencil.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

namespace mx {

// Periodic images of the owned point (j0, j1, j2) along the folded axes: offset of each axis's
// image (0 if none) and its sign.
struct Images {
  int64_t o[3];
  int s[3];
  bool any;
};
__device__ __forceinline__ Images images_of(const Grid &g, const AxisSlot &ax, int j0, int j1, int j2) {
  Images m;
  const int j[3] = {j0, j1, j2};
  m.any = false;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int d = g.wrap[k] ? zwrap_dst(ax.ysh[k], j[k], g.nax[k]) : -1;
    m.o[k] = d >= 0 ? int64_t(d - j[k]) * g.ast[k] : 0;
    m.s[k] = d >= 0 ? g.wrap[k] : 0;
    m.any = m.any || d >= 0;
  }
  return m;
}
template <class R> __device__ __forceinline__ R signed_value(R v, int s) { return s > 0 ? v : -v; }
// write v at i and at every combination of its images
template <class R> __device__ __forceinline__ void put(R *f, uint32_t i, R v, const Grid &g, const AxisSlot &ax, int j0, int j1, int j2) {
  f[i] = v;
  if (!(g.wrap[0] | g.wrap[1] | g.wrap[2])) return;
  const Images m = images_of(g, ax, j0, j1, j2);
  if (!m.any) return;
#pragma unroll
  for (int sub = 1; sub < 8; ++sub) {
    int64_t off = 0;
    int sg = 1;
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if ((sub >> k) & 1) {
        ok = ok && m.s[k] != 0;
        off += m.o[k];
        sg *= m.s[k];
      }
    if (ok) f[int64_t(i) + off] = signed_value(v, sg);
  }
}

// vrow: the step's row of source values
template <class R> __device__ __forceinline__ R apply_tab(const SrcTab &t, int c, uint32_t i, R v, const double *vrow) {
#pragma unroll
  for (int s = 0; s < SRC_TAB; ++s)
    if (s < t.n && t.comp[s] == c && t.idx[s] == int64_t(i)) v = R(double(v) - vrow[t.vbase + s]);
  return v;
}

// value of box source S at loop-axis indices (j0, j1, j2); contraction is off so the complex
// product rounds as the host's does
template <int S> __device__ __forceinline__ double box_value(const SrcBoxes &b, int cmp, int j0, int j1, int j2, const double *vrow) {
#pragma clang fp contract(off)
  const int64_t k = int64_t(j0 - b.lo[S][0]) * b.bst[S][0] + int64_t(j1 - b.lo[S][1]) * b.bst[S][1] +
                    int64_t(j2 - b.lo[S][2]) * b.bst[S][2];
  const double2 am = b.amp[S][k];
  const double sx = vrow[b.vbase + 2 * S], sy = vrow[b.vbase + 2 * S + 1];
  const double re = am.x * sx - am.y * sy, im = am.x * sy + am.y * sx;
  return cmp ? im * b.mult[S] : re * b.mult[S];
}
template <int S> __device__ __forceinline__ bool in_box(const SrcBoxes &b, int c, int j0, int j1, int j2) {
  return S < b.n && b.comp[S] == c && j0 >= b.lo[S][0] && j0 <= b.hi[S][0] && j1 >= b.lo[S][1] && j1 <= b.hi[S][1] &&
         j2 >= b.lo[S][2] && j2 <= b.hi[S][2];
}
template <class R>
__device__ __forceinline__ R apply_boxes(const SrcBoxes &b, int cmp, int c, int j0, int j1, int j2, R v, const double *vrow) {
  static_assert(SRC_BOX == 2, "one term per box");
  if (in_box<0>(b, c, j0, j1, j2)) v = R(double(v) - box_value<0>(b, cmp, j0, j1, j2, vrow));
  if (in_box<1>(b, c, j0, j1, j2)) v = R(double(v) - box_value<1>(b, cmp, j0, j1, j2, vrow));
  return v;
}
// whether the V points from (j0, j1, z0) along the row meet the launch's source box
template <int V, class R> __device__ __forceinline__ bool in_src_box(const StepArgs<R> &a, int j0, int j1, int z0) {
  return j0 >= a.sbox_lo[0] && j0 <= a.sbox_hi[0] && j1 >= a.sbox_lo[1] && j1 <= a.sbox_hi[1] && z0 + V - 1 >= a.sbox_lo[2] &&
         z0 <= a.sbox_hi[2];
}
// currents (dip false) or integrated dipoles (dip true) of slot C at point i: the source set
// is read from device memory only for indices inside the launch's source range
template <class R, bool DIP, int C>
__device__ __forceinline__ R apply_sources(const StepArgs<R> &a, uint32_t i, int j0, int j1, int j2, R v, const double *vrow) {
  const SrcSet &S = *a.src;
  const SrcTab &t = DIP ? S.dip : S.cur;
  const SrcBoxes &b = DIP ? S.dipb : S.curb;
  if (t.n && int64_t(i) >= t.lo && int64_t(i) <= t.hi) v = apply_tab(t, C, i, v, vrow);
  if (b.n) v = apply_boxes(b, a.cmp, C, j0, j1, j2, v, vrow);
  return v;
}

// ---------------------------------------------------------------------------------------
// generic kernel: one point per thread, full uPML logic, over the launch box

template <class R, bool CURL, bool UPD, int C>
__device__ __forceinline__ void cell_update(const StepArgs<R> &a, uint32_t i, int j0, int j1, int j2, R cu, const double *vrow) {
  R fv = R(0);
  if (CURL) {
    const CurlComp<R> &k = a.cc[C];
    fv = k.f[i];
    const int kk = k.ksig >= 0 ? 2 * asel(j0, j1, j2, k.ksig) + k.ysig : 0;
    const int ku = k.ksigu >= 0 ? 2 * asel(j0, j1, j2, k.ksigu) + k.ysigu : 0;
    const bool pml = pml_on(k.sig, k.kap, kk);
    const bool pmlu = pml_on(k.sigu, k.kapu, ku);
    if (k.cnd) {
      // conductivity: the curl term is damped by (1 - dt/2 cnd) / (1 + dt/2 cnd); where uPML
      // acts on f it goes through fcnd, whose increment the sigma update takes
      // (src/step_generic.cpp:85-96,108-128,158-177,198-220)
      // the auxiliaries f_u and fcnd are FP64 arrays, updated in FP64
      const double cn = k.cnd[i], ci = k.cndinv[i], dt2 = k.dt2, dtdx = k.dtdx, cud = cu;
      double fprev = 0.0, g = fv;
      if (pmlu) fprev = g = k.fu[i];
      if (pml) {
        const double fc = k.fcnd[i];
        const double fcn = ((1.0 - dt2 * cn) * fc - dtdx * cud) * ci;
        k.fcnd[i] = fcn;
        g = ((double(k.kap[kk]) - double(k.sig[kk])) * g + (fcn - fc)) * double(k.siginv[kk]);
      }
      else
        g = ((1.0 - dt2 * cn) * g - dtdx * cud) * ci;
      if (pmlu) {
        k.fu[i] = g;
        fv = R(double(k.siginvu[ku]) * ((double(k.kapu[ku]) - double(k.sigu[ku])) * double(fv) + g - fprev));
      }
      else
        fv = R(g);
    }
    else if (!pmlu) {
      if (!pml)
        fv -= k.dtdx * cu;
      else
        fv = ((k.kap[kk] - k.sig[kk]) * fv - k.dtdx * cu) * k.siginv[kk];
    }
    else {
      // the uPML auxiliary f_u is an FP64 array, updated in FP64
      const double fprev = k.fu[i], dtdx = k.dtdx, cud = cu;
      double fun;
      if (!pml)
        fun = fprev - dtdx * cud;
      else
        fun = ((double(k.kap[kk]) - double(k.sig[kk])) * fprev - dtdx * cud) * double(k.siginv[kk]);
      k.fu[i] = fun;
      fv = R(double(k.siginvu[ku]) * ((double(k.kapu[ku]) - double(k.sigu[ku])) * double(fv) + fun - fprev));
    }
    if (a.src && in_src_box<1>(a, j0, j1, j2)) fv = apply_sources<R, false, C>(a, i, j0, j1, j2, fv, vrow);
    if (slot_metal(k.ax, j0, j1, j2)) fv = R(0);
    put(k.f, i, fv, a.g, k.ax, j0, j1, j2);
  }
  if (UPD && a.uc[C].f) {
    const UpdComp<R> &u = a.uc[C];
    R gv = CURL ? fv : u.g[i];
    bool pin[MAXPOLE]; // inside pole p's box
#pragma unroll
    for (int p = 0; p < MAXPOLE; ++p) {
      pin[p] = p < a.npole && u.P[p] && j0 >= a.pole[p].lo[0] && j0 <= a.pole[p].hi[0] && j1 >= a.pole[p].lo[1] &&
               j1 <= a.pole[p].hi[1] && j2 >= a.pole[p].lo[2] && j2 <= a.pole[p].hi[2];
      if (pin[p]) gv -= u.P[p][i];
    }
    if (a.src && in_src_box<1>(a, j0, j1, j2)) gv = apply_sources<R, true, C>(a, i, j0, j1, j2, gv, vrow);
    R ev, wv;
    const int kw = u.ksigw >= 0 ? 2 * asel(j0, j1, j2, u.ksigw) + u.ysigw : 0;
    if (pml_on(u.sigw, u.kapw, kw)) {
      // the uPML auxiliary f_w is an FP64 array, updated in FP64
      const double kapw = u.kapw[kw], sigw = u.sigw[kw];
      const double fwprev = u.fw[i];
      const R fwr = u.u ? gv * u.u[i] : gv;
      const double fwn = fwr;
      u.fw[i] = fwn;
      ev = R(double(u.f[i]) + ((kapw + sigw) * fwn - (kapw - sigw) * fwprev));
      wv = fwr;
    }
    else {
      ev = u.u ? gv * u.u[i] : gv;
      wv = ev;
    }
#pragma unroll
    for (int p = 0; p < MAXPOLE; ++p)
      if (pin[p] && u.sigma[p]) {
        const PoleCoef<R> &q = a.pole[p];
        const R pcur = u.P[p][i];
        const R pn = q.gamma1inv * (pcur * (R(2) - q.omega0dtsqr_denom) - q.gamma1 * u.Pp[p][i] +
                                    q.omega0dtsqr * (u.sigma[p][i] * wv));
        if (a.pswap)
          put(u.Pp[p], i, pn, a.g, u.ax, j0, j1, j2);
        else {
          put(u.P[p], i, pn, a.g, u.ax, j0, j1, j2);
          u.Pp[p][i] = pcur;
        }
      }
    if (slot_metal(u.ax, j0, j1, j2)) ev = R(0);
    put(u.f, i, ev, a.g, u.ax, j0, j1, j2);
  }
}

// one component slot C (a template argument, so every access to the slot's parameters has a
// constant offset)
template <class R, bool CURL, bool UPD, int C>
__device__ __forceinline__ void step_slot(const StepArgs<R> &a, uint32_t i, int j0, int j1, int j2, const double *vrow) {
  if (C >= a.nc) return;
  if (CURL ? !a.cc[C].f : !a.uc[C].f) return; // no component along this direction
  if (!slot_owned(CURL ? a.cc[C].ax : a.uc[C].ax, j0, j1, j2)) return;
  R cu = R(0);
  if (CURL) {
    const CurlComp<R> &k = a.cc[C];
    if (k.g2)
      cu = k.g1[i + k.s1] - k.g1[i] + k.g2[i] - k.g2[i + k.s2];
    else if (k.g1)
      cu = k.g1[i + k.s1] - k.g1[i];
  }
  cell_update<R, CURL, UPD, C>(a, i, j0, j1, j2, cu, vrow);
}

// point t of the launch box of a
template <class R, bool CURL, bool UPD>
__device__ __forceinline__ void step_body(const StepArgs<R> &a, uint32_t t, const double *vrow) {
  const Grid &g = a.g;
  const uint32_t r = t / uint32_t(g.an[2]);
  const int q2 = int(t - r * uint32_t(g.an[2]));
  const int q0 = int(r / uint32_t(g.an[1]));
  const int q1 = int(r - uint32_t(q0) * uint32_t(g.an[1]));
  const int j0 = g.alo[0] + q0, j1 = g.alo[1] + q1, j2 = g.alo[2] + q2;
  const uint32_t i = uint32_t(j0 * g.ast[0] + j1 * g.ast[1] + j2 * g.ast[2]);
  step_slot<R, CURL, UPD, 0>(a, i, j0, j1, j2, vrow);
  step_slot<R, CURL, UPD, 1>(a, i, j0, j1, j2, vrow);
  step_slot<R, CURL, UPD, 2>(a, i, j0, j1, j2, vrow);
}

template <class R, bool CURL, bool UPD>
__global__ __launch_bounds__(BLOCK) void k_step(const StepArgs<R> a, uint32_t count) {
  const uint32_t t = xcd_block() * uint32_t(BLOCK) + threadIdx.x;
  if (t >= count) return;
  step_body<R, CURL, UPD>(a, t, a.vrow);
}

// ---------------------------------------------------------------------------------------
// vectorised kernel: V consecutive points of a padded row per thread, no uPML

template <class R, int V> struct Vec {
  R v[V];
};
// one 16-byte (or, for 8-byte groups, 8-byte) access per group; groups are aligned to their size
template <int B> struct VecWord;
template <> struct VecWord<16> { using T = float4; };
template <> struct VecWord<8> { using T = float2; };
template <class R, int V> __device__ __forceinline__ Vec<R, V> vload(const R *p) {
  using W = typename VecWord<V * sizeof(R)>::T;
  Vec<R, V> r;
  const W q = *reinterpret_cast<const W *>(p);
  __builtin_memcpy(r.v, &q, sizeof(W));
  return r;
}
template <class R, int V> __device__ __forceinline__ void vstore(R *p, const Vec<R, V> &x, unsigned mask) {
  using W = typename VecWord<V * sizeof(R)>::T;
  if (mask == (1u << V) - 1) {
    W q;
    __builtin_memcpy(&q, x.v, sizeof(W));
    *reinterpret_cast<W *>(p) = q;
  }
  else {
#pragma unroll
    for (int k = 0; k < V; ++k)
      if (mask & (1u << k)) p[k] = x.v[k];
  }
}

// the FP64 auxiliary arrays (f_u, f_w) of a group: V doubles in 16-byte pieces (groups are
// aligned to V values, so to 8 V bytes)
template <int V> __device__ __forceinline__ Vec<double, V> vload_aux(const double *p) {
  Vec<double, V> r;
  if (V == 1) {
    r.v[0] = p[0];
    return r;
  }
#pragma unroll
  for (int h = 0; h < V / 2; ++h) {
    const double2 q = reinterpret_cast<const double2 *>(p)[h];
    r.v[2 * h] = q.x;
    r.v[2 * h + 1] = q.y;
  }
  return r;
}
template <int V> __device__ __forceinline__ void vstore_aux(double *p, const Vec<double, V> &x, unsigned mask) {
  if (V > 1 && mask == (1u << V) - 1) {
#pragma unroll
    for (int h = 0; h < V / 2; ++h) reinterpret_cast<double2 *>(p)[h] = make_double2(x.v[2 * h], x.v[2 * h + 1]);
    return;
  }
#pragma unroll
  for (int k = 0; k < V; ++k)
    if (mask & (1u << k)) p[k] = x.v[k];
}

// partner values at i + s for the group starting at i (s = +-1 along the row: a shift of the
// centre vector plus one element that is read only when an owned point needs it)
template <class R, int V>
__device__ __forceinline__ Vec<R, V> neighbour(const R *g, uint32_t i, int64_t s, const Vec<R, V> &c, unsigned mask) {
  if (s == 1) {
    Vec<R, V> r;
#pragma unroll
    for (int k = 0; k < V - 1; ++k) r.v[k] = c.v[k + 1];
    r.v[V - 1] = ((mask >> (V - 1)) & 1) ? g[i + V] : R(0);
    return r;
  }
  if (s == -1) {
    Vec<R, V> r;
#pragma unroll
    for (int k = 1; k < V; ++k) r.v[k] = c.v[k - 1];
    r.v[0] = (mask & 1) ? g[i - 1] : R(0);
    return r;
  }
  return vload<R, V>(g + (int64_t(i) + s));
}

template <class R, int V> __device__ __forceinline__ Vec<R, V> signed_vec(const Vec<R, V> &x, int s) {
  Vec<R, V> r;
#pragma unroll
  for (int k = 0; k < V; ++k) r.v[k] = signed_value(x.v[k], s);
  return r;
}

// store the owned points of a group and their periodic images: whole-row images along loop
// axes 0 and 1 (uniform over the row), single-point images along the row's own axis
template <class R, int V>
__device__ __forceinline__ void put_row(R *f, uint32_t i, const Vec<R, V> &x, unsigned mask, const Grid &g, int ybits, int j0,
                                        int j1, int z0) {
  vstore<R, V>(f + i, x, mask);
  if (!(g.wrap[0] | g.wrap[1] | g.wrap[2])) return;
  const int e0 = g.wrap[0] ? zwrap_dst(ybits & 1, j0, g.nax[0]) : -1;
  const int e1 = g.wrap[1] ? zwrap_dst((ybits >> 1) & 1, j1, g.nax[1]) : -1;
  const int64_t o0 = e0 >= 0 ? int64_t(e0 - j0) * g.ast[0] : 0, o1 = e1 >= 0 ? int64_t(e1 - j1) * g.ast[1] : 0;
  if (e0 >= 0) vstore<R, V>(f + (int64_t(i) + o0), signed_vec(x, g.wrap[0]), mask);
  if (e1 >= 0) vstore<R, V>(f + (int64_t(i) + o1), signed_vec(x, g.wrap[1]), mask);
  if (e0 >= 0 && e1 >= 0) vstore<R, V>(f + (int64_t(i) + o0 + o1), signed_vec(x, g.wrap[0] * g.wrap[1]), mask);
  if (!g.wrap[2]) return;
#pragma unroll
  for (int k = 0; k < V; ++k)
    if ((mask >> k) & 1) {
      const int zd = zwrap_dst((ybits >> 2) & 1, z0 + k, g.nax[2]);
      if (zd < 0) continue;
      const int64_t b = int64_t(i) + k + (zd - (z0 + k));
      const R v = x.v[k];
      f[b] = signed_value(v, g.wrap[2]);
      if (e0 >= 0) f[b + o0] = signed_value(v, g.wrap[2] * g.wrap[0]);
      if (e1 >= 0) f[b + o1] = signed_value(v, g.wrap[2] * g.wrap[1]);
      if (e0 >= 0 && e1 >= 0) f[b + o0 + o1] = signed_value(v, g.wrap[2] * g.wrap[0] * g.wrap[1]);
    }
}

// Owned range and metal planes of a slot along loop axis k, from its Yee-shift bits (the
// reference's owned range [1 - ysh, n - ysh] and on_metal_boundary, src/vec.cpp:432-463,
// src/boundaries.cpp:191-205); derived rather than loaded to keep the kernel's scalar
// registers for the pointers.
__device__ __forceinline__ bool owns_at(const Grid &g, int ybits, int k, int j) {
  const int n = g.nax[k], y = (ybits >> k) & 1;
  return n ? (j >= 1 - y && j <= n - y) : j == 0;
}
__device__ __forceinline__ bool metal_at(const Grid &g, int ybits, int k, int j) {
  const int n = g.nax[k], y = (ybits >> k) & 1;
  return n && ((((g.mhi_bits >> k) & 1) && y == 0 && j == n) || (((g.mlo_bits >> k) & 1) && y == 1 && j == 0));
}
// points of a group that are owned by a slot / on its metal planes
template <int V> __device__ __forceinline__ unsigned group_owned(const Grid &g, int ybits, int j0, int j1, int z0, unsigned boxmask) {
  if (!owns_at(g, ybits, 0, j0) || !owns_at(g, ybits, 1, j1)) return 0u;
  unsigned m = boxmask;
#pragma unroll
  for (int k = 0; k < V; ++k)
    if (!owns_at(g, ybits, 2, z0 + k)) m &= ~(1u << k);
  return m;
}
template <int V> __device__ __forceinline__ unsigned group_metal(const Grid &g, int ybits, int j0, int j1, int z0) {
  if (!(g.mhi_bits | g.mlo_bits)) return 0u;
  if (metal_at(g, ybits, 0, j0) || metal_at(g, ybits, 1, j1)) return (1u << V) - 1;
  unsigned m = 0;
#pragma unroll
  for (int k = 0; k < V; ++k)
    if (metal_at(g, ybits, 2, z0 + k)) m |= 1u << k;
  return m;
}

// slot C of the vectorised kernel; F0..F2 are the partner centre rows by direction
// uPML coefficients of a group's V points along loop axis ks (-1: none): index 2 j + y with j
// the point's coordinate along ks (the row's own axis varies per point, the others not);
// bit k of `on` is set where the profile is active (sig != 0 or kap != 1)
template <class R, int V> struct ProfV {
  R sig[V], kap[V], sinv[V];
  unsigned on;
};
template <class R, int V>
__device__ __forceinline__ void load_prof(ProfV<R, V> &p, const Grid &g, const R *sig, const R *kap, const R *siginv, int ks, int y,
                                          int j0, int j1, int z0) {
  p.on = 0;
#pragma unroll
  for (int k = 0; k < V; ++k) p.sig[k] = R(0), p.kap[k] = R(1), p.sinv[k] = R(1);
  if (!sig || ks < 0) return;
  if (ks == 2) {
    if (z0 >= g.qlo[2] && z0 + V - 1 <= g.qhi[2]) return; // the whole group is quiet
    const int nz = g.nax[2];
#pragma unroll
    for (int k = 0; k < V; ++k) {
      const int idx = 2 * min(z0 + k, nz) + y; // points past the row's end are masked off
      p.sig[k] = sig[idx];
      p.kap[k] = kap[idx];
      if (siginv) p.sinv[k] = siginv[idx];
      if (p.sig[k] != R(0) || p.kap[k] != R(1)) p.on |= 1u << k;
    }
  }
  else {
    const int j = ks == 0 ? j0 : j1;
    if (j >= (ks == 0 ? g.qlo[0] : g.qlo[1]) && j <= (ks == 0 ? g.qhi[0] : g.qhi[1])) return;
    const int idx = 2 * j + y;
    const R sg = sig[idx], kp = kap[idx], si = siginv ? siginv[idx] : R(1);
#pragma unroll
    for (int k = 0; k < V; ++k) p.sig[k] = sg, p.kap[k] = kp, p.sinv[k] = si;
    if (sg != R(0) || kp != R(1)) p.on = (1u << V) - 1;
  }
}

// lanes of a group outside the quiet range of loop axis ks (where a uPML profile of that axis
// may be active); decided from the launch's quiet ranges, without loading the profile
template <int V> __device__ __forceinline__ unsigned prof_lanes(const Grid &g, int ks, int j0, int j1, int z0) {
  if (ks < 0) return 0u;
  if (ks == 2) {
    unsigned m = 0;
#pragma unroll
    for (int k = 0; k < V; ++k)
      if (z0 + k < g.qlo[2] || z0 + k > g.qhi[2]) m |= 1u << k;
    return m;
  }
  const int j = ks == 0 ? j0 : j1;
  return (j >= (ks == 0 ? g.qlo[0] : g.qlo[1]) && j <= (ks == 0 ? g.qhi[0] : g.qhi[1])) ? 0u : (1u << V) - 1;
}

// One slot (component) of a group: curl (B or D) and/or update (H or E, and the poles).
// NP: the poles the update handles (0, 1, 2 or MAXPOLE; registers are allocated for NP).
// Every array load of the slot is issued before its first store: the compiler may not move a
// load above a store it cannot prove independent, so a thread would otherwise wait for memory
// several times per slot.  Profile loads (cached 1-D arrays) stay with the arithmetic.
template <class R, int V, bool CURL, bool UPD, bool PML, int NP, int C>
__device__ __forceinline__ void fast_slot(const StepArgs<R> &a, uint32_t i, int j0, int j1, int z0, unsigned mask,
                                          const Vec<R, V> &F0, const Vec<R, V> &F1, const Vec<R, V> &F2, const double *vrow) {
  if (!mask) return;
  const Grid &g = a.g;
  const int ybits = CURL ? a.cc[C].ax.ybits : a.uc[C].ax.ybits;
  const unsigned metal = group_metal<V>(g, ybits, j0, j1, z0);
  const bool upd = UPD && a.uc[C].f;
  // slot C = direction C: g1 = F[(C+2)%3] along C+1, g2 = F[(C+1)%3] along C+2
  constexpr int dp = (C + 1) % 3, dm = (C + 2) % 3;
  const Vec<R, V> &g1c = C == 0 ? F2 : (C == 1 ? F0 : F1);
  const Vec<R, V> &g2c = C == 0 ? F1 : (C == 1 ? F2 : F0);

  // ---- loads ----
  Vec<R, V> fv, g1n, g2n;
  Vec<double, V> fuv; // f_u is FP64 whatever R is
  unsigned upot = 0;  // lanes where the f_u profile may be active
  if (CURL) {
    const CurlComp<R> &kc = a.cc[C];
    if (a.hasp[C]) g1n = neighbour<R, V>(a.F[(C + 2) % 3], i, a.sgn * g.st[dp], g1c, mask);
    if (a.hasm[C]) g2n = neighbour<R, V>(a.F[(C + 1) % 3], i, a.sgn * g.st[dm], g2c, mask);
    fv = vload<R, V>(kc.f + i);
    if (PML && kc.sigu) {
      upot = prof_lanes<V>(g, kc.ksigu, j0, j1, z0) & mask;
      if (upot) fuv = vload_aux<V>(kc.fu + i);
    }
  }
  Vec<R, V> gv, uu, Pc[NP > 0 ? NP : 1], Ppv[NP > 0 ? NP : 1], sgv[NP > 0 ? NP : 1], eold;
  Vec<double, V> fwv; // f_w is FP64 whatever R is
  bool pin[NP > 0 ? NP : 1]; // the group meets pole p's box
  unsigned wpot = 0;         // lanes where the f_w profile may be active
  if (upd) {
    const UpdComp<R> &u = a.uc[C];
    if (!CURL) gv = vload<R, V>(u.g + i);
    if (u.u) uu = vload<R, V>(u.u + i);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      pin[p] = p < a.npole && u.P[p] && j0 >= a.pole[p].lo[0] && j0 <= a.pole[p].hi[0] && j1 >= a.pole[p].lo[1] &&
               j1 <= a.pole[p].hi[1] && z0 + V - 1 >= a.pole[p].lo[2] && z0 <= a.pole[p].hi[2];
      if (pin[p]) {
        Pc[p] = vload<R, V>(u.P[p] + i);
        if (u.sigma[p]) {
          Ppv[p] = vload<R, V>(u.Pp[p] + i);
          if ((u.sigone >> p) & 1) { // one value over the pole's box, zero outside it
#pragma unroll
            for (int k = 0; k < V; ++k)
              sgv[p].v[k] = z0 + k >= a.pole[p].lo[2] && z0 + k <= a.pole[p].hi[2] ? u.sigval[p] : R(0);
          }
          else
            sgv[p] = vload<R, V>(u.sigma[p] + i);
        }
      }
    }
    if (PML && u.sigw) {
      wpot = prof_lanes<V>(g, u.ksigw, j0, j1, z0) & mask;
      if (wpot) {
        fwv = vload_aux<V>(u.fw + i);
        eold = vload<R, V>(u.f + i);
      }
    }
  }

  // ---- arithmetic ----
  unsigned um = 0;
  if (CURL) {
    const CurlComp<R> &kc = a.cc[C];
    Vec<R, V> cu;
#pragma unroll
    for (int k = 0; k < V; ++k) cu.v[k] = R(0);
    if (a.hasp[C]) {
#pragma unroll
      for (int k = 0; k < V; ++k) cu.v[k] = g1n.v[k] - g1c.v[k];
    }
    if (a.hasm[C]) {
      if (a.hasp[C]) {
#pragma unroll
        for (int k = 0; k < V; ++k) cu.v[k] = cu.v[k] + g2c.v[k] - g2n.v[k];
      }
      else {
#pragma unroll
        for (int k = 0; k < V; ++k) cu.v[k] = g2c.v[k] - g2n.v[k];
      }
    }
    if (!PML) {
#pragma unroll
      for (int k = 0; k < V; ++k) fv.v[k] -= a.dtdx * cu.v[k];
    }
    else {
      // the four step_curl branches per point (src/step_generic.cpp:65-249)
      ProfV<R, V> P, U;
      load_prof<R, V>(P, g, kc.sig, kc.kap, kc.siginv, kc.ksig, kc.ysig, j0, j1, z0);
      load_prof<R, V>(U, g, kc.sigu, kc.kapu, kc.siginvu, kc.ksigu, kc.ysigu, j0, j1, z0);
      um = U.on & mask & upot;
      if (!(P.on | um)) {
        // no active profile in the group: the plain update (a wave with no uPML point skips
        // the per-point branches below entirely)
#pragma unroll
        for (int k = 0; k < V; ++k) fv.v[k] -= a.dtdx * cu.v[k];
      }
      else
#pragma unroll
      for (int k = 0; k < V; ++k) {
        const bool pml = (P.on >> k) & 1, pmlu = (um >> k) & 1;
        if (!pmlu) {
          if (!pml)
            fv.v[k] -= a.dtdx * cu.v[k];
          else
            fv.v[k] = ((P.kap[k] - P.sig[k]) * fv.v[k] - a.dtdx * cu.v[k]) * P.sinv[k];
        }
        else {
          // f_u in FP64
          const double fprev = fuv.v[k], dtdx = a.dtdx, cud = cu.v[k];
          double fun;
          if (!pml)
            fun = fprev - dtdx * cud;
          else
            fun = ((double(P.kap[k]) - double(P.sig[k])) * fprev - dtdx * cud) * double(P.sinv[k]);
          fuv.v[k] = fun;
          fv.v[k] = R(double(U.sinv[k]) * ((double(U.kap[k]) - double(U.sig[k])) * double(fv.v[k]) + fun - fprev));
        }
      }
    }
    if (a.src && in_src_box<V>(a, j0, j1, z0)) {
#pragma unroll
      for (int k = 0; k < V; ++k) fv.v[k] = apply_sources<R, false, C>(a, i + k, j0, j1, z0 + k, fv.v[k], vrow);
    }
    if (metal) {
#pragma unroll
      for (int k = 0; k < V; ++k)
        if ((metal >> k) & 1) fv.v[k] = R(0);
    }
  }
  Vec<R, V> ev, Pn[NP > 0 ? NP : 1];
  unsigned wm = 0;
  if (upd) {
    const UpdComp<R> &u = a.uc[C];
    if (CURL) gv = fv;
#pragma unroll
    for (int p = 0; p < NP; ++p)
      if (pin[p]) {
#pragma unroll
        for (int k = 0; k < V; ++k) gv.v[k] -= Pc[p].v[k];
      }
    if (a.src && in_src_box<V>(a, j0, j1, z0)) {
#pragma unroll
      for (int k = 0; k < V; ++k) gv.v[k] = apply_sources<R, true, C>(a, i + k, j0, j1, z0 + k, gv.v[k], vrow);
    }
    ev = gv;
    if (u.u) {
#pragma unroll
      for (int k = 0; k < V; ++k) ev.v[k] = gv.v[k] * uu.v[k];
    }
    const Vec<R, V> wv = ev; // W of the polarisations: E, or f_w in the uPML
    if (PML && wpot) {
      // step_update_EDHB's uPML branch per point (src/step_generic.cpp:682-701)
      ProfV<R, V> W;
      load_prof<R, V>(W, g, u.sigw, u.kapw, nullptr, u.ksigw, u.ysigw, j0, j1, z0);
      wm = W.on & mask & wpot;
#pragma unroll
      for (int k = 0; k < V; ++k)
        if ((wm >> k) & 1) {
          // f_w in FP64
          const double fwprev = fwv.v[k], fwn = ev.v[k], kw = W.kap[k], sw = W.sig[k];
          fwv.v[k] = fwn;
          ev.v[k] = R(double(eold.v[k]) + ((kw + sw) * fwn - (kw - sw) * fwprev));
        }
    }
#pragma unroll
    for (int p = 0; p < NP; ++p)
      if (pin[p] && u.sigma[p]) {
        const PoleCoef<R> &q = a.pole[p];
#pragma unroll
        for (int k = 0; k < V; ++k)
          Pn[p].v[k] = q.gamma1inv * (Pc[p].v[k] * (R(2) - q.omega0dtsqr_denom) - q.gamma1 * Ppv[p].v[k] +
                                      q.omega0dtsqr * (sgv[p].v[k] * wv.v[k]));
      }
    if (metal) {
#pragma unroll
      for (int k = 0; k < V; ++k)
        if ((metal >> k) & 1) ev.v[k] = R(0);
    }
  }

  // ---- stores ----
  if (CURL) {
    const CurlComp<R> &kc = a.cc[C];
    if (PML && um) vstore_aux<V>(kc.fu + i, fuv, um);
    put_row<R, V>(kc.f, i, fv, mask, g, ybits, j0, j1, z0);
  }
  if (upd) {
    const UpdComp<R> &u = a.uc[C];
    if (PML && wm) vstore_aux<V>(u.fw + i, fwv, wm);
#pragma unroll
    for (int p = 0; p < NP; ++p)
      if (pin[p] && u.sigma[p]) {
        if (a.pswap)
          put_row<R, V>(u.Pp[p], i, Pn[p], mask, g, ybits, j0, j1, z0);
        else {
          put_row<R, V>(u.P[p], i, Pn[p], mask, g, ybits, j0, j1, z0);
          vstore<R, V>(u.Pp[p] + i, Pc[p], mask);
        }
      }
    put_row<R, V>(u.f, i, ev, mask, g, ybits, j0, j1, z0);
  }
}

// group t of the launch box of a (groups of V points along the rows, from group g0)
template <class R, int V, bool CURL, bool UPD, bool PML, int NP>
__device__ __forceinline__ void fast_body(const StepArgs<R> &a, uint32_t t, int g0, int ngroups, const double *vrow) {
  const Grid &g = a.g;
  const uint32_t r = t / uint32_t(ngroups);
  const int grp = int(t - r * uint32_t(ngroups));
  const int q0 = int(r / uint32_t(g.an[1]));
  const int q1 = int(r - uint32_t(q0) * uint32_t(g.an[1]));
  const int j0 = g.alo[0] + q0, j1 = g.alo[1] + q1;
  const int z0 = (g0 + grp) * V;
  const uint32_t i = uint32_t(j0 * g.ast[0] + j1 * g.ast[1] + z0);
  // points of the group inside the launch box along the row
  unsigned boxmask = 0;
#pragma unroll
  for (int k = 0; k < V; ++k)
    if (z0 + k >= g.alo[2] && z0 + k < g.alo[2] + g.an[2]) boxmask |= 1u << k;
  // ownership of each component slot in this group
  unsigned m0 = 0, m1 = 0, m2 = 0;
  if (CURL ? a.cc[0].f : a.uc[0].f) m0 = group_owned<V>(g, CURL ? a.cc[0].ax.ybits : a.uc[0].ax.ybits, j0, j1, z0, boxmask);
  if (CURL ? a.cc[1].f : a.uc[1].f) m1 = group_owned<V>(g, CURL ? a.cc[1].ax.ybits : a.uc[1].ax.ybits, j0, j1, z0, boxmask);
  if (CURL ? a.cc[2].f : a.uc[2].f) m2 = group_owned<V>(g, CURL ? a.cc[2].ax.ybits : a.uc[2].ax.ybits, j0, j1, z0, boxmask);
  if (!(m0 | m1 | m2)) return;
  // centre rows of the partner components (by direction), loaded once
  Vec<R, V> F0, F1, F2;
  if (CURL) {
    if (a.F[0]) F0 = vload<R, V>(a.F[0] + i);
    if (a.F[1]) F1 = vload<R, V>(a.F[1] + i);
    if (a.F[2]) F2 = vload<R, V>(a.F[2] + i);
  }
  fast_slot<R, V, CURL, UPD, PML, NP, 0>(a, i, j0, j1, z0, m0, F0, F1, F2, vrow);
  fast_slot<R, V, CURL, UPD, PML, NP, 1>(a, i, j0, j1, z0, m1, F0, F1, F2, vrow);
  fast_slot<R, V, CURL, UPD, PML, NP, 2>(a, i, j0, j1, z0, m2, F0, F1, F2, vrow);
}


// group t of the launch box for slot C alone
template <class R, int V, bool CURL, bool UPD, bool PML, int NP, int C>
__device__ __forceinline__ void fast_body_slot(const StepArgs<R> &a, uint32_t t, int g0, int ngroups) {
  const Grid &g = a.g;
  const uint32_t r = t / uint32_t(ngroups);
  const int grp = int(t - r * uint32_t(ngroups));
  const int q0 = int(r / uint32_t(g.an[1]));
  const int q1 = int(r - uint32_t(q0) * uint32_t(g.an[1]));
  const int j0 = g.alo[0] + q0, j1 = g.alo[1] + q1;
  const int z0 = (g0 + grp) * V;
  const uint32_t i = uint32_t(j0 * g.ast[0] + j1 * g.ast[1] + z0);
  unsigned boxmask = 0;
#pragma unroll
  for (int k = 0; k < V; ++k)
    if (z0 + k >= g.alo[2] && z0 + k < g.alo[2] + g.an[2]) boxmask |= 1u << k;
  if (!(CURL ? a.cc[C].f : a.uc[C].f)) return;
  const unsigned m = group_owned<V>(g, CURL ? a.cc[C].ax.ybits : a.uc[C].ax.ybits, j0, j1, z0, boxmask);
  if (!m) return;
  Vec<R, V> F0, F1, F2;
  if (CURL) {
    if (a.F[0]) F0 = vload<R, V>(a.F[0] + i);
    if (a.F[1]) F1 = vload<R, V>(a.F[1] + i);
    if (a.F[2]) F2 = vload<R, V>(a.F[2] + i);
  }
  fast_slot<R, V, CURL, UPD, PML, NP, C>(a, i, j0, j1, z0, m, F0, F1, F2, a.vrow);
}

// With an update (UPD), one slot per block (logical block mod 3): a thread carries one
// slot's registers, and the three blocks of a tile, dealt to the same XCD, share the partner
// rows in its L2.  A curl alone keeps the three slots in one thread (measured faster there:
// its threads are light, and the shared centre rows are loaded once).
template <class R, int V, bool CURL, bool UPD, bool PML, int NP>
__global__ __launch_bounds__(BLOCK) void k_fast(const StepArgs<R> a, uint32_t count, int g0, int ngroups) {
  if (!UPD) {
    const uint32_t t = xcd_block(a.rev) * uint32_t(BLOCK) + threadIdx.x;
    if (t >= count) return;
    fast_body<R, V, CURL, UPD, PML, NP>(a, t, g0, ngroups, a.vrow);
    return;
  }
  const uint32_t lb = xcd_block(a.rev);
  const int C = int(lb % 3u);
  const uint32_t t = (lb / 3u) * uint32_t(BLOCK) + threadIdx.x;
  if (t >= count) return;
  if (C == 0) fast_body_slot<R, V, CURL, UPD, PML, NP, 0>(a, t, g0, ngroups);
  else if (C == 1) fast_body_slot<R, V, CURL, UPD, PML, NP, 1>(a, t, g0, ngroups);
  else fast_body_slot<R, V, CURL, UPD, PML, NP, 2>(a, t, g0, ngroups);
}

// ---------------------------------------------------------------------------------------
// 3D row kernel: V consecutive z points of one (x, y) row per thread, for launch boxes of a
// 3D chunk with no active x or y uPML profile and no folded periodic images (the interior
// box of a chunk stepped alone).  In 3D the loop axes are the directions, so each slot's Yee
// shifts, curl partners and uPML directions are compile-time facts: slot C reads its
// partners F[(C+2)%3] along C+1 and F[(C+1)%3] along C+2 (src/fields.cpp:428-456); the z
// uPML acts through f_u on slot 0 (dsigu = z), on f itself for slot 1 (dsig = z) and through
// f_w on slot 2's update (dsigw = z) (src/step_db.cpp:56-59, src/update_eh.cpp:162-163), lane
// by lane and only in the groups outside the z quiet range.  The generic per-slot logic of
// k_fast (runtime loop axes, periodic images, x/y profiles) is absent, so the per-point
// instruction count is a fraction of k_fast's; sources run in a non-inlined helper entered
// only by the groups inside the source box.  Each block updates one slot (logical block mod
// 3), so a thread carries one slot's registers and the three blocks of a tile, dealt to the
// same XCD, share the partner rows in its L2.  Arithmetic is the reference's, as in k_fast.

// t / d for every 32-bit t by a multiply-high, a subtraction and two shifts (the round-up
// method of Granlund and Montgomery, "Division by invariant integers using multiplication",
// PLDI 1994): with l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1,
// t / d = (h + ((t - h) >> 1)) >> (l - 1) where h = mulhi(m, t); d = 1 takes m = 0, no shifts.
struct Div32 {
  uint32_t m;
  int s1, s2;
};
static inline Div32 div32(uint32_t d) {
  Div32 r{0, 0, 0};
  if (d <= 1) return r;
  int l = 0;
  while ((uint64_t(1) << l) < d) ++l;
  r.m = uint32_t(((uint64_t(1) << 32) * ((uint64_t(1) << l) - d)) / d + 1);
  r.s1 = 1;
  r.s2 = l - 1;
  return r;
}
__host__ __device__ __forceinline__ uint32_t div32_apply(const Div32 &q, uint32_t t) {
#ifdef __HIP_DEVICE_COMPILE__
  const uint32_t h = __umulhi(q.m, t);
#else
  const uint32_t h = uint32_t((uint64_t(q.m) * t) >> 32);
#endif
  return (h + ((t - h) >> q.s1)) >> q.s2;
}

struct Row3Geom {
  uint32_t sx, sy;      // strides along x and y (z is stride 1; a chunk has < 2^32 points)
  int nx, ny, nz;       // cells per direction
  int x0, y0;           // launch box low corner in x and y
  uint32_t nyb;         // rows along y in the box
  int g0;               // first group along z
  uint32_t ngroups;     // groups per row
  Div32 dgroups, dnyb;  // division by ngroups and by nyb
  int zlo, zhi;         // launch box along z
  int qzlo, qzhi;       // z indices where no z uPML profile is active
  int mhi_bits, mlo_bits; // bit d: the High wall normal to d is Metallic / the Low wall Magnetic
  int rev;              // visit logical blocks in descending order
  int cmp;              // real (0) or imaginary (1) parts
};
template <class R, int NP> struct Row3Args {
  Row3Geom g;
  R dtdx;
  const R *F[3];        // curl partners by direction
  R *f[3];              // B or D by direction (null: absent)
  double *fu0;          // f_u of slot 0 (its dsigu is z, FP64), null when none
  R *e[3];              // H or E (null: no separate update array)
  const R *chi[3];      // diagonal chi1inv / mu1inv (null: 1)
  double *fw2;          // f_w of slot 2 (its dsigw is z, FP64), null when none
  const R *zsig, *zkap, *zsiginv; // the z uPML profile (null: no z uPML)
  // poles: P and P_prev of pole p for component c are at pofs[p] and ppofs[p] (values) from
  // pbase[c], its sigma at sofs[p] from sbase[c]: six pointers instead of three per pole and
  // component, offsets the same for every component
  R *pbase[3];
  const R *sbase[3];
  int64_t pofs[NP > 0 ? NP : 1], ppofs[NP > 0 ? NP : 1], sofs[NP > 0 ? NP : 1];
  int pmask;            // bit 3 p + c: pole p has a polarisation for component c
  int smask;            // bit 3 p + c: ... and a sigma array
  R sigval[NP > 0 ? NP : 1][3];
  int sigone;           // bit 3 p + c: pole p's sigma of component c is sigval over its box
  PoleCoef<R> pole[NP > 0 ? NP : 1];
  int pswap;
  const SrcSet *src;    // null: no sources in this launch
  int sbox_lo[3], sbox_hi[3];
  const double *vrow;
};

// Yee shift of slot C along direction d: electric components (the D phase, PH 1) are shifted
// along their own direction, magnetic ones (the B phase, PH 0) along the others
template <int PH, int C, int D> constexpr int r3_ysh() { return PH ? (C == D ? 1 : 0) : (C == D ? 0 : 1); }

// lanes of a group owned by slot C and inside the box along z (x and y are checked per thread)
template <int V, int PH, int C> __device__ __forceinline__ unsigned r3_zmask(const Row3Geom &g, int z0) {
  constexpr int y = r3_ysh<PH, C, 2>();
  unsigned m = 0;
#pragma unroll
  for (int k = 0; k < V; ++k) {
    const int z = z0 + k;
    if (z >= g.zlo && z <= g.zhi && z >= 1 - y && z <= g.nz - y) m |= 1u << k;
  }
  return m;
}
template <int PH, int C> __device__ __forceinline__ bool r3_xyown(const Row3Geom &g, int x, int y) {
  constexpr int yx = r3_ysh<PH, C, 0>(), yy = r3_ysh<PH, C, 1>();
  return x >= 1 - yx && x <= g.nx - yx && y >= 1 - yy && y <= g.ny - yy;
}
// metal planes (on_metal_boundary, src/boundaries.cpp:191-205): all lanes when x or y is on
// one, else the z lanes on one
template <int V, int PH, int C> __device__ __forceinline__ unsigned r3_metal(const Row3Geom &g, int x, int y, int z0) {
  if (!(g.mhi_bits | g.mlo_bits)) return 0u;
  constexpr int yx = r3_ysh<PH, C, 0>(), yy = r3_ysh<PH, C, 1>(), yz = r3_ysh<PH, C, 2>();
  const bool mx_ = (((g.mhi_bits >> 0) & 1) && yx == 0 && x == g.nx) || (((g.mlo_bits >> 0) & 1) && yx == 1 && x == 0);
  const bool my_ = (((g.mhi_bits >> 1) & 1) && yy == 0 && y == g.ny) || (((g.mlo_bits >> 1) & 1) && yy == 1 && y == 0);
  if (mx_ || my_) return (1u << V) - 1;
  unsigned m = 0;
#pragma unroll
  for (int k = 0; k < V; ++k) {
    const int z = z0 + k;
    if ((((g.mhi_bits >> 2) & 1) && yz == 0 && z == g.nz) || (((g.mlo_bits >> 2) & 1) && yz == 1 && z == 0)) m |= 1u << k;
  }
  return m;
}

// sources of slot C on the lanes of one group (entered only inside the launch's source box);
// the group's values travel by value, so the caller's registers never need an address
template <class R, int V, bool DIP, int C>
__device__ __noinline__ Vec<R, V> r3_sources(const SrcSet *S, int cmp, const double *vrow, uint32_t i, int x, int y, int z0,
                                             unsigned mask, Vec<R, V> v) {
  const SrcTab &t = DIP ? S->dip : S->cur;
  const SrcBoxes &b = DIP ? S->dipb : S->curb;
#pragma unroll
  for (int k = 0; k < V; ++k) {
    if (!((mask >> k) & 1)) continue;
    R w = v.v[k];
    if (t.n && int64_t(i + k) >= t.lo && int64_t(i + k) <= t.hi) w = apply_tab(t, C, i + k, w, vrow);
    if (b.n) w = apply_boxes(b, cmp, C, x, y, z0 + k, w, vrow);
    v.v[k] = w;
  }
  return v;
}

template <class R, int V, int PH, bool UPD, int NP, bool ZP, int C>
__device__ __forceinline__ void r3_slot(const Row3Args<R, NP> &a, uint32_t i, int x, int y, int z0, unsigned zm, const Vec<R, V> &F0,
                                        const Vec<R, V> &F1, const Vec<R, V> &F2, bool srcs, unsigned pinxy) {
  const Row3Geom &g = a.g;
  R *const fp = a.f[C];
  if (!fp || !zm || !r3_xyown<PH, C>(g, x, y)) return;
  const unsigned mask = zm;
  const unsigned metal = r3_metal<V, PH, C>(g, x, y, z0);
  // curl partners: g1 = F[(C+2)%3] along dp, g2 = F[(C+1)%3] along dm; forward differences
  // for B (i + stride), backward for D (i - stride), the D update's negated strides
  // (src/step_db.cpp:81-84)
  constexpr int dp = (C + 1) % 3, dm = (C + 2) % 3;
  const Vec<R, V> &g1c = C == 0 ? F2 : (C == 1 ? F0 : F1);
  const Vec<R, V> &g2c = C == 0 ? F1 : (C == 1 ? F2 : F0);
  const R *g1p = a.F[(C + 2) % 3], *g2p = a.F[(C + 1) % 3];
  constexpr int64_t sg = PH ? -1 : 1;
  const bool zq = !ZP || (z0 >= g.qzlo && z0 + V - 1 <= g.qzhi); // no z uPML lane in the group
  // ---- loads ----
  Vec<R, V> g1n, g2n;
  if (dp == 2) g1n = neighbour<R, V>(g1p, i, sg, g1c, mask);
  else g1n = vload<R, V>(g1p + (PH ? i - (dp == 0 ? g.sx : g.sy) : i + (dp == 0 ? g.sx : g.sy)));
  if (dm == 2) g2n = neighbour<R, V>(g2p, i, sg, g2c, mask);
  else g2n = vload<R, V>(g2p + (PH ? i - (dm == 0 ? g.sx : g.sy) : i + (dm == 0 ? g.sx : g.sy)));
  Vec<R, V> fv = vload<R, V>(fp + i);
  Vec<double, V> fuv, fwv; // the FP64 uPML auxiliaries
  const bool zfu = ZP && C == 0 && !zq && a.fu0;
  if (zfu) fuv = vload_aux<V>(a.fu0 + i);
  R *const ep = UPD ? a.e[C] : nullptr;
  Vec<R, V> uu, Pc[NP > 0 ? NP : 1], Ppv[NP > 0 ? NP : 1], sgv[NP > 0 ? NP : 1], eold;
  bool pin[NP > 0 ? NP : 1];
  const bool zfw = UPD && ZP && C == 2 && !zq && a.fw2 && ep;
  if (ep) {
    if (a.chi[C]) uu = vload<R, V>(a.chi[C] + i);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const PoleCoef<R> &q = a.pole[p];
      pin[p] = ((a.pmask >> (3 * p + C)) & 1) && ((pinxy >> p) & 1) && z0 + V - 1 >= q.lo[2] && z0 <= q.hi[2];
      if (pin[p]) {
        Pc[p] = vload<R, V>(a.pbase[C] + (a.pofs[p] + i));
        if ((a.smask >> (3 * p + C)) & 1) {
          Ppv[p] = vload<R, V>(a.pbase[C] + (a.ppofs[p] + i));
          if ((a.sigone >> (3 * p + C)) & 1) {
#pragma unroll
            for (int k = 0; k < V; ++k) sgv[p].v[k] = z0 + k >= q.lo[2] && z0 + k <= q.hi[2] ? a.sigval[p][C] : R(0);
          }
          else
            sgv[p] = vload<R, V>(a.sbase[C] + (a.sofs[p] + i));
        }
      }
    }
    if (zfw) {
      fwv = vload_aux<V>(a.fw2 + i);
      eold = vload<R, V>(ep + i);
    }
  }
  // ---- curl and the z uPML of slots 0 and 1 ----
  Vec<R, V> cu;
#pragma unroll
  for (int k = 0; k < V; ++k) cu.v[k] = g1n.v[k] - g1c.v[k] + g2c.v[k] - g2n.v[k];
  unsigned um = 0; // lanes whose f_u takes part
  if (ZP && C != 2 && !zq && a.zsig) {
    // per-lane profile at half-pixel index 2 z + ysh_z (src/meep_internals.hpp:217-226)
    constexpr int yz = r3_ysh<PH, C, 2>();
    R sg_[V], kp_[V], si_[V];
    unsigned on = 0;
#pragma unroll
    for (int k = 0; k < V; ++k) {
      const int idx = 2 * min(z0 + k, g.nz) + yz;
      sg_[k] = a.zsig[idx];
      kp_[k] = a.zkap[idx];
      si_[k] = a.zsiginv[idx];
      if (sg_[k] != R(0) || kp_[k] != R(1)) on |= 1u << k;
    }
    on &= mask;
    if (C == 0 && zfu) {
      // step_curl through f_u (dsigu = z, no uPML on f): src/step_generic.cpp:130-151
      um = on;
#pragma unroll
      for (int k = 0; k < V; ++k) {
        if ((um >> k) & 1) {
          // f_u in FP64
          const double fprev = fuv.v[k], dtdx = a.dtdx, cud = cu.v[k];
          const double fun = fprev - dtdx * cud;
          fuv.v[k] = fun;
          fv.v[k] = R(double(si_[k]) * ((double(kp_[k]) - double(sg_[k])) * double(fv.v[k]) + fun - fprev));
        }
        else
          fv.v[k] -= a.dtdx * cu.v[k];
      }
    }
    else if (C == 1) {
      // uPML on f (dsig = z): src/step_generic.cpp:179-192
#pragma unroll
      for (int k = 0; k < V; ++k) {
        if ((on >> k) & 1)
          fv.v[k] = ((kp_[k] - sg_[k]) * fv.v[k] - a.dtdx * cu.v[k]) * si_[k];
        else
          fv.v[k] -= a.dtdx * cu.v[k];
      }
    }
    else {
#pragma unroll
      for (int k = 0; k < V; ++k) fv.v[k] -= a.dtdx * cu.v[k];
    }
  }
  else {
#pragma unroll
    for (int k = 0; k < V; ++k) fv.v[k] -= a.dtdx * cu.v[k];
  }
  if (srcs) fv = r3_sources<R, V, false, C>(a.src, g.cmp, a.vrow, i, x, y, z0, mask, fv);
  if (metal) {
#pragma unroll
    for (int k = 0; k < V; ++k)
      if ((metal >> k) & 1) fv.v[k] = R(0);
  }
  // ---- update: E = chi1inv (D - sum P - dipoles), f_w of slot 2, the poles ----
  Vec<R, V> ev, Pn[NP > 0 ? NP : 1];
  unsigned wm = 0;
  if (ep) {
    Vec<R, V> gv = fv;
#pragma unroll
    for (int p = 0; p < NP; ++p)
      if (pin[p]) {
#pragma unroll
        for (int k = 0; k < V; ++k) gv.v[k] -= Pc[p].v[k];
      }
    if (srcs) gv = r3_sources<R, V, true, C>(a.src, g.cmp, a.vrow, i, x, y, z0, mask, gv);
    ev = gv;
    if (a.chi[C]) {
#pragma unroll
      for (int k = 0; k < V; ++k) ev.v[k] = gv.v[k] * uu.v[k];
    }
    const Vec<R, V> wv = ev; // W of the polarisations: E, or f_w where the profile acts
    if (zfw) {
      // step_update_EDHB's uPML branch (src/step_generic.cpp:682-701), dsigw = z
      constexpr int yz = r3_ysh<PH, C, 2>();
#pragma unroll
      for (int k = 0; k < V; ++k) {
        const int idx = 2 * min(z0 + k, g.nz) + yz;
        const R sw = a.zsig[idx], kw = a.zkap[idx];
        if (((mask >> k) & 1) && (sw != R(0) || kw != R(1))) {
          wm |= 1u << k;
          // f_w in FP64
          const double fwprev = fwv.v[k], fwn = ev.v[k], kwd = kw, swd = sw;
          fwv.v[k] = fwn;
          ev.v[k] = R(double(eold.v[k]) + ((kwd + swd) * fwn - (kwd - swd) * fwprev));
        }
      }
    }
#pragma unroll
    for (int p = 0; p < NP; ++p)
      if (pin[p] && ((a.smask >> (3 * p + C)) & 1)) {
        const PoleCoef<R> &q = a.pole[p];
#pragma unroll
        for (int k = 0; k < V; ++k)
          Pn[p].v[k] = q.gamma1inv * (Pc[p].v[k] * (R(2) - q.omega0dtsqr_denom) - q.gamma1 * Ppv[p].v[k] +
                                      q.omega0dtsqr * (sgv[p].v[k] * wv.v[k]));
      }
    if (metal) {
#pragma unroll
      for (int k = 0; k < V; ++k)
        if ((metal >> k) & 1) ev.v[k] = R(0);
    }
  }
  // ---- stores ----
  if (um) vstore_aux<V>(a.fu0 + i, fuv, um);
  vstore<R, V>(fp + i, fv, mask);
  if (ep) {
    if (wm) vstore_aux<V>(a.fw2 + i, fwv, wm);
#pragma unroll
    for (int p = 0; p < NP; ++p)
      if (pin[p] && ((a.smask >> (3 * p + C)) & 1)) {
        if (a.pswap)
          vstore<R, V>(a.pbase[C] + (a.ppofs[p] + i), Pn[p], mask);
        else {
          vstore<R, V>(a.pbase[C] + (a.pofs[p] + i), Pn[p], mask);
          vstore<R, V>(a.pbase[C] + (a.ppofs[p] + i), Pc[p], mask);
        }
      }
    vstore<R, V>(ep + i, ev, mask);
  }
}

template <class R, int V, int PH, bool UPD, int NP, bool ZP>
__global__ __launch_bounds__(BLOCK) void k_row3(const Row3Args<R, NP> a, uint32_t count) {
  const uint32_t lb = xcd_block(a.g.rev);
  const int C = int(lb % 3u);
  const uint32_t t = (lb / 3u) * uint32_t(BLOCK) + threadIdx.x;
  if (t >= count) return;
  const Row3Geom &g = a.g;
  const uint32_t r = div32_apply(g.dgroups, t);
  const int grp = int(t - r * g.ngroups);
  const uint32_t qx = div32_apply(g.dnyb, r);
  const int x = g.x0 + int(qx), y = g.y0 + int(r - qx * g.nyb);
  const int z0 = (g.g0 + grp) * V;
  const uint32_t i = uint32_t(x) * g.sx + uint32_t(y) * g.sy + uint32_t(z0);
  const bool srcs = a.src && x >= a.sbox_lo[0] && x <= a.sbox_hi[0] && y >= a.sbox_lo[1] && y <= a.sbox_hi[1] &&
                    z0 + V - 1 >= a.sbox_lo[2] && z0 <= a.sbox_hi[2];
  unsigned pinxy = 0;
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const PoleCoef<R> &q = a.pole[p];
    if (x >= q.lo[0] && x <= q.hi[0] && y >= q.lo[1] && y <= q.hi[1]) pinxy |= 1u << p;
  }
  if (C == 0) {
    const unsigned m = r3_zmask<V, PH, 0>(g, z0);
    if (!m) return;
    const Vec<R, V> F1 = vload<R, V>(a.F[1] + i), F2 = vload<R, V>(a.F[2] + i);
    r3_slot<R, V, PH, UPD, NP, ZP, 0>(a, i, x, y, z0, m, F1, F1, F2, srcs, pinxy);
  }
  else if (C == 1) {
    const unsigned m = r3_zmask<V, PH, 1>(g, z0);
    if (!m) return;
    const Vec<R, V> F0 = vload<R, V>(a.F[0] + i), F2 = vload<R, V>(a.F[2] + i);
    r3_slot<R, V, PH, UPD, NP, ZP, 1>(a, i, x, y, z0, m, F0, F2, F2, srcs, pinxy);
  }
  else {
    const unsigned m = r3_zmask<V, PH, 2>(g, z0);
    if (!m) return;
    const Vec<R, V> F0 = vload<R, V>(a.F[0] + i), F1 = vload<R, V>(a.F[1] + i);
    r3_slot<R, V, PH, UPD, NP, ZP, 2>(a, i, x, y, z0, m, F0, F1, F1, srcs, pinxy);
  }
}

// The row kernel for a's launch box when the box qualifies (3D, loop axes = directions, no
// folded images, no x or y uPML profile active in the box, no conductivity, the 3D curl plan
// in every slot); false leaves the launch to k_fast.
template <class R, int NP>
static bool r3_fill(const StepArgs<R> &a, bool curl, bool upd, Row3Args<R, NP> &b, int V) {
  const Grid &g = a.g;
  if (!curl || a.nc != 3) return false;
  for (int k = 0; k < 3; ++k)
    if (g.axdir[k] != k || g.wrap[k] || g.nax[k] < 1) return false;
  for (int k = 0; k < 2; ++k) // x and y inside their quiet ranges
    if (g.qlo[k] > g.qhi[k] || g.alo[k] < g.qlo[k] || g.alo[k] + g.an[k] - 1 > g.qhi[k]) return false;
  if (a.sgn != 1 && a.sgn != -1) return false;
  for (int d = 0; d < 3; ++d)
    if (!a.F[d]) return false;
  const int PH = a.sgn < 0 ? 1 : 0;
  memset(&b, 0, sizeof b);
  Row3Geom &q = b.g;
  if (g.st[0] <= 0 || g.st[1] <= 0 || g.st[0] >= (int64_t(1) << 32) || g.st[1] >= (int64_t(1) << 32)) return false;
  q.sx = uint32_t(g.st[0]);
  q.sy = uint32_t(g.st[1]);
  if (g.st[2] != 1) return false;
  q.nx = g.nax[0];
  q.ny = g.nax[1];
  q.nz = g.nax[2];
  q.x0 = g.alo[0];
  q.y0 = g.alo[1];
  q.nyb = uint32_t(g.an[1]);
  q.zlo = g.alo[2];
  q.zhi = g.alo[2] + g.an[2] - 1;
  q.g0 = g.alo[2] / V;
  q.ngroups = uint32_t((g.alo[2] + g.an[2] - 1) / V - q.g0 + 1);
  q.dgroups = div32(q.ngroups);
  q.dnyb = div32(q.nyb);
  q.qzlo = g.qlo[2];
  q.qzhi = g.qhi[2];
  q.mhi_bits = g.mhi_bits;
  q.mlo_bits = g.mlo_bits;
  q.rev = a.rev;
  q.cmp = a.cmp;
  b.dtdx = a.dtdx;
  const R *zs = a.pml.sig[2], *zk = a.pml.kap[2], *zi = a.pml.siginv[2];
  b.zsig = zs;
  b.zkap = zk;
  b.zsiginv = zi;
  for (int c = 0; c < 3; ++c) {
    const CurlComp<R> &k = a.cc[c];
    b.F[c] = a.F[c];
    b.f[c] = k.f;
    if (!k.f) continue;
    if (k.cnd || !a.hasp[c] || !a.hasm[c] || !a.F[(c + 1) % 3] || !a.F[(c + 2) % 3]) return false;
    // the 3D uPML plan: dsig = c + 1, dsigu = c + 2 (src/step_db.cpp:56-59), with the z
    // profile acting only through slot 0's f_u and on slot 1's f
    if (k.dsig >= 0 && k.dsig != (c + 1) % 3) return false;
    if (k.dsigu >= 0 && k.dsigu != (c + 2) % 3) return false;
    if ((k.dsig == 2 || k.dsigu == 2) && !zs) return false;
    if (c == 0 && k.dsigu == 2) b.fu0 = k.fu;
    if (c == 0 && k.dsigu == 2 && !k.fu) return false;
    const int ysh = PH ? 1 : 0; // electric: shifted along its own direction; magnetic: not
    if (k.ysh[c] != ysh) return false;
    if (upd && a.uc[c].f) {
      const UpdComp<R> &u = a.uc[c];
      if (u.dsigw >= 0 && u.dsigw != c) return false;
      if (u.dsigw == 2) {
        if (!u.fw || !zs) return false;
        b.fw2 = u.fw;
      }
      b.e[c] = u.f;
      b.chi[c] = u.u;
      if (a.npole > NP) return false;
      for (int p = 0; p < NP && p < a.npole; ++p) {
        if (!u.P[p]) continue;
        if (!u.Pp[p]) return false;
        b.pmask |= 1 << (3 * p + c);
        b.sigval[p][c] = u.sigval[p];
        if ((u.sigone >> p) & 1) b.sigone |= 1 << (3 * p + c);
        if (u.sigma[p]) b.smask |= 1 << (3 * p + c);
      }
    }
  }
  // pole arrays as offsets in per-component blocks: the base is the lowest address of the
  // component's arrays, and every array must sit a whole number of arrays from it, with the
  // same offsets for every component (else the launch goes to k_fast)
  if (NP > 0) {
    for (int c = 0; c < 3; ++c) {
      if (!b.e[c]) continue;
      const UpdComp<R> &u = a.uc[c];
      const R *lo = nullptr, *slo = nullptr;
      for (int p = 0; p < NP && p < a.npole; ++p) {
        if (u.P[p]) {
          if (!lo || u.P[p] < lo) lo = u.P[p];
          if (u.Pp[p] < lo) lo = u.Pp[p];
        }
        if (u.sigma[p] && (!slo || u.sigma[p] < slo)) slo = u.sigma[p];
      }
      b.pbase[c] = const_cast<R *>(lo);
      b.sbase[c] = slo;
    }
    int64_t st = -1;
    for (int c = 0; c < 3; ++c) {
      if (!b.e[c]) continue;
      const UpdComp<R> &u = a.uc[c];
      for (int p = 0; p < NP && p < a.npole; ++p) {
        const int64_t offs[3] = {u.P[p] ? int64_t(u.P[p] - b.pbase[c]) : -1, u.Pp[p] ? int64_t(u.Pp[p] - b.pbase[c]) : -1,
                                 u.sigma[p] ? int64_t(u.sigma[p] - b.sbase[c]) : -1};
        for (int w = 0; w < 3; ++w) {
          if (offs[w] < 0) continue;
          if (offs[w] > 0) {
            if (st < 0) st = offs[w];
            if (offs[w] % st) return false;
          }
        }
      }
    }
    if (st < 0) st = 1;
    int poff[NP > 0 ? NP : 1], ppoff[NP > 0 ? NP : 1], soff[NP > 0 ? NP : 1];
    for (int p = 0; p < NP; ++p) poff[p] = ppoff[p] = soff[p] = -1;
    for (int c = 0; c < 3; ++c) {
      if (!b.e[c]) continue;
      const UpdComp<R> &u = a.uc[c];
      for (int p = 0; p < NP && p < a.npole; ++p) {
        const int64_t o[3] = {u.P[p] ? int64_t(u.P[p] - b.pbase[c]) / st : -1, u.Pp[p] ? int64_t(u.Pp[p] - b.pbase[c]) / st : -1,
                              u.sigma[p] ? int64_t(u.sigma[p] - b.sbase[c]) / st : -1};
        int *dst[3] = {&poff[p], &ppoff[p], &soff[p]};
        for (int w = 0; w < 3; ++w) {
          if (o[w] < 0) continue;
          if (o[w] > 2 * NP) return false;
          if (*dst[w] < 0)
            *dst[w] = int(o[w]);
          else if (*dst[w] != int(o[w]))
            return false;
        }
      }
    }
    for (int p = 0; p < NP; ++p) {
      b.pofs[p] = poff[p] < 0 ? 0 : poff[p] * st;
      b.ppofs[p] = ppoff[p] < 0 ? 0 : ppoff[p] * st;
      b.sofs[p] = soff[p] < 0 ? 0 : soff[p] * st;
    }
  }
  for (int p = 0; p < NP && p < a.npole; ++p) b.pole[p] = a.pole[p];
  b.pswap = a.pswap;
  b.src = a.src;
  for (int k = 0; k < 3; ++k) {
    b.sbox_lo[k] = a.sbox_lo[k];
    b.sbox_hi[k] = a.sbox_hi[k];
  }
  b.vrow = a.vrow;
  return true;
}

template <class R, int V, int PH, bool UPD, int NP>
static bool r3_launch(const StepArgs<R> &a, bool curl, bool upd, hipStream_t s) {
  Row3Args<R, NP> b;
  if (!r3_fill<R, NP>(a, curl, upd, b, V)) return false;
  const size_t n = size_t(b.g.ngroups) * size_t(a.g.an[0]) * size_t(a.g.an[1]);
  if (n >= (size_t(1) << 32)) return false;
  const uint32_t c32 = uint32_t(n);
  const unsigned nb = (3u * nblocks(n) + 7u) / 8u * 8u;
  const bool zp = b.zsig != nullptr;
  if (zp)
    hipLaunchKernelGGL((k_row3<R, V, PH, UPD, NP, true>), dim3(nb), dim3(BLOCK), 0, s, b, c32);
  else
    hipLaunchKernelGGL((k_row3<R, V, PH, UPD, NP, false>), dim3(nb), dim3(BLOCK), 0, s, b, c32);
  return true;
}

template <class R> bool launch_row3(const StepArgs<R> &a, bool curl, bool upd, hipStream_t s) {
  static const bool off = getenv("MEEP_AMD_NO_ROW3") != nullptr;
  if (off) return false;
  constexpr int V = 16 / sizeof(R), VP = V > 2 ? 2 : V; // the pole-carrying updates take 8-byte groups in FP32
  const int PH = a.sgn < 0 ? 1 : 0;
  const bool u = upd;
  if (PH == 0) return u ? r3_launch<R, V, 0, true, 0>(a, curl, upd, s) : r3_launch<R, V, 0, false, 0>(a, curl, upd, s);
  if (!u) return r3_launch<R, V, 1, false, 0>(a, curl, upd, s);
  switch (a.npole) {
    case 0: return r3_launch<R, V, 1, true, 0>(a, curl, upd, s);
    case 1: return r3_launch<R, VP, 1, true, 1>(a, curl, upd, s);
    case 2: return r3_launch<R, VP, 1, true, 2>(a, curl, upd, s);
    default: return false;
  }
}

// ---------------------------------------------------------------------------------------

template <class R> __device__ __forceinline__ void conn_one(const Conn *conns, const double2 *phases, int t, R *const *slots, int ncmp) {
  const Conn e = conns[t];
  R *a0 = slots[2 * e.slot], *a1 = slots[2 * e.slot + 1];
  if (e.cls == 0) { // CONNECT_PHASE: complex<realnum> product
    const double2 ph = phases[t];
    const R re = a0[e.src], im = a1[e.src];
    const R pr = R(ph.x), pi = R(ph.y);
    a0[e.dst] = pr * re - pi * im;
    a1[e.dst] = pr * im + pi * re;
  }
  else if (e.cls == 1) {
    a0[e.dst] = -a0[e.src];
    if (ncmp > 1) a1[e.dst] = -a1[e.src];
  }
  else {
    a0[e.dst] = a0[e.src];
    if (ncmp > 1) a1[e.dst] = a1[e.src];
  }
}
template <class R>
__global__ __launch_bounds__(BLOCK) void k_conn(const Conn *conns, const double2 *phases, int n, R *const *slots, int ncmp) {
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= n) return;
  conn_one<R>(conns, phases, t, slots, ncmp);
}

template <class R>
__global__ __launch_bounds__(BLOCK) void k_src(R *f, const int64_t *idx, const double2 *amp, int n, double2 scal, double mult, int cmp) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= n) return;
  const double2 am = amp[t];
  const double re = am.x * scal.x - am.y * scal.y, im = am.x * scal.y + am.y * scal.x;
  f[idx[t]] = R(double(f[idx[t]]) - (cmp ? im * mult : re * mult));
}

template <class R> __device__ __forceinline__ R avg_value(const R *F, int64_t idx, double w, int64_t a1, int64_t a2) {
  if (a2) return R((w * 0.25) * double(F[idx] + F[idx + a1] + F[idx + a2] + F[idx + (a1 + a2)]));
  if (a1) return R((w * 0.5) * double(F[idx] + F[idx + a1]));
  return R(w * double(F[idx]));
}

// DFT accumulation (src/dft.cpp:270-306).  Each block first forms the phases
// exp(i omega t) * scale of all frequencies once into LDS (src/dft.cpp:270-271); lists longer
// than the table form them per accumulator.
constexpr int DFT_TAB = 512;
__device__ __forceinline__ double2 dft_phase(double omega, double time, double scale_re, double scale_im) {
  double sn, cs;
  sincos(omega * time, &sn, &cs);
  return make_double2(cs * scale_re - sn * scale_im, cs * scale_im + sn * scale_re);
}
// One accumulator in registers, in the chunk's accumulator form: FP64 (d), Kahan-compensated
// FP32 (s, c), or realnum (re, im).
template <class R> struct DftAcc {
  double2 d;
  float2 s, c;
  R re, im;
};
template <class R> __device__ __forceinline__ void dft_load(const DftArgs<R> &a, size_t o, DftAcc<R> &v) {
  if (a.acc_kind == 1)
    v.d = reinterpret_cast<const double2 *>(a.acc)[o];
  else if (a.acc_kind == 2) {
    v.s = reinterpret_cast<const float2 *>(a.acc)[o];
    v.c = reinterpret_cast<const float2 *>(a.comp)[o];
  }
  else {
    v.re = reinterpret_cast<const R *>(a.acc)[2 * o];
    v.im = reinterpret_cast<const R *>(a.acc)[2 * o + 1];
  }
}
template <class R> __device__ __forceinline__ void dft_store(const DftArgs<R> &a, size_t o, const DftAcc<R> &v) {
  if (a.acc_kind == 1)
    reinterpret_cast<double2 *>(a.acc)[o] = v.d;
  else if (a.acc_kind == 2) {
    reinterpret_cast<float2 *>(a.acc)[o] = v.s;
    reinterpret_cast<float2 *>(a.comp)[o] = v.c;
  }
  else {
    reinterpret_cast<R *>(a.acc)[2 * o] = v.re;
    reinterpret_cast<R *>(a.acc)[2 * o + 1] = v.im;
  }
}
// v += phase * (f0 + i f1) (src/dft.cpp:296-306)
template <class R> __device__ __forceinline__ void dft_add(const DftArgs<R> &a, DftAcc<R> &v, double2 phv, R f0, R f1) {
  const double pr = phv.x, pim = phv.y;
  if (a.acc_kind == 1) {
    v.d.x += pr * double(f0) - pim * double(f1);
    v.d.y += pr * double(f1) + pim * double(f0);
    return;
  }
  const R rr = R(pr), ri = R(pim);
  R tr, ti;
  if (a.f1) { // complex<realnum> product
    tr = rr * f0 - ri * f1;
    ti = rr * f1 + ri * f0;
  }
  else {
    tr = f0 * rr;
    ti = f0 * ri;
  }
  if (a.acc_kind == 2) { // Kahan-compensated complex<float>
    const float yr = float(tr) - v.c.x, yi = float(ti) - v.c.y;
    const float sr = v.s.x + yr, si = v.s.y + yi;
    v.c.x = (sr - v.s.x) - yr;
    v.c.y = (si - v.s.y) - yi;
    v.s.x = sr;
    v.s.y = si;
    return;
  }
  v.re += tr;
  v.im += ti;
}
// accumulator o += phase * (f0 + i f1), in the chunk's accumulator form
template <class R> __device__ __forceinline__ void dft_acc(const DftArgs<R> &a, size_t o, double2 phv, R f0, R f1) {
  DftAcc<R> v;
  dft_load(a, o, v);
  dft_add(a, v, phv, f0, f1);
  dft_store(a, o, v);
}
// the averaged, weighted field value of point k (src/dft.cpp:278-294)
template <class R> __device__ __forceinline__ void dft_point(const DftArgs<R> &a, int k, R &f0, R &f1) {
  const int64_t idx = a.idx[k];
  const double w = a.w[k];
  f0 = avg_value(a.f0, idx, w, a.avg1, a.avg2);
  f1 = a.f1 ? avg_value(a.f1, idx, w, a.avg1, a.avg2) : R(0);
}
// accumulator o of a chunk (one thread per accumulator: the batched kernel)
template <class R> __device__ __forceinline__ void dft_one(const DftArgs<R> &a, size_t o, double2 phv) {
  const int k = int(o / size_t(a.nfreq));
  R f0, f1;
  dft_point(a, k, f0, f1);
  dft_acc(a, o, phv, f0, f1);
}

// One thread per point and DFT_Q consecutive frequencies: the point's field values are read
// once, and for Kahan accumulators with nfreq a multiple of DFT_Q the sums and compensations
// move as 16-byte vectors.  Grid-stride over (point, frequency quad).
constexpr int DFT_Q = 4;
template <class R> __global__ __launch_bounds__(BLOCK) void k_dft(const DftArgs<R> a) {
  __shared__ double2 ph[DFT_TAB];
  const int nf = a.nfreq;
  const bool tab = nf <= DFT_TAB;
  if (tab)
    for (int q = threadIdx.x; q < nf; q += BLOCK) ph[q] = dft_phase(a.omega[q], a.time, a.scale_re, a.scale_im);
  __syncthreads();
  const uint32_t nq = uint32_t(nf + DFT_Q - 1) / DFT_Q;
  const uint32_t n = uint32_t(a.npts) * nq;
  const bool vec = a.acc_kind == 2 && tab && nf % DFT_Q == 0;
  for (uint32_t o = blockIdx.x * BLOCK + threadIdx.x; o < n; o += gridDim.x * BLOCK) {
    const uint32_t k = o / nq;
    const int fb = int(o - k * nq) * DFT_Q;
    R f0, f1;
    dft_point(a, int(k), f0, f1);
    const size_t o0 = size_t(k) * size_t(nf) + size_t(fb);
    if (vec) {
      // four complex<float> sums and compensations: 32 bytes each, 32-byte aligned
      float4 *acc = reinterpret_cast<float4 *>(reinterpret_cast<float2 *>(a.acc) + o0);
      float4 *cp = reinterpret_cast<float4 *>(reinterpret_cast<float2 *>(a.comp) + o0);
      float4 s[2] = {acc[0], acc[1]}, c[2] = {cp[0], cp[1]};
      float *sv = reinterpret_cast<float *>(s), *cv = reinterpret_cast<float *>(c);
#pragma unroll
      for (int j = 0; j < DFT_Q; ++j) {
        const double2 phv = ph[fb + j];
        const R rr = R(phv.x), ri = R(phv.y);
        R tr, ti;
        if (a.f1) {
          tr = rr * f0 - ri * f1;
          ti = rr * f1 + ri * f0;
        }
        else {
          tr = f0 * rr;
          ti = f0 * ri;
        }
        const float yr = float(tr) - cv[2 * j], yi = float(ti) - cv[2 * j + 1];
        const float sr = sv[2 * j] + yr, si = sv[2 * j + 1] + yi;
        cv[2 * j] = (sr - sv[2 * j]) - yr;
        cv[2 * j + 1] = (si - sv[2 * j + 1]) - yi;
        sv[2 * j] = sr;
        sv[2 * j + 1] = si;
      }
      acc[0] = s[0];
      acc[1] = s[1];
      cp[0] = c[0];
      cp[1] = c[1];
      continue;
    }
    const int cnt = min(DFT_Q, nf - fb);
    for (int j = 0; j < cnt; ++j) {
      const double2 phv = tab ? ph[fb + j] : dft_phase(a.omega[fb + j], a.time, a.scale_re, a.scale_im);
      dft_acc(a, o0 + size_t(j), phv, f0, f1);
    }
  }
}

// one step's value of every point of a chunk, into its deferred-step slot
template <class R> __global__ __launch_bounds__(BLOCK) void k_dft_snap(const DftArgs<R> a, R *s0, R *s1) {
  const int k = int(blockIdx.x) * BLOCK + int(threadIdx.x);
  if (k >= a.npts) return;
  R f0, f1;
  dft_point(a, k, f0, f1);
  s0[k] = f0;
  if (s1) s1[k] = f1;
}
// the deferred steps of a chunk: a thread per point and DFT_Q consecutive frequencies loads
// each accumulator once, applies the steps' terms in step order, and stores it once
template <class R> __global__ __launch_bounds__(BLOCK) void k_dft_flush(const DftArgs<R> a, const DftDefer<R> d) {
  __shared__ double2 ph[DFT_FLUSH_TAB];
  __shared__ double tt[DFT_DEFER];
  const int nf = a.nfreq, ns = d.nsteps;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < DFT_DEFER; ++j) tt[j] = d.time[j];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < ns * nf; q += BLOCK) {
    const int j = q / nf;
    ph[q] = dft_phase(a.omega[q - j * nf], tt[j], a.scale_re, a.scale_im);
  }
  __syncthreads();
  const uint32_t nq = uint32_t(nf + DFT_Q - 1) / DFT_Q;
  const uint32_t n = uint32_t(a.npts) * nq;
  for (uint32_t o = blockIdx.x * BLOCK + threadIdx.x; o < n; o += gridDim.x * BLOCK) {
    const uint32_t k = o / nq;
    const int fb = int(o - k * nq) * DFT_Q;
    const int cnt = min(DFT_Q, nf - fb);
    R f0[DFT_DEFER], f1[DFT_DEFER];
#pragma unroll
    for (int j = 0; j < DFT_DEFER; ++j)
      if (j < ns) {
        f0[j] = d.snap0[size_t(j) * a.npts + k];
        f1[j] = d.snap1 ? d.snap1[size_t(j) * a.npts + k] : R(0);
      }
    const size_t o0 = size_t(k) * size_t(nf) + size_t(fb);
    for (int m = 0; m < cnt; ++m) {
      DftAcc<R> v;
      dft_load(a, o0 + m, v);
#pragma unroll
      for (int j = 0; j < DFT_DEFER; ++j)
        if (j < ns) dft_add(a, v, ph[j * nf + fb + m], f0[j], f1[j]);
      dft_store(a, o0 + m, v);
    }
  }
}

template <class R>
__global__ void k_check(const R *const *ptrs, const int64_t *idx, const double *w, int n, int *flag) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const R v = ptrs[t][idx[t]];
  if (w[t] != 0.0 && !isfinite(double(v))) atomicOr(flag, 1);
}


template <class R> void launch_step(const StepArgs<R> &a, bool curl, bool upd, hipStream_t s) {
  const size_t count = size_t(a.g.an[0]) * size_t(a.g.an[1]) * size_t(a.g.an[2]);
  const unsigned nb = nblocks8(count);
  const uint32_t c32 = uint32_t(count);
  if (curl && upd)
    hipLaunchKernelGGL((k_step<R, true, true>), dim3(nb), dim3(BLOCK), 0, s, a, c32);
  else if (curl)
    hipLaunchKernelGGL((k_step<R, true, false>), dim3(nb), dim3(BLOCK), 0, s, a, c32);
  else if (upd)
    hipLaunchKernelGGL((k_step<R, false, true>), dim3(nb), dim3(BLOCK), 0, s, a, c32);
}

// groups of V points along the rows of a's launch box: first group, groups per row, threads
template <int V> static inline void fast_groups(const Grid &g, int &g0, int &ngroups, uint32_t &count) {
  g0 = g.alo[2] / V;
  ngroups = (g.alo[2] + g.an[2] - 1) / V - g0 + 1;
  const size_t n = size_t(ngroups) * size_t(g.an[0]) * size_t(g.an[1]);
  if (n >= (size_t(1) << 32)) throw std::runtime_error("meep: launch box too large for one launch");
  count = uint32_t(n);
}
template <class R, int V, bool PML, int NP> void launch_fast_np(const StepArgs<R> &a, bool curl, bool upd, hipStream_t s) {
  int g0, ngroups;
  uint32_t c32;
  fast_groups<V>(a.g, g0, ngroups, c32);
  const unsigned nb = upd ? (3u * nblocks(c32) + 7u) / 8u * 8u : nblocks8(c32); // with an update, one slot per block
  if (curl && upd)
    hipLaunchKernelGGL((k_fast<R, V, true, true, PML, NP>), dim3(nb), dim3(BLOCK), 0, s, a, c32, g0, ngroups);
  else if (curl)
    hipLaunchKernelGGL((k_fast<R, V, true, false, PML, NP>), dim3(nb), dim3(BLOCK), 0, s, a, c32, g0, ngroups);
  else if (upd)
    hipLaunchKernelGGL((k_fast<R, V, false, true, PML, NP>), dim3(nb), dim3(BLOCK), 0, s, a, c32, g0, ngroups);
}
// The pole-carrying updates hold P, P_prev and sigma of every pole in registers: in FP32 they
// take 8-byte groups (V = 2) to keep the occupancy of the 16-byte form's plain kernels.
template <class R, bool PML> void launch_fast_t(const StepArgs<R> &a, bool curl, bool upd, hipStream_t s) {
  constexpr int V = 16 / sizeof(R), VP = V > 2 ? 2 : V;
  if (!upd || a.npole == 0) {
    launch_fast_np<R, V, PML, 0>(a, curl, upd, s);
    return;
  }
  switch (a.npole) {
    case 1: launch_fast_np<R, VP, PML, 1>(a, curl, upd, s); break;
    case 2: launch_fast_np<R, VP, PML, 2>(a, curl, upd, s); break;
    default: launch_fast_np<R, VP, PML, MAXPOLE>(a, curl, upd, s); break;
  }
}
template <class R> void launch_fast(const StepArgs<R> &a, bool curl, bool upd, bool pml, hipStream_t s) {
  if (pml)
    launch_fast_t<R, true>(a, curl, upd, s);
  else
    launch_fast_t<R, false>(a, curl, upd, s);
}

template <class R>
void launch_conn(const Conn *conns, const double2 *phases, int n, R *const *slots, int ncmp, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_conn<R>, dim3(nblocks(n)), dim3(BLOCK), 0, s, conns, phases, n, slots, ncmp);
}

template <class R>
void launch_src(R *f, const int64_t *idx, const double2 *amp, int n, double2 scal, double mult, int cmp, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_src<R>, dim3(nblocks(n)), dim3(BLOCK), 0, s, f, idx, amp, n, scal, mult, cmp);
}

template <class R>
__global__ __launch_bounds__(BLOCK) void k_src_images(R *f, const int64_t *idx, const int *k, const int64_t *dst, const int8_t *sgn, int n) {
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= n) return;
  const R v = f[idx[k[t]]];
  f[dst[t]] = sgn[t] > 0 ? v : -v;
}
template <class R>
void launch_src_images(R *f, const int64_t *idx, const int *k, const int64_t *dst, const int8_t *sgn, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_src_images<R>, dim3(nblocks(n)), dim3(BLOCK), 0, s, f, idx, k, dst, sgn, n);
}
template <class R> __global__ __launch_bounds__(BLOCK) void k_average(R *f, const R *b, size_t n) {
  const size_t t = size_t(blockIdx.x) * BLOCK + threadIdx.x;
  if (t < n) f[t] = R(0.5 * double(f[t] + b[t]));
}
template <class R> void launch_average(R *f, const R *b, size_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_average<R>, dim3(nblocks(n)), dim3(BLOCK), 0, s, f, b, n);
}
template <class R> __global__ __launch_bounds__(BLOCK) void k_widen(double *dst, const R *src, size_t n) {
  const size_t t = size_t(blockIdx.x) * BLOCK + threadIdx.x;
  if (t < n) dst[t] = double(src[t]);
}
template <class R> void launch_widen(double *dst, const R *src, size_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_widen<R>, dim3(nblocks(n)), dim3(BLOCK), 0, s, dst, src, n);
}
template <class R> __global__ __launch_bounds__(BLOCK) void k_gather(R *dst, const R *f, const int64_t *idx, int n) {
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t < n) dst[t] = f[idx[t]];
}
template <class R> __global__ __launch_bounds__(BLOCK) void k_scatter(R *f, const int64_t *idx, const R *src, int n) {
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t < n) f[idx[t]] = src[t];
}
template <class R> void launch_gather(R *dst, const R *f, const int64_t *idx, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_gather<R>, dim3(nblocks(n)), dim3(BLOCK), 0, s, dst, f, idx, n);
}
template <class R> void launch_scatter(R *f, const int64_t *idx, const R *src, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_scatter<R>, dim3(nblocks(n)), dim3(BLOCK), 0, s, f, idx, src, n);
}

template <class R> void launch_dft_snap(const DftArgs<R> &a, R *snap0, R *snap1, hipStream_t s) {
  if (a.npts > 0) hipLaunchKernelGGL(k_dft_snap<R>, dim3(nblocks(size_t(a.npts))), dim3(BLOCK), 0, s, a, snap0, snap1);
}
template <class R> void launch_dft_flush(const DftArgs<R> &a, const DftDefer<R> &d, hipStream_t s) {
  if (d.nsteps < 1 || d.nsteps > DFT_DEFER || d.nsteps * a.nfreq > DFT_FLUSH_TAB)
    throw std::runtime_error("meep: internal: bad deferred DFT flush");
  const size_t n = size_t(a.npts) * size_t((a.nfreq + DFT_Q - 1) / DFT_Q);
  if (n >= (size_t(1) << 31)) throw std::runtime_error("meep: DFT chunk too large for one launch");
  const unsigned nb = std::min<unsigned>(nblocks(n), 4096u);
  if (n > 0) hipLaunchKernelGGL(k_dft_flush<R>, dim3(nb), dim3(BLOCK), 0, s, a, d);
}
template <class R> void launch_dft(const DftArgs<R> &a, hipStream_t s) {
  const size_t n = size_t(a.npts) * size_t((a.nfreq + DFT_Q - 1) / DFT_Q);
  if (n >= (size_t(1) << 31)) throw std::runtime_error("meep: DFT chunk too large for one launch");
  // enough blocks to fill the chip several times over; each then walks a few strides
  const unsigned nb = std::min<unsigned>(nblocks(n), 4096u);
  if (n > 0) hipLaunchKernelGGL(k_dft<R>, dim3(nb), dim3(BLOCK), 0, s, a);
}

template <class R>
void launch_check(const R *const *ptrs, const int64_t *idx, const double *w, int n, int *flag, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_check<R>, dim3(nblocks(n)), dim3(BLOCK), 0, s, ptrs, idx, w, n, flag);
}

// a field's value at a centred point: its four Yee points summed in realnum, then halved twice
// (the reference's 0.25 * (f + f + f + f), rounded back to realnum as its fvals are)
template <class R> __device__ __forceinline__ double flux_avg(const R *p, int64_t i, int64_t o1, int64_t o2) {
  if (!p) return 0.0;
  const R s = p[i] + p[i + o1] + p[i + o2] + p[i + o1 + o2];
  return double(R(0.25 * double(s)));
}
template <class R> __global__ __launch_bounds__(BLOCK) void k_flux_plane(FluxPlaneArgs<R> a, double *partial) {
  __shared__ double red[BLOCK];
  double acc = 0.0;
  for (int j = blockIdx.x * BLOCK + threadIdx.x; j < a.n; j += gridDim.x * BLOCK) {
    const int64_t i = a.idx[j];
    double v = 0.0;
    for (int p = 0; p < a.npair; ++p) {
      const double er = flux_avg(a.e[p][0], i, a.eo[p][0], a.eo[p][1]), ei = flux_avg(a.e[p][1], i, a.eo[p][0], a.eo[p][1]);
      const double hr = flux_avg(a.h[p][0], i, a.ho[p][0], a.ho[p][1]), hi = flux_avg(a.h[p][1], i, a.ho[p][0], a.ho[p][1]);
      v += a.sign[p] * (er * hr + ei * hi);
    }
    acc += v * a.w[j];
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = BLOCK / 2; h > 0; h >>= 1) {
    if (int(threadIdx.x) < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}
template <class R> void launch_flux_plane(const FluxPlaneArgs<R> &a, double *partial, hipStream_t s) {
  hipLaunchKernelGGL(k_flux_plane<R>, dim3(FLUX_BLOCKS), dim3(BLOCK), 0, s, a, partial);
}

// ---- batched launches -----------------------------------------------------------------
// Logical blocks are dealt to XCDs in contiguous runs as in the single-chunk kernels (nblk is a
// multiple of 8); each logical block belongs to one item.
// KIND: 0 generic, 1 vectorised, 2 vectorised with per-point uPML, 3 and 4 the same two with
// the Lorentz/Drude poles
template <class R, int BS, int KIND, bool CURL, bool UPD>
__global__ __launch_bounds__(BS) void k_step_batch(const BatchItem<R> *items, const uint32_t *blk_item, const uint32_t *blk_first,
                                                   const double *vrow) {
  const uint32_t nb = gridDim.x, b = blockIdx.x;
  const uint32_t lb = (b & 7u) * (nb >> 3) + (b >> 3);
  const uint32_t k = blk_item[lb];
  const BatchItem<R> &it = items[k];
  const uint32_t t = blk_first[lb] + threadIdx.x;
  if (t >= it.count) return;
  const double *row = vrow + it.a.vofs;
  if (KIND == 1)
    fast_body<R, 16 / sizeof(R), CURL, UPD, false, 0>(it.a, t, it.g0, it.ngroups, row);
  else if (KIND == 2)
    fast_body<R, 16 / sizeof(R), CURL, UPD, true, 0>(it.a, t, it.g0, it.ngroups, row);
  else if (KIND == 3)
    fast_body<R, 16 / sizeof(R), CURL, UPD, false, MAXPOLE>(it.a, t, it.g0, it.ngroups, row);
  else if (KIND == 4)
    fast_body<R, 16 / sizeof(R), CURL, UPD, true, MAXPOLE>(it.a, t, it.g0, it.ngroups, row);
  else
    step_body<R, CURL, UPD>(it.a, t, row);
}

template <class R, int BS, int KIND>
void launch_step_batch_k(const BatchItem<R> *items, const uint32_t *blk_item, const uint32_t *blk_first, unsigned nblk, bool curl,
                         bool upd, const double *vrow, hipStream_t s) {
  const dim3 grid(nblk), blk(BS);
  if (curl && upd)
    hipLaunchKernelGGL((k_step_batch<R, BS, KIND, true, true>), grid, blk, 0, s, items, blk_item, blk_first, vrow);
  else if (curl)
    hipLaunchKernelGGL((k_step_batch<R, BS, KIND, true, false>), grid, blk, 0, s, items, blk_item, blk_first, vrow);
  else
    hipLaunchKernelGGL((k_step_batch<R, BS, KIND, false, true>), grid, blk, 0, s, items, blk_item, blk_first, vrow);
}
template <class R, int BS>
void launch_step_batch_bs(const BatchItem<R> *items, const uint32_t *blk_item, const uint32_t *blk_first, unsigned nblk, int kind,
                          bool curl, bool upd, const double *vrow, hipStream_t s) {
  if (kind == 4)
    launch_step_batch_k<R, BS, 4>(items, blk_item, blk_first, nblk, curl, upd, vrow, s);
  else if (kind == 3)
    launch_step_batch_k<R, BS, 3>(items, blk_item, blk_first, nblk, curl, upd, vrow, s);
  else if (kind == 2)
    launch_step_batch_k<R, BS, 2>(items, blk_item, blk_first, nblk, curl, upd, vrow, s);
  else if (kind == 1)
    launch_step_batch_k<R, BS, 1>(items, blk_item, blk_first, nblk, curl, upd, vrow, s);
  else
    launch_step_batch_k<R, BS, 0>(items, blk_item, blk_first, nblk, curl, upd, vrow, s);
}
template <class R>
void launch_step_batch(const BatchItem<R> *items, const uint32_t *blk_item, const uint32_t *blk_first, unsigned nblk, int kind,
                       bool curl, bool upd, int block, const double *vrow, hipStream_t s) {
  if (!nblk) return;
  switch (block) {
    case 64: launch_step_batch_bs<R, 64>(items, blk_item, blk_first, nblk, kind, curl, upd, vrow, s); break;
    case 128: launch_step_batch_bs<R, 128>(items, blk_item, blk_first, nblk, kind, curl, upd, vrow, s); break;
    case 256: launch_step_batch_bs<R, 256>(items, blk_item, blk_first, nblk, kind, curl, upd, vrow, s); break;
    default: launch_step_batch_bs<R, 512>(items, blk_item, blk_first, nblk, kind, curl, upd, vrow, s); break;
  }
}

template <class R>
__global__ __launch_bounds__(BLOCK) void k_dft_batch(const BatchDft<R> *items, const uint32_t *blk_item, const uint32_t *blk_first, long step) {
  __shared__ double2 ph[DFT_TAB];
  const BatchDft<R> &it = items[blk_item[blockIdx.x]];
  const int64_t tt = it.t_start + step + 1;
  if (tt % it.decimation) return; // the whole block belongs to this chunk
  const DftArgs<R> &a = it.a;
  const double time = double(tt) * it.dt - (it.half ? 0.5 * it.dt : 0.0);
  const int nf = a.nfreq;
  if (nf > DFT_TAB) return; // the host keeps such chunks out of batches
  for (int q = threadIdx.x; q < nf; q += BLOCK) ph[q] = dft_phase(a.omega[q], time, a.scale_re, a.scale_im);
  __syncthreads();
  const size_t o = size_t(blk_first[blockIdx.x]) + threadIdx.x;
  if (o >= size_t(a.npts) * size_t(nf)) return;
  dft_one<R>(a, o, ph[int(o % size_t(nf))]);
}
template <class R>
__global__ __launch_bounds__(BLOCK) void k_dft_snap_batch(const BatchDft<R> *items, const uint32_t *blk_item, const uint32_t *blk_first,
                                                        int slot) {
  const BatchDft<R> &it = items[blk_item[blockIdx.x]];
  const uint32_t p = blk_first[blockIdx.x] + threadIdx.x; // padding blocks start past every item
  if (p >= uint32_t(it.a.npts)) return;
  R f0, f1;
  dft_point(it.a, int(p), f0, f1);
  const size_t o = size_t(slot) * size_t(it.a.npts) + p;
  it.snap0[o] = f0;
  if (it.snap1) it.snap1[o] = f1;
}
template <class R>
__global__ __launch_bounds__(BLOCK) void k_dft_flush_batch(const BatchDft<R> *items, const uint32_t *blk_item, const uint32_t *blk_first,
                                                         long first, int ns) {
  __shared__ double2 ph[DFT_FLUSH_TAB];
  const BatchDft<R> &it = items[blk_item[blockIdx.x]];
  const DftArgs<R> &a = it.a;
  const int nf = a.nfreq;
  if (nf * ns > DFT_FLUSH_TAB) return; // the host keeps such chunks out of deferral
  for (int q = threadIdx.x; q < ns * nf; q += BLOCK) {
    const int j = q / nf;
    const int64_t tt = it.t_start + first + j + 1;
    const double time = double(tt) * it.dt - (it.half ? 0.5 * it.dt : 0.0);
    ph[q] = dft_phase(a.omega[q - j * nf], time, a.scale_re, a.scale_im);
  }
  __syncthreads();
  const uint32_t nq = uint32_t(nf + DFT_Q - 1) / DFT_Q;
  const uint32_t o = blk_first[blockIdx.x] + threadIdx.x;
  if (o >= uint32_t(a.npts) * nq) return;
  const uint32_t k = o / nq;
  const int fb = int(o - k * nq) * DFT_Q;
  const int cnt = min(DFT_Q, nf - fb);
  R f0[DFT_DEFER], f1[DFT_DEFER];
#pragma unroll
  for (int j = 0; j < DFT_DEFER; ++j)
    if (j < ns) {
      f0[j] = it.snap0[size_t(j) * a.npts + k];
      f1[j] = it.snap1 ? it.snap1[size_t(j) * a.npts + k] : R(0);
    }
  const size_t o0 = size_t(k) * size_t(nf) + size_t(fb);
  for (int m = 0; m < cnt; ++m) {
    DftAcc<R> v;
    dft_load(a, o0 + m, v);
#pragma unroll
    for (int j = 0; j < DFT_DEFER; ++j)
      if (j < ns) dft_add(a, v, ph[j * nf + fb + m], f0[j], f1[j]);
    dft_store(a, o0 + m, v);
  }
}
template <class R>
void launch_dft_snap_batch(const BatchDft<R> *items, const uint32_t *blk_item, const uint32_t *blk_first, unsigned nblk, int slot,
                           hipStream_t s) {
  if (slot < 0 || slot >= DFT_DEFER) throw std::runtime_error("meep: internal: bad deferred DFT slot");
  if (nblk) hipLaunchKernelGGL(k_dft_snap_batch<R>, dim3(nblk), dim3(BLOCK), 0, s, items, blk_item, blk_first, slot);
}
template <class R>
void launch_dft_flush_batch(const BatchDft<R> *items, const uint32_t *blk_item, const uint32_t *blk_first, unsigned nblk, long first,
                            int nsteps, hipStream_t s) {
  if (nsteps < 1 || nsteps > DFT_DEFER) throw std::runtime_error("meep: internal: bad deferred DFT flush");
  if (nblk) hipLaunchKernelGGL(k_dft_flush_batch<R>, dim3(nblk), dim3(BLOCK), 0, s, items, blk_item, blk_first, first, nsteps);
}
template <class R>
void launch_dft_batch(const BatchDft<R> *items, const uint32_t *blk_item, const uint32_t *blk_first, unsigned nblk, long step,
                      hipStream_t s) {
  if (nblk) hipLaunchKernelGGL(k_dft_batch<R>, dim3(nblk), dim3(BLOCK), 0, s, items, blk_item, blk_first, step);
}

template <class R>
__global__ __launch_bounds__(BLOCK) void k_conn_batch(const BatchConn *items, const uint32_t *blk_item, const uint32_t *blk_first) {
  const BatchConn &it = items[blk_item[blockIdx.x]];
  const uint32_t t = blk_first[blockIdx.x] + threadIdx.x; // padding blocks start past every item
  if (t >= uint32_t(it.n)) return;
  conn_one<R>(it.conns, it.phases, int(t), reinterpret_cast<R *const *>(it.slots), it.ncmp);
}
template <class R>
void launch_conn_batch(const BatchConn *items, const uint32_t *blk_item, const uint32_t *blk_first, unsigned nblk, hipStream_t s) {
  if (nblk) hipLaunchKernelGGL(k_conn_batch<R>, dim3(nblk), dim3(BLOCK), 0, s, items, blk_item, blk_first);
}

#define INSTANTIATE_BATCH(R)                                                                       \
  template void launch_step_batch<R>(const BatchItem<R> *, const uint32_t *, const uint32_t *, unsigned, int, bool, bool, int, \
                                     const double *, hipStream_t);                                 \
  template void launch_dft_batch<R>(const BatchDft<R> *, const uint32_t *, const uint32_t *, unsigned, long, hipStream_t); \
  template void launch_dft_snap_batch<R>(const BatchDft<R> *, const uint32_t *, const uint32_t *, unsigned, int, hipStream_t); \
  template void launch_dft_flush_batch<R>(const BatchDft<R> *, const uint32_t *, const uint32_t *, unsigned, long, int, hipStream_t); \
  template void launch_conn_batch<R>(const BatchConn *, const uint32_t *, const uint32_t *, unsigned, hipStream_t);
#define INSTANTIATE(R)                                                                             \
  template void launch_step<R>(const StepArgs<R> &, bool, bool, hipStream_t);                      \
  template void launch_fast<R>(const StepArgs<R> &, bool, bool, bool, hipStream_t);                \
  template bool launch_row3<R>(const StepArgs<R> &, bool, bool, hipStream_t);                      \
  template void launch_conn<R>(const Conn *, const double2 *, int, R *const *, int, hipStream_t);  \
  template void launch_src<R>(R *, const int64_t *, const double2 *, int, double2, double, int,      \
                              hipStream_t);                                                        \
  template void launch_average<R>(R *, const R *, size_t, hipStream_t);                            \
  template void launch_widen<R>(double *, const R *, size_t, hipStream_t);                         \
  template void launch_src_images<R>(R *, const int64_t *, const int *, const int64_t *, const int8_t *, int, \
                                     hipStream_t);                                                 \
  template void launch_gather<R>(R *, const R *, const int64_t *, int, hipStream_t);               \
  template void launch_scatter<R>(R *, const int64_t *, const R *, int, hipStream_t);              \
  template void launch_dft<R>(const DftArgs<R> &, hipStream_t);                                    \
  template void launch_dft_snap<R>(const DftArgs<R> &, R *, R *, hipStream_t);                     \
  template void launch_dft_flush<R>(const DftArgs<R> &, const DftDefer<R> &, hipStream_t);         \
  template void launch_check<R>(const R *const *, const int64_t *, const double *, int, int *,     \
                                hipStream_t);                                                      \
  template void launch_flux_plane<R>(const FluxPlaneArgs<R> &, double *, hipStream_t);
// The build compiles this file once per field precision (MX_REAL) and part (MX_PART 1: the
// single-chunk launchers, 2: the batched ones), in parallel.
#ifdef MX_REAL
#if MX_PART != 2
INSTANTIATE(MX_REAL)
#endif
#if MX_PART != 1
INSTANTIATE_BATCH(MX_REAL)
#endif
#else
INSTANTIATE(float)
INSTANTIATE(double)
INSTANTIATE_BATCH(float)
INSTANTIATE_BATCH(double)
#endif

} // namespace mx
