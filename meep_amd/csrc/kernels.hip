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

#include <hip/hip_runtime.h>

namespace mx {

constexpr int BLOCK = 256;

// logical block index, dealt to XCDs in contiguous runs (gridDim.x is a multiple of 8);
// rev walks them in descending order
__device__ __forceinline__ uint32_t xcd_block(int rev = 0) {
  const uint32_t nb = gridDim.x, b = blockIdx.x;
  const uint32_t lb = (b & 7u) * (nb >> 3) + (b >> 3);
  return rev ? nb - 1 - lb : lb;
}

// halo point along the stride-1 direction fed by the owned point z of a component with Yee
// shift ysh there; -1 when z is not the wrapped end
__device__ __forceinline__ int zwrap_dst(int ysh, int z, int n) {
  if (ysh == 0) return z == n ? 0 : -1;
  return z == 0 ? n : -1;
}

__device__ __forceinline__ int jsel(const int j[3], int d) { return d == 0 ? j[0] : d == 1 ? j[1] : j[2]; }
__device__ __forceinline__ int ysel(const int y[3], int d) { return d == 0 ? y[0] : d == 1 ? y[1] : y[2]; }
template <class R> __device__ __forceinline__ const R *psel(const R *const p[3], int d) {
  return d == 0 ? p[0] : d == 1 ? p[1] : p[2];
}

// on_metal_boundary for an owned point (src/boundaries.cpp:191-205): owned points reach the
// High wall only with no Yee shift along d, and the Low magnetic plane only with one.
__device__ __forceinline__ bool on_metal(const Grid &g, const int j[3], const int ysh[3]) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (((g.hi_metal >> d) & 1) && ysh[d] == 0 && j[d] == g.num[d]) return true;
    if (((g.lo_magnetic >> d) & 1) && ysh[d] == 1 && j[d] == 0) return true;
  }
  return false;
}

template <class R> __device__ __forceinline__ bool pml_active(const PmlProfile<R> &p, int d, int k) {
  return psel(p.sig, d)[k] != R(0) || psel(p.kap, d)[k] != R(1);
}

__device__ __forceinline__ bool owned(const int j[3], const int lo[3], const int hi[3]) {
  return j[0] >= lo[0] && j[0] <= hi[0] && j[1] >= lo[1] && j[1] <= hi[1] && j[2] >= lo[2] && j[2] <= hi[2];
}

__device__ __forceinline__ bool in_tab(const SrcTab &t, uint32_t i) {
  return t.n && int64_t(i) >= t.lo && int64_t(i) <= t.hi;
}
template <class R> __device__ __forceinline__ R apply_tab(const SrcTab &t, int c, uint32_t i, R v) {
#pragma unroll
  for (int s = 0; s < SRC_TAB; ++s)
    if (s < t.n && t.comp[s] == c && t.idx[s] == int64_t(i)) v = R(double(v) - t.val[s]);
  return v;
}

// ---------------------------------------------------------------------------------------
// generic kernel: one point per thread, full uPML logic, over the launch box

template <class R, bool CURL, bool UPD>
__device__ __forceinline__ void cell_update(const StepArgs<R> &a, int c, uint32_t i, const int j[3], R cu) {
  R fv = R(0);
  if (CURL) {
    const CurlComp<R> &k = a.cc[c];
    fv = k.f[i];
    const int kk = k.dsig >= 0 ? 2 * jsel(j, k.dsig) + ysel(k.ysh, k.dsig) : 0;
    const int ku = k.dsigu >= 0 ? 2 * jsel(j, k.dsigu) + ysel(k.ysh, k.dsigu) : 0;
    const bool pml = k.dsig >= 0 && pml_active(a.pml, k.dsig, kk);
    const bool pmlu = k.dsigu >= 0 && pml_active(a.pml, k.dsigu, ku);
    if (!pmlu) {
      if (!pml)
        fv -= k.dtdx * cu;
      else
        fv = ((psel(a.pml.kap, k.dsig)[kk] - psel(a.pml.sig, k.dsig)[kk]) * fv - k.dtdx * cu) * psel(a.pml.siginv, k.dsig)[kk];
    }
    else {
      const R fprev = k.fu[i];
      R fun;
      if (!pml)
        fun = fprev - k.dtdx * cu;
      else
        fun = ((psel(a.pml.kap, k.dsig)[kk] - psel(a.pml.sig, k.dsig)[kk]) * fprev - k.dtdx * cu) * psel(a.pml.siginv, k.dsig)[kk];
      k.fu[i] = fun;
      fv = psel(a.pml.siginv, k.dsigu)[ku] * ((psel(a.pml.kap, k.dsigu)[ku] - psel(a.pml.sig, k.dsigu)[ku]) * fv + fun - fprev);
    }
    if (in_tab(a.cur, i)) fv = apply_tab(a.cur, c, i, fv);
    if (on_metal(a.g, j, k.ysh)) fv = R(0);
    k.f[i] = fv;
    if (a.zwrap) {
      const int d2 = a.g.axdir[2], z = jsel(j, d2);
      const int zd = zwrap_dst(ysel(k.ysh, d2), z, a.g.num[d2]);
      if (zd >= 0) k.f[i - z + zd] = a.zsign > 0 ? fv : -fv;
    }
  }
  if (UPD && a.uc[c].f) {
    const UpdComp<R> &u = a.uc[c];
    R gv = CURL ? fv : u.g[i];
#pragma unroll
    for (int p = 0; p < MAXPOLE; ++p)
      if (p < a.npole && u.P[p]) gv -= u.P[p][i];
    if (in_tab(a.dip, i)) gv = apply_tab(a.dip, c, i, gv);
    R ev, wv;
    const int kw = u.dsigw >= 0 ? 2 * jsel(j, u.dsigw) + ysel(u.ysh, u.dsigw) : 0;
    if (u.dsigw >= 0 && pml_active(a.pml, u.dsigw, kw)) {
      const R kapw = psel(a.pml.kap, u.dsigw)[kw], sigw = psel(a.pml.sig, u.dsigw)[kw];
      const R fwprev = u.fw[i];
      const R fwn = u.u ? gv * u.u[i] : gv;
      u.fw[i] = fwn;
      ev = u.f[i] + ((kapw + sigw) * fwn - (kapw - sigw) * fwprev);
      wv = fwn;
    }
    else {
      ev = u.u ? gv * u.u[i] : gv;
      wv = ev;
    }
#pragma unroll
    for (int p = 0; p < MAXPOLE; ++p)
      if (p < a.npole && u.P[p] && u.sigma[p]) {
        const PoleCoef<R> &q = a.pole[p];
        const R pcur = u.P[p][i];
        const R pn = q.gamma1inv * (pcur * (R(2) - q.omega0dtsqr_denom) - q.gamma1 * u.Pp[p][i] +
                                    q.omega0dtsqr * (u.sigma[p][i] * wv));
        u.P[p][i] = pn;
        u.Pp[p][i] = pcur;
        if (a.zwrap) {
          const int d2 = a.g.axdir[2], z = jsel(j, d2);
          const int zd = zwrap_dst(ysel(u.ysh, d2), z, a.g.num[d2]);
          if (zd >= 0) u.P[p][i - z + zd] = a.zsign > 0 ? pn : -pn;
        }
      }
    if (on_metal(a.g, j, u.ysh)) ev = R(0);
    u.f[i] = ev;
    if (a.zwrap) {
      const int d2 = a.g.axdir[2], z = jsel(j, d2);
      const int zd = zwrap_dst(ysel(u.ysh, d2), z, a.g.num[d2]);
      if (zd >= 0) u.f[i - z + zd] = a.zsign > 0 ? ev : -ev;
    }
  }
}

template <class R, bool CURL, bool UPD>
__global__ __launch_bounds__(BLOCK) void k_step(const StepArgs<R> a, uint32_t count) {
  const uint32_t t = xcd_block() * uint32_t(BLOCK) + threadIdx.x;
  if (t >= count) return;
  const Grid &g = a.g;
  int e[3], j[3] = {0, 0, 0};
#pragma unroll
  for (int k = 0; k < 3; ++k) e[k] = g.axdir[k] >= 0 ? g.bhi[g.axdir[k]] - g.blo[g.axdir[k]] + 1 : 1;
  const uint32_t r = t / uint32_t(e[2]);
  const int q2 = int(t - r * uint32_t(e[2]));
  const int q0 = int(r / uint32_t(e[1]));
  const int q1 = int(r - uint32_t(q0) * uint32_t(e[1]));
  if (g.axdir[0] >= 0) j[g.axdir[0]] = g.blo[g.axdir[0]] + q0;
  if (g.axdir[1] >= 0) j[g.axdir[1]] = g.blo[g.axdir[1]] + q1;
  if (g.axdir[2] >= 0) j[g.axdir[2]] = g.blo[g.axdir[2]] + q2;
  const uint32_t i = uint32_t(j[0] * g.st[0] + j[1] * g.st[1] + j[2] * g.st[2]);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (c >= a.nc) break;
    if (CURL ? !a.cc[c].f : !a.uc[c].f) continue; // no component along this direction
    const int *lo = CURL ? a.cc[c].lo : a.uc[c].lo;
    const int *hi = CURL ? a.cc[c].hi : a.uc[c].hi;
    if (!owned(j, lo, hi)) continue;
    R cu = R(0);
    if (CURL) {
      const CurlComp<R> &k = a.cc[c];
      if (k.g2)
        cu = k.g1[i + k.s1] - k.g1[i] + k.g2[i] - k.g2[i + k.s2];
      else if (k.g1)
        cu = k.g1[i + k.s1] - k.g1[i];
    }
    cell_update<R, CURL, UPD>(a, c, i, j, cu);
  }
}

// ---------------------------------------------------------------------------------------
// vectorised kernel: V consecutive points of a padded row per thread, no uPML

template <class R, int V> struct Vec {
  R v[V];
};
template <class R, int V> __device__ __forceinline__ Vec<R, V> vload(const R *p) {
  static_assert(V * sizeof(R) == 16, "16-byte groups");
  Vec<R, V> r;
  const float4 q = *reinterpret_cast<const float4 *>(p);
  __builtin_memcpy(r.v, &q, 16);
  return r;
}
template <class R, int V> __device__ __forceinline__ void vstore(R *p, const Vec<R, V> &x, unsigned mask) {
  if (mask == (1u << V) - 1) {
    float4 q;
    __builtin_memcpy(&q, x.v, 16);
    *reinterpret_cast<float4 *>(p) = q;
  }
  else {
#pragma unroll
    for (int k = 0; k < V; ++k)
      if (mask & (1u << k)) p[k] = x.v[k];
  }
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

// the periodic image of an owned end point of the row, written into the row's halo point
template <class R, int V>
__device__ __forceinline__ void wrap_store(R *f, uint32_t i, int z0, int ysh, int n, int sign, const Vec<R, V> &x, unsigned mask) {
#pragma unroll
  for (int k = 0; k < V; ++k)
    if ((mask >> k) & 1) {
      const int zd = zwrap_dst(ysh, z0 + k, n);
      if (zd >= 0) f[i - z0 + zd] = sign > 0 ? x.v[k] : -x.v[k];
    }
}

// metal points of a component along the row of a group (the uniform directions d0, d1 give
// all or nothing)
__device__ __forceinline__ unsigned metal_mask(const Grid &g, const int j[3], const int ysh[3], int d0, int d1, int d2, int z0, int V) {
  auto hit = [&](int d, int jd) {
    return (((g.hi_metal >> d) & 1) && ysh[d] == 0 && jd == g.num[d]) || (((g.lo_magnetic >> d) & 1) && ysh[d] == 1 && jd == 0);
  };
  if ((d0 >= 0 && hit(d0, j[d0])) || (d1 >= 0 && hit(d1, j[d1]))) return (1u << V) - 1;
  unsigned m = 0;
  for (int k = 0; k < V; ++k)
    if (hit(d2, z0 + k)) m |= 1u << k;
  return m;
}

template <class R, int V, bool CURL, bool UPD>
__global__ __launch_bounds__(BLOCK) void k_fast(const StepArgs<R> a, uint32_t count, int g0, int ngroups) {
  const uint32_t t = xcd_block(a.rev) * uint32_t(BLOCK) + threadIdx.x;
  if (t >= count) return;
  const Grid &g = a.g;
  const int d0 = g.axdir[0], d1 = g.axdir[1], d2 = g.axdir[2];
  const int e1 = d1 >= 0 ? g.bhi[d1] - g.blo[d1] + 1 : 1;
  const uint32_t r = t / uint32_t(ngroups);
  const int grp = int(t - r * uint32_t(ngroups));
  const int q0 = int(r / uint32_t(e1));
  const int q1 = int(r - uint32_t(q0) * uint32_t(e1));
  int j[3] = {0, 0, 0};
  if (d0 >= 0) j[d0] = g.blo[d0] + q0;
  if (d1 >= 0) j[d1] = g.blo[d1] + q1;
  const int z0 = (g0 + grp) * V;
  const uint32_t i = uint32_t((d0 >= 0 ? j[d0] * g.st[d0] : 0) + (d1 >= 0 ? j[d1] * g.st[d1] : 0) + z0);
  // points of the group inside the launch box along the row
  unsigned boxmask = 0;
#pragma unroll
  for (int k = 0; k < V; ++k)
    if (z0 + k >= g.blo[d2] && z0 + k <= g.bhi[d2]) boxmask |= 1u << k;
  // ownership of each component in this group
  unsigned cmask[3] = {0, 0, 0}, any = 0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (CURL ? !a.cc[c].f : !a.uc[c].f) continue;
    const int *lo = CURL ? a.cc[c].lo : a.uc[c].lo;
    const int *hi = CURL ? a.cc[c].hi : a.uc[c].hi;
    if ((d0 >= 0 && (j[d0] < lo[d0] || j[d0] > hi[d0])) || (d1 >= 0 && (j[d1] < lo[d1] || j[d1] > hi[d1]))) continue;
    unsigned m = boxmask;
#pragma unroll
    for (int k = 0; k < V; ++k)
      if (z0 + k < lo[d2] || z0 + k > hi[d2]) m &= ~(1u << k);
    cmask[c] = m;
    any |= m;
  }
  if (!any) return;
  // centre rows of the partner components (by direction), loaded once
  Vec<R, V> F0, F1, F2;
  if (CURL) {
    if (a.F[0]) F0 = vload<R, V>(a.F[0] + i);
    if (a.F[1]) F1 = vload<R, V>(a.F[1] + i);
    if (a.F[2]) F2 = vload<R, V>(a.F[2] + i);
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const unsigned mask = cmask[c];
    if (!mask) continue;
    const int *ysh = CURL ? a.cc[c].ysh : a.uc[c].ysh;
    const unsigned metal = (g.hi_metal | g.lo_magnetic) ? metal_mask(g, j, ysh, d0, d1, d2, z0, V) : 0u;
    Vec<R, V> fv;
    if (CURL) {
      const CurlComp<R> &kc = a.cc[c];
      // slot c = direction c: g1 = F[(c+2)%3] along c+1, g2 = F[(c+1)%3] along c+2
      const int dp = (c + 1) % 3, dm = (c + 2) % 3;
      const R *g1 = a.F[(c + 2) % 3], *g2 = a.F[(c + 1) % 3];
      const Vec<R, V> &g1c = c == 0 ? F2 : (c == 1 ? F0 : F1);
      const Vec<R, V> &g2c = c == 0 ? F1 : (c == 1 ? F2 : F0);
      Vec<R, V> cu;
#pragma unroll
      for (int k = 0; k < V; ++k) cu.v[k] = R(0);
      if (a.hasp[c]) {
        const Vec<R, V> g1n = neighbour<R, V>(g1, i, a.sgn * g.st[dp], g1c, mask);
#pragma unroll
        for (int k = 0; k < V; ++k) cu.v[k] = g1n.v[k] - g1c.v[k];
      }
      if (a.hasm[c]) {
        const Vec<R, V> g2n = neighbour<R, V>(g2, i, a.sgn * g.st[dm], g2c, mask);
        if (a.hasp[c]) {
#pragma unroll
          for (int k = 0; k < V; ++k) cu.v[k] = cu.v[k] + g2c.v[k] - g2n.v[k];
        }
        else {
#pragma unroll
          for (int k = 0; k < V; ++k) cu.v[k] = g2c.v[k] - g2n.v[k];
        }
      }
      fv = vload<R, V>(kc.f + i);
#pragma unroll
      for (int k = 0; k < V; ++k) fv.v[k] -= a.dtdx * cu.v[k];
      if (a.cur.n && int64_t(i) + V - 1 >= a.cur.lo && int64_t(i) <= a.cur.hi) {
#pragma unroll
        for (int k = 0; k < V; ++k) fv.v[k] = apply_tab(a.cur, c, i + k, fv.v[k]);
      }
      if (metal) {
#pragma unroll
        for (int k = 0; k < V; ++k)
          if ((metal >> k) & 1) fv.v[k] = R(0);
      }
      vstore<R, V>(kc.f + i, fv, mask);
      if (a.zwrap) wrap_store<R, V>(kc.f, i, z0, ysh[d2], g.num[d2], a.zsign, fv, mask);
    }
    if (UPD && a.uc[c].f) {
      const UpdComp<R> &u = a.uc[c];
      Vec<R, V> gv = CURL ? fv : vload<R, V>(u.g + i);
#pragma unroll
      for (int p = 0; p < MAXPOLE; ++p)
        if (p < a.npole && u.P[p]) {
          const Vec<R, V> P = vload<R, V>(u.P[p] + i);
#pragma unroll
          for (int k = 0; k < V; ++k) gv.v[k] -= P.v[k];
        }
      if (a.dip.n && int64_t(i) + V - 1 >= a.dip.lo && int64_t(i) <= a.dip.hi) {
#pragma unroll
        for (int k = 0; k < V; ++k) gv.v[k] = apply_tab(a.dip, c, i + k, gv.v[k]);
      }
      Vec<R, V> ev = gv;
      if (u.u) {
        const Vec<R, V> uu = vload<R, V>(u.u + i);
#pragma unroll
        for (int k = 0; k < V; ++k) ev.v[k] = gv.v[k] * uu.v[k];
      }
#pragma unroll
      for (int p = 0; p < MAXPOLE; ++p)
        if (p < a.npole && u.P[p] && u.sigma[p]) {
          const PoleCoef<R> &q = a.pole[p];
          const Vec<R, V> P = vload<R, V>(u.P[p] + i), Pp = vload<R, V>(u.Pp[p] + i), sg = vload<R, V>(u.sigma[p] + i);
          Vec<R, V> Pn;
#pragma unroll
          for (int k = 0; k < V; ++k)
            Pn.v[k] = q.gamma1inv * (P.v[k] * (R(2) - q.omega0dtsqr_denom) - q.gamma1 * Pp.v[k] + q.omega0dtsqr * (sg.v[k] * ev.v[k]));
          vstore<R, V>(u.P[p] + i, Pn, mask);
          vstore<R, V>(u.Pp[p] + i, P, mask);
          if (a.zwrap) wrap_store<R, V>(u.P[p], i, z0, ysh[d2], g.num[d2], a.zsign, Pn, mask);
        }
      if (metal) {
#pragma unroll
        for (int k = 0; k < V; ++k)
          if ((metal >> k) & 1) ev.v[k] = R(0);
      }
      vstore<R, V>(u.f + i, ev, mask);
      if (a.zwrap) wrap_store<R, V>(u.f, i, z0, ysh[d2], g.num[d2], a.zsign, ev, mask);
    }
  }
}

// ---------------------------------------------------------------------------------------

template <class R>
__global__ __launch_bounds__(BLOCK) void k_conn(const Conn *conns, const double2 *phases, int n, R *const *slots, int ncmp) {
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= n) return;
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
__global__ __launch_bounds__(BLOCK) void k_src(R *f, const int64_t *idx, const double *val, int n) {
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= n) return;
  f[idx[t]] = R(double(f[idx[t]]) - val[t]);
}

template <class R> __device__ __forceinline__ R avg_value(const R *F, int64_t idx, double w, int64_t a1, int64_t a2) {
  if (a2) return R((w * 0.25) * double(F[idx] + F[idx + a1] + F[idx + a2] + F[idx + (a1 + a2)]));
  if (a1) return R((w * 0.5) * double(F[idx] + F[idx + a1]));
  return R(w * double(F[idx]));
}

template <class R> __global__ __launch_bounds__(BLOCK) void k_dft(const DftArgs<R> a) {
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  const int k = t / a.nfreq;
  if (k >= a.npts) return;
  const int fi = t - k * a.nfreq;
  const int64_t idx = a.idx[k];
  const double w = a.w[k];
  const R f0 = avg_value(a.f0, idx, w, a.avg1, a.avg2);
  const R f1 = a.f1 ? avg_value(a.f1, idx, w, a.avg1, a.avg2) : R(0);
  double sn, cs;
  sincos(a.omega[fi] * a.time, &sn, &cs);
  const double pr = cs * a.scale_re - sn * a.scale_im, pim = cs * a.scale_im + sn * a.scale_re;
  const size_t o = size_t(k) * a.nfreq + fi;
  if (a.acc_kind == 1) { // FP64 accumulators
    double2 *acc = reinterpret_cast<double2 *>(a.acc);
    double2 v = acc[o];
    v.x += pr * double(f0) - pim * double(f1);
    v.y += pr * double(f1) + pim * double(f0);
    acc[o] = v;
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
    float2 *acc = reinterpret_cast<float2 *>(a.acc);
    float2 *cp = reinterpret_cast<float2 *>(a.comp);
    float2 s = acc[o], c = cp[o];
    const float yr = float(tr) - c.x, yi = float(ti) - c.y;
    const float sr = s.x + yr, si = s.y + yi;
    c.x = (sr - s.x) - yr;
    c.y = (si - s.y) - yi;
    s.x = sr;
    s.y = si;
    acc[o] = s;
    cp[o] = c;
    return;
  }
  R *acc = reinterpret_cast<R *>(a.acc);
  acc[2 * o] += tr;
  acc[2 * o + 1] += ti;
}

template <class R>
__global__ void k_check(const R *const *ptrs, const int64_t *idx, const double *w, int n, int *flag) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const R v = ptrs[t][idx[t]];
  if (w[t] != 0.0 && !isfinite(double(v))) atomicOr(flag, 1);
}

static inline unsigned nblocks(size_t n) { return unsigned((n + BLOCK - 1) / BLOCK); }
static inline unsigned nblocks8(size_t n) { return (nblocks(n) + 7u) / 8u * 8u; }

template <class R> void launch_step(const StepArgs<R> &a, bool curl, bool upd, hipStream_t s) {
  size_t count = 1;
  for (int k = 0; k < 3; ++k)
    if (a.g.axdir[k] >= 0) count *= size_t(a.g.bhi[a.g.axdir[k]] - a.g.blo[a.g.axdir[k]] + 1);
  const unsigned nb = nblocks8(count);
  const uint32_t c32 = uint32_t(count);
  if (curl && upd)
    hipLaunchKernelGGL((k_step<R, true, true>), dim3(nb), dim3(BLOCK), 0, s, a, c32);
  else if (curl)
    hipLaunchKernelGGL((k_step<R, true, false>), dim3(nb), dim3(BLOCK), 0, s, a, c32);
  else if (upd)
    hipLaunchKernelGGL((k_step<R, false, true>), dim3(nb), dim3(BLOCK), 0, s, a, c32);
}

template <class R> void launch_fast(const StepArgs<R> &a, bool curl, bool upd, hipStream_t s) {
  constexpr int V = 16 / sizeof(R);
  const int d0 = a.g.axdir[0], d1 = a.g.axdir[1], d2 = a.g.axdir[2];
  const int g0 = a.g.blo[d2] / V, g1 = a.g.bhi[d2] / V;
  const int ngroups = g1 - g0 + 1;
  size_t count = size_t(ngroups);
  if (d0 >= 0) count *= size_t(a.g.bhi[d0] - a.g.blo[d0] + 1);
  if (d1 >= 0) count *= size_t(a.g.bhi[d1] - a.g.blo[d1] + 1);
  const unsigned nb = nblocks8(count);
  const uint32_t c32 = uint32_t(count);
  if (curl && upd)
    hipLaunchKernelGGL((k_fast<R, V, true, true>), dim3(nb), dim3(BLOCK), 0, s, a, c32, g0, ngroups);
  else if (curl)
    hipLaunchKernelGGL((k_fast<R, V, true, false>), dim3(nb), dim3(BLOCK), 0, s, a, c32, g0, ngroups);
  else if (upd)
    hipLaunchKernelGGL((k_fast<R, V, false, true>), dim3(nb), dim3(BLOCK), 0, s, a, c32, g0, ngroups);
}

template <class R>
void launch_conn(const Conn *conns, const double2 *phases, int n, R *const *slots, int ncmp, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_conn<R>, dim3(nblocks(n)), dim3(BLOCK), 0, s, conns, phases, n, slots, ncmp);
}

template <class R> void launch_src(R *f, const int64_t *idx, const double *val, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_src<R>, dim3(nblocks(n)), dim3(BLOCK), 0, s, f, idx, val, n);
}

template <class R> void launch_dft(const DftArgs<R> &a, hipStream_t s) {
  const size_t n = size_t(a.npts) * a.nfreq;
  if (n > 0) hipLaunchKernelGGL(k_dft<R>, dim3(nblocks(n)), dim3(BLOCK), 0, s, a);
}

template <class R>
void launch_check(const R *const *ptrs, const int64_t *idx, const double *w, int n, int *flag, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_check<R>, dim3(nblocks(n)), dim3(BLOCK), 0, s, ptrs, idx, w, n, flag);
}

#define INSTANTIATE(R)                                                                             \
  template void launch_step<R>(const StepArgs<R> &, bool, bool, hipStream_t);                      \
  template void launch_fast<R>(const StepArgs<R> &, bool, bool, hipStream_t);                      \
  template void launch_conn<R>(const Conn *, const double2 *, int, R *const *, int, hipStream_t);  \
  template void launch_src<R>(R *, const int64_t *, const double *, int, hipStream_t);            \
  template void launch_dft<R>(const DftArgs<R> &, hipStream_t);                                    \
  template void launch_check<R>(const R *const *, const int64_t *, const double *, int, int *,     \
                                hipStream_t);
INSTANTIATE(float)
INSTANTIATE(double)

} // namespace mx
