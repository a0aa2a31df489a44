// kernels.hip — CDNA4 (gfx950) kernels of the FDTD time step.
//
// The step is HBM-bound stencil arithmetic (no contraction, so no MFMA): each kernel streams
// the reference-layout arrays once, one thread per array element, 64-lane waves walking the
// stride-1 axis so every load and store is a coalesced 256-byte (FP32) / 512-byte (FP64) wave
// access.  Neighbour reads of the curl (i+s along the two slower axes) hit L2 within one XCD:
// consecutive blocks cover consecutive rows.
//
// k_step fuses, per cell and per component, what the reference does in separate passes:
//   curl + uPML            step_curl             src/step_generic.cpp:65-249
//   current injection      fields_chunk::step_source  src/step.cpp:295-318
//   metal walls            zero_metal            src/boundaries.cpp:310-313
//   f_minus_p = D - P - p  update_eh             src/update_eh.cpp:114-138
//   E = chi1inv f_minus_p  step_update_EDHB      src/step_generic.cpp:682-701,774-782
//   Lorentz/Drude P        update_P              src/susceptibility.cpp:188-262
// in the reference's order, so the arithmetic of each value is the reference's.
#include "engine.hpp"

#include <hip/hip_runtime.h>

namespace mx {

constexpr int BLOCK = 256;

__device__ __forceinline__ void cell_index(const Grid &g, uint32_t i, int j[3]) {
  j[0] = j[1] = j[2] = 0;
  const uint32_t e2 = uint32_t(g.ext[2]), e1 = uint32_t(g.ext[1]);
  const uint32_t r = i / e2;
  const uint32_t q2 = i - r * e2;
  const uint32_t q0 = r / e1;
  const uint32_t q1 = r - q0 * e1;
  if (g.axdir[0] >= 0) j[g.axdir[0]] = int(q0);
  if (g.axdir[1] >= 0) j[g.axdir[1]] = int(q1);
  if (g.axdir[2] >= 0) j[g.axdir[2]] = int(q2);
}

__device__ __forceinline__ bool owned(const int j[3], const int lo[3], const int hi[3]) {
  return j[0] >= lo[0] && j[0] <= hi[0] && j[1] >= lo[1] && j[1] <= hi[1] && j[2] >= lo[2] &&
         j[2] <= hi[2];
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

template <class R>
__device__ __forceinline__ bool pml_active(const PmlProfile<R> &p, int d, int k) {
  return p.sig[d][k] != R(0) || p.kap[d][k] != R(1);
}

template <class R, bool CURL, bool UPD>
__global__ __launch_bounds__(BLOCK) void k_step(const StepArgs<R> a) {
  const uint32_t i = blockIdx.x * uint32_t(BLOCK) + threadIdx.x;
  if (i >= a.g.ntot) return;
  int j[3];
  cell_index(a.g, i, j);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (c >= a.nc) break;
    const int *lo = CURL ? a.cc[c].lo : a.uc[c].lo;
    const int *hi = CURL ? a.cc[c].hi : a.uc[c].hi;
    if (!owned(j, lo, hi)) continue;
    R fv = R(0);
    if (CURL) {
      const CurlComp<R> &k = a.cc[c];
      R cu = R(0);
      if (k.g2)
        cu = k.g1[i + k.s1] - k.g1[i] + k.g2[i] - k.g2[i + k.s2];
      else if (k.g1)
        cu = k.g1[i + k.s1] - k.g1[i];
      fv = k.f[i];
      const bool pml = k.dsig >= 0 && pml_active(a.pml, k.dsig, 2 * j[k.dsig] + k.ysh[k.dsig]);
      const bool pmlu = k.dsigu >= 0 && pml_active(a.pml, k.dsigu, 2 * j[k.dsigu] + k.ysh[k.dsigu]);
      if (!pmlu) {
        if (!pml)
          fv -= k.dtdx * cu;
        else {
          const int kk = 2 * j[k.dsig] + k.ysh[k.dsig];
          fv = ((a.pml.kap[k.dsig][kk] - a.pml.sig[k.dsig][kk]) * fv - k.dtdx * cu) *
               a.pml.siginv[k.dsig][kk];
        }
      }
      else {
        const int ku = 2 * j[k.dsigu] + k.ysh[k.dsigu];
        const R fprev = k.fu[i];
        R fun;
        if (!pml)
          fun = fprev - k.dtdx * cu;
        else {
          const int kk = 2 * j[k.dsig] + k.ysh[k.dsig];
          fun = ((a.pml.kap[k.dsig][kk] - a.pml.sig[k.dsig][kk]) * fprev - k.dtdx * cu) *
                a.pml.siginv[k.dsig][kk];
        }
        k.fu[i] = fun;
        fv = a.pml.siginv[k.dsigu][ku] *
             ((a.pml.kap[k.dsigu][ku] - a.pml.sig[k.dsigu][ku]) * fv + fun - fprev);
      }
#pragma unroll
      for (int s = 0; s < SRC_TAB; ++s)
        if (s < a.cur.n && a.cur.comp[s] == c && a.cur.idx[s] == int64_t(i))
          fv = R(double(fv) - a.cur.val[s]);
      if (on_metal(a.g, j, k.ysh)) fv = R(0);
      k.f[i] = fv;
    }
    if (UPD && a.uc[c].f) {
      const UpdComp<R> &u = a.uc[c];
      R gv = CURL ? fv : u.g[i];
#pragma unroll
      for (int p = 0; p < MAXPOLE; ++p)
        if (p < a.npole && u.P[p]) gv -= u.P[p][i];
#pragma unroll
      for (int s = 0; s < SRC_TAB; ++s)
        if (s < a.dip.n && a.dip.comp[s] == c && a.dip.idx[s] == int64_t(i))
          gv = R(double(gv) - a.dip.val[s]);
      R ev, wv;
      const bool pmlw = u.dsigw >= 0 && pml_active(a.pml, u.dsigw, 2 * j[u.dsigw] + u.ysh[u.dsigw]);
      if (pmlw) {
        const int kw = 2 * j[u.dsigw] + u.ysh[u.dsigw];
        const R kapw = a.pml.kap[u.dsigw][kw], sigw = a.pml.sig[u.dsigw][kw];
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
          u.P[p][i] = q.gamma1inv * (pcur * (R(2) - q.omega0dtsqr_denom) - q.gamma1 * u.Pp[p][i] +
                                     q.omega0dtsqr * (u.sigma[p][i] * wv));
          u.Pp[p][i] = pcur;
        }
      if (on_metal(a.g, j, u.ysh)) ev = R(0);
      u.f[i] = ev;
    }
  }
}

template <class R>
__global__ __launch_bounds__(BLOCK) void k_conn(const Conn *conns, int n, R *const *slots, int ncmp) {
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= n) return;
  const Conn e = conns[t];
  R *a0 = slots[2 * e.slot], *a1 = slots[2 * e.slot + 1];
  if (e.cls == 0) { // CONNECT_PHASE: complex<realnum> product
    const R re = a0[e.src], im = a1[e.src];
    const R pr = R(e.ph_re), pi = R(e.ph_im);
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
    float yr = float(tr) - c.x, yi = float(ti) - c.y;
    float sr = s.x + yr, si = s.y + yi;
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

template <class R> void launch_step(const StepArgs<R> &a, bool curl, bool upd, hipStream_t s) {
  const unsigned nb = nblocks(a.g.ntot);
  if (curl && upd)
    hipLaunchKernelGGL((k_step<R, true, true>), dim3(nb), dim3(BLOCK), 0, s, a);
  else if (curl)
    hipLaunchKernelGGL((k_step<R, true, false>), dim3(nb), dim3(BLOCK), 0, s, a);
  else if (upd)
    hipLaunchKernelGGL((k_step<R, false, true>), dim3(nb), dim3(BLOCK), 0, s, a);
}

template <class R> void launch_conn(const Conn *conns, int n, R *const *slots, int ncmp, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_conn<R>, dim3(nblocks(n)), dim3(BLOCK), 0, s, conns, n, slots, ncmp);
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
  template void launch_conn<R>(const Conn *, int, R *const *, int, hipStream_t);                   \
  template void launch_src<R>(R *, const int64_t *, const double *, int, hipStream_t);            \
  template void launch_dft<R>(const DftArgs<R> &, hipStream_t);                                    \
  template void launch_check<R>(const R *const *, const int64_t *, const double *, int, int *,     \
                                hipStream_t);
INSTANTIATE(float)
INSTANTIATE(double)

} // namespace mx
