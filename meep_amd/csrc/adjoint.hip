// adjoint.hip — the reverse-mode (time-reversed) FDTD step on gfx950.
//
// The step is linear in the fields, so the gradient of an objective of the final DFT
// accumulators is a backward sweep of the transposed step over cotangent ("bar") arrays.
// Each forward update of kernels.hip has its transpose here, written per point from the
// update's own formula (the reference's step_curl / step_update_EDHB branches,
// src/step_generic.cpp:65-249,682-701,774-782):
//
//   curl, plain            f' = f - d C                      f~ = f~',  C~ = -d f~'
//   curl, uPML in f        f' = s (k - g) f - s d C          f~ = s (k - g) f~',  C~ = -s d f~'
//   curl, uPML through f_u u' = u - d C ;  f' = su (ku - gu) f + su (u' - u)
//                          with G = su f~':  u~ = u~',  C~ = -d (u~' + G),  f~ = su (ku - gu) f~'
//   curl, both             u' = s (k - g) u - s d C ;  f' as above
//                          u~ = s (k - g) (u~' + G) - G,  C~ = -s d (u~' + G)
//   update, plain          e' = x g                          g~ += x e~',  e~ = 0
//   update, uPML W         w' = x g ;  e' = e + (kw + gw) w' - (kw - gw) w
//                          W = w~' + (kw + gw) e~':  g~ += x W,  w~ = -(kw - gw) e~',  e~ = e~'
//
// (d = Courant, s = 1/(k + g), x = chi1inv).  Metal points are zeroed after the forward
// update, so their cotangents are zeroed before the transpose.  The curl term's cotangent C~
// is then gathered into the partner arrays (the transpose of the difference stencil), halo
// images fold back into their owned sources, and the DFT objective injects its cotangent.
#include "engine.hpp"
#include "stencil.hpp"

#include <hip/hip_runtime.h>

namespace mx {

// the thread's point of the full grid, by loop axis
struct Pt {
  int j0, j1, j2;
  uint32_t i;
};
__device__ __forceinline__ bool grid_point(const Grid &g, uint32_t t, Pt &p) {
  const uint32_t r = t / uint32_t(g.an[2]);
  const int q2 = int(t - r * uint32_t(g.an[2]));
  const int q0 = int(r / uint32_t(g.an[1]));
  const int q1 = int(r - uint32_t(q0) * uint32_t(g.an[1]));
  p.j0 = g.alo[0] + q0;
  p.j1 = g.alo[1] + q1;
  p.j2 = g.alo[2] + q2;
  p.i = uint32_t(p.j0 * g.ast[0] + p.j1 * g.ast[1] + p.j2 * g.ast[2]);
  return true;
}

// ---- update transpose ---------------------------------------------------------------------
template <class R, int C>
__device__ __forceinline__ void adj_upd_slot(const StepArgs<R> &a, const AdjSlots<R> &b, const AdjDesign<R> &dz, const Pt &p) {
  if (C >= a.nc || !a.uc[C].f) return;
  const UpdComp<R> &u = a.uc[C];
  if (!slot_owned(u.ax, p.j0, p.j1, p.j2)) return;
  const uint32_t i = p.i;
  R eb = b.fb[C][i];
  if (slot_metal(u.ax, p.j0, p.j1, p.j2)) eb = R(0);
  const int kw = u.ksigw >= 0 ? 2 * asel(p.j0, p.j1, p.j2, u.ksigw) + u.ysigw : 0;
  R wb; // cotangent of u * g
  if (pml_on(u.sigw, u.kapw, kw)) {
    const R kapw = u.kapw[kw], sigw = u.sigw[kw];
    wb = b.fwb[C][i] + (kapw + sigw) * eb;
    b.fwb[C][i] = -(kapw - sigw) * eb;
    b.fb[C][i] = eb;
  }
  else {
    wb = eb;
    b.fb[C][i] = R(0);
  }
  b.gb[C][i] += u.u ? u.u[i] * wb : wb;
  if (dz.slot == C && p.j0 >= dz.lo[0] && p.j0 <= dz.hi[0] && p.j1 >= dz.lo[1] && p.j1 <= dz.hi[1] && p.j2 >= dz.lo[2] &&
      p.j2 <= dz.hi[2]) {
    const int64_t k = (int64_t(p.j0 - dz.lo[0]) * (dz.hi[1] - dz.lo[1] + 1) + (p.j1 - dz.lo[1])) * (dz.hi[2] - dz.lo[2] + 1) +
                      (p.j2 - dz.lo[2]);
    dz.grad[k] += double(wb) * double(dz.hist[k]);
  }
}

template <class R>
__global__ __launch_bounds__(BLOCK) void k_adj_upd(const StepArgs<R> a, const AdjSlots<R> b, const AdjDesign<R> dz, uint32_t count) {
  const uint32_t t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= count) return;
  Pt p;
  grid_point(a.g, t, p);
  adj_upd_slot<R, 0>(a, b, dz, p);
  adj_upd_slot<R, 1>(a, b, dz, p);
  adj_upd_slot<R, 2>(a, b, dz, p);
}

// ---- curl transpose (local part) ------------------------------------------------------------
template <class R, int C> __device__ __forceinline__ void adj_curl_slot(const StepArgs<R> &a, const AdjSlots<R> &b, const Pt &p) {
  if (C >= a.nc || !a.cc[C].f) return;
  const CurlComp<R> &k = a.cc[C];
  if (!slot_owned(k.ax, p.j0, p.j1, p.j2)) return;
  const uint32_t i = p.i;
  R fb = b.fb[C][i];
  if (slot_metal(k.ax, p.j0, p.j1, p.j2)) fb = R(0);
  const int kk = k.ksig >= 0 ? 2 * asel(p.j0, p.j1, p.j2, k.ksig) + k.ysig : 0;
  const int ku = k.ksigu >= 0 ? 2 * asel(p.j0, p.j1, p.j2, k.ksigu) + k.ysigu : 0;
  const bool pml = pml_on(k.sig, k.kap, kk);
  const bool pmlu = pml_on(k.sigu, k.kapu, ku);
  R cb;
  if (!pmlu) {
    if (!pml) {
      cb = -k.dtdx * fb;
      b.fb[C][i] = fb;
    }
    else {
      cb = -k.siginv[kk] * k.dtdx * fb;
      b.fb[C][i] = (k.kap[kk] - k.sig[kk]) * k.siginv[kk] * fb;
    }
  }
  else {
    const R G = k.siginvu[ku] * fb;
    const R ut = b.fub[C][i] + G;
    if (!pml) {
      cb = -k.dtdx * ut;
      b.fub[C][i] = ut - G;
    }
    else {
      cb = -k.siginv[kk] * k.dtdx * ut;
      b.fub[C][i] = (k.kap[kk] - k.sig[kk]) * k.siginv[kk] * ut - G;
    }
    b.fb[C][i] = k.siginvu[ku] * (k.kapu[ku] - k.sigu[ku]) * fb;
  }
  b.cbar[C][i] = cb;
}

template <class R> __global__ __launch_bounds__(BLOCK) void k_adj_curl(const StepArgs<R> a, const AdjSlots<R> b, uint32_t count) {
  const uint32_t t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= count) return;
  Pt p;
  grid_point(a.g, t, p);
  adj_curl_slot<R, 0>(a, b, p);
  adj_curl_slot<R, 1>(a, b, p);
  adj_curl_slot<R, 2>(a, b, p);
}

// ---- curl transpose (gather into a partner) -------------------------------------------------
template <class R, int E> __device__ __forceinline__ R gather_entry(const Grid &g, const AdjGather<R> &ga, const Pt &p, int64_t i) {
  if (E >= ga.n) return R(0);
  // the point j - s exists when its coordinate along ax stays inside 0 .. n
  const int jc = asel(p.j0, p.j1, p.j2, ga.ax[E]) - ga.step[E];
  const int n = asel(g.nax[0], g.nax[1], g.nax[2], ga.ax[E]);
  const R behind = (jc >= 0 && jc <= n) ? ga.cbar[E][i - ga.s[E]] : R(0);
  const R here = ga.cbar[E][i];
  return ga.plus[E] ? behind - here : here - behind;
}
template <class R> __global__ __launch_bounds__(BLOCK) void k_adj_gather(const Grid g, const AdjGather<R> ga, uint32_t count) {
  const uint32_t t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= count) return;
  Pt p;
  grid_point(g, t, p);
  const int64_t i = p.i;
  R v = gather_entry<R, 0>(g, ga, p, i);
  v += gather_entry<R, 1>(g, ga, p, i);
  v += gather_entry<R, 2>(g, ga, p, i);
  v += gather_entry<R, 3>(g, ga, p, i);
  ga.gb[i] += v;
}

// ---- halo transpose ------------------------------------------------------------------------
template <class R>
__global__ __launch_bounds__(BLOCK) void k_adj_halo(const int *first, const uint32_t *src, const uint32_t *dst, const int8_t *sgn,
                                                   const uint16_t *slot, int nsrc, R *const *bars) {
  const int q = blockIdx.x * BLOCK + threadIdx.x;
  if (q >= nsrc) return;
  for (int e = first[q]; e < first[q + 1]; ++e) {
    R *bar = bars[slot[e]];
    const R v = bar[dst[e]];
    bar[src[q]] += sgn[e] > 0 ? v : -v;
    bar[dst[e]] = R(0);
  }
}

// ---- DFT objective cotangent -----------------------------------------------------------------
template <class R>
__global__ __launch_bounds__(BLOCK) void k_adj_dft(const double *w, const double2 *acc, const double *omega, int npts, int nfreq,
                                                  double time, double scale_re, double scale_im, double *G) {
  const int k = blockIdx.x * BLOCK + threadIdx.x;
  if (k >= npts) return;
  double sum = 0;
  for (int f = 0; f < nfreq; ++f) {
    double sn, cs;
    sincos(omega[f] * time, &sn, &cs);
    const double pr = cs * scale_re - sn * scale_im, pim = cs * scale_im + sn * scale_re;
    const double2 a = acc[size_t(k) * nfreq + f];
    sum += 2.0 * (a.x * pr + a.y * pim); // 2 Re(conj(a) ph)
  }
  G[k] = w[k] * sum;
}
template <class R>
__global__ __launch_bounds__(BLOCK) void k_adj_dft_scatter(R *bar, const int64_t *idx, const double *G, int npts, int64_t off, double coef) {
  const int k = blockIdx.x * BLOCK + threadIdx.x;
  if (k >= npts) return;
  R *p = bar + idx[k] + off;
  *p = R(double(*p) + coef * G[k]);
}

// ---- forward history of D at the design points ------------------------------------------------
template <class R>
__global__ __launch_bounds__(BLOCK) void k_record(R *hist, const R *D, const Grid g, int lo0, int lo1, int lo2, int n0, int n1, int n2) {
  const uint32_t t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= uint32_t(n0) * uint32_t(n1) * uint32_t(n2)) return;
  const uint32_t r = t / uint32_t(n2);
  const int q2 = int(t - r * uint32_t(n2));
  const int q0 = int(r / uint32_t(n1));
  const int q1 = int(r - uint32_t(q0) * uint32_t(n1));
  const int64_t i = int64_t(lo0 + q0) * g.ast[0] + int64_t(lo1 + q1) * g.ast[1] + int64_t(lo2 + q2) * g.ast[2];
  hist[t] = D[i];
}

// ---- launchers ---------------------------------------------------------------------------------
static inline uint32_t grid_count(const Grid &g) { return uint32_t(g.an[0]) * uint32_t(g.an[1]) * uint32_t(g.an[2]); }

template <class R> void launch_adj_upd(const StepArgs<R> &a, const AdjSlots<R> &b, const AdjDesign<R> &dz, hipStream_t s) {
  const uint32_t n = grid_count(a.g);
  hipLaunchKernelGGL(k_adj_upd<R>, dim3(nblocks(n)), dim3(BLOCK), 0, s, a, b, dz, n);
}
template <class R> void launch_adj_curl(const StepArgs<R> &a, const AdjSlots<R> &b, hipStream_t s) {
  const uint32_t n = grid_count(a.g);
  hipLaunchKernelGGL(k_adj_curl<R>, dim3(nblocks(n)), dim3(BLOCK), 0, s, a, b, n);
}
template <class R> void launch_adj_gather(const Grid &g, const AdjGather<R> &ga, hipStream_t s) {
  const uint32_t n = grid_count(g);
  if (ga.n > 0) hipLaunchKernelGGL(k_adj_gather<R>, dim3(nblocks(n)), dim3(BLOCK), 0, s, g, ga, n);
}
template <class R>
void launch_adj_halo(const int *first, const uint32_t *src, const uint32_t *dst, const int8_t *sgn, const uint16_t *slot, int nsrc,
                     R *const *bars, hipStream_t s) {
  if (nsrc > 0) hipLaunchKernelGGL(k_adj_halo<R>, dim3(nblocks(nsrc)), dim3(BLOCK), 0, s, first, src, dst, sgn, slot, nsrc, bars);
}
template <class R>
void launch_adj_dft(const int64_t *, const double *w, const double2 *acc, const double *omega, int npts, int nfreq, double time,
                    double scale_re, double scale_im, double *G, hipStream_t s) {
  if (npts > 0)
    hipLaunchKernelGGL(k_adj_dft<R>, dim3(nblocks(npts)), dim3(BLOCK), 0, s, w, acc, omega, npts, nfreq, time, scale_re, scale_im, G);
}
template <class R> void launch_adj_dft_scatter(R *bar, const int64_t *idx, const double *G, int npts, int64_t off, double coef, hipStream_t s) {
  if (npts > 0) hipLaunchKernelGGL(k_adj_dft_scatter<R>, dim3(nblocks(npts)), dim3(BLOCK), 0, s, bar, idx, G, npts, off, coef);
}
template <class R> void launch_record(R *hist, const R *D, const Grid &g, const int *lo, const int *hi, hipStream_t s) {
  const int n0 = hi[0] - lo[0] + 1, n1 = hi[1] - lo[1] + 1, n2 = hi[2] - lo[2] + 1;
  const size_t n = size_t(n0) * n1 * n2;
  if (n > 0) hipLaunchKernelGGL(k_record<R>, dim3(nblocks(n)), dim3(BLOCK), 0, s, hist, D, g, lo[0], lo[1], lo[2], n0, n1, n2);
}

#define INSTANTIATE_ADJ(R)                                                                                                 \
  template void launch_adj_upd<R>(const StepArgs<R> &, const AdjSlots<R> &, const AdjDesign<R> &, hipStream_t);           \
  template void launch_adj_curl<R>(const StepArgs<R> &, const AdjSlots<R> &, hipStream_t);                                \
  template void launch_adj_gather<R>(const Grid &, const AdjGather<R> &, hipStream_t);                                    \
  template void launch_adj_halo<R>(const int *, const uint32_t *, const uint32_t *, const int8_t *, const uint16_t *, int, \
                                   R *const *, hipStream_t);                                                               \
  template void launch_adj_dft<R>(const int64_t *, const double *, const double2 *, const double *, int, int, double, double, \
                                  double, double *, hipStream_t);                                                          \
  template void launch_adj_dft_scatter<R>(R *, const int64_t *, const double *, int, int64_t, double, hipStream_t);       \
  template void launch_record<R>(R *, const R *, const Grid &, const int *, const int *, hipStream_t);
// The build compiles this file once per field precision (MX_REAL), in parallel.
#ifdef MX_REAL
INSTANTIATE_ADJ(MX_REAL)
#else
INSTANTIATE_ADJ(float)
INSTANTIATE_ADJ(double)
#endif

} // namespace mx
