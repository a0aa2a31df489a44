// near2far.hip — the near-to-far transformation's field sums on gfx950.
//
// dft_near2far::farfield_lowlevel (src/near2far.cpp:350-397) sums, for every far point and
// frequency, the homogeneous-medium Green's function of every equivalent surface current (a
// DFT accumulator of a tangential field, at its grid location and its periodic copies).  The
// work is a dense (far points x frequencies) by (surface currents) product of a transcendental
// kernel: FP64 vector arithmetic (sincos, or Bessel functions in 2D), no contraction to feed
// MFMA; the bound is VALU instruction issue.  Each block takes one far point, a group of
// consecutive frequencies and one segment of the current list; its 256 threads stride over the
// segment, compute each current's geometry once for the group, keep the six complex fields of
// each frequency in registers, and reduce them through the wave and LDS into one partial per
// frequency.  A second pass adds the partials of each (point, frequency) in segment order, so
// the result does not depend on timing.
//
// The currents come as SoA lists prepared on the host, sorted by class (Ex Ey Ez Hx Hy Hz):
// position x, y, z and one amplitude per frequency (frequency-major, so a block's reads of one
// frequency are contiguous).  The periodic copies are separate entries whose amplitude carries
// their Bloch phase.
#include "engine.hpp"

#include <hip/hip_runtime.h>

namespace mx {

constexpr int FF_BLOCK = 256;
constexpr double FF_PI = 3.141592653589793238462643383279502884197;

struct Cx {
  double re, im;
};
__device__ __forceinline__ Cx cmul(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ Cx cscale(Cx a, double s) { return {a.re * s, a.im * s}; }

// component numbering of the reference (Ex Ey Er Ep Ez, Hx Hy Hr Hp Hz): its direction
// (0 X, 1 Y, 2 Z) and kind
__device__ __forceinline__ int ff_dir(int c) {
  const int m = c % 5;
  return m == 0 ? 0 : m == 1 ? 1 : 2;
}

// green3d (src/near2far.cpp:133-186) for a current of direction PD (0 X, 1 Y, 2 Z) and kind
// ELECTRIC at offset (rh r) from the source, written as the reference's terms with
// 1/(ikr) = -i/(kr) and 1/(ikr)^2 = -1/(kr)^2 expanded: the fields of the current's own kind
// are pre (t1 p + t2 (p.r) r), those of the other kind pre t3 (r x p) divided (J) or
// multiplied (magnetic current, negated) by the impedance, pre = i e^{ikr} k n f0 / (4 pi r
// (eps or mu)).  (cs, sn) = e^{ikr}; ikr = 1/(kr); amp = k n / (4 pi r (eps or mu)).
template <int PD, bool ELECTRIC>
__device__ __forceinline__ void green3d_term(Cx acc[6], const double rh[3], double ikr, double cs, double sn, double amp,
                                             double zf, Cx f0) {
  const Cx pre = cmul(f0, Cx{-sn * amp, cs * amp});
  const double ikr2 = ikr * ikr, prh = rh[PD];
  const double a1 = 1.0 - ikr2, a2 = (-1.0 + 3.0 * ikr2) * prh, b2 = -3.0 * ikr * prh;
  const Cx pt1 = {pre.re * a1 - pre.im * ikr, pre.re * ikr + pre.im * a1};
  const Cx pt2 = {pre.re * a2 - pre.im * b2, pre.re * b2 + pre.im * a2};
  const Cx pt3 = {(pre.re - pre.im * ikr) * zf, (pre.im + pre.re * ikr) * zf};
  // r x p for p the unit vector of direction PD
  const double rxp0 = PD == 2 ? rh[1] : PD == 1 ? -rh[2] : 0.0;
  const double rxp1 = PD == 0 ? rh[2] : PD == 2 ? -rh[0] : 0.0;
  const double rxp2 = PD == 1 ? rh[0] : PD == 0 ? -rh[1] : 0.0;
  const double rxp[3] = {rxp0, rxp1, rxp2};
  constexpr int same = ELECTRIC ? 0 : 3, other = ELECTRIC ? 3 : 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    acc[same + d].re += pt2.re * rh[d] + (d == PD ? pt1.re : 0.0);
    acc[same + d].im += pt2.im * rh[d] + (d == PD ? pt1.im : 0.0);
    if (d != PD) {
      acc[other + d].re += pt3.re * rxp[d];
      acc[other + d].im += pt3.im * rxp[d];
    }
  }
}

// green2d (src/near2far.cpp:207-267) for one current, with H_n = J_n + i Y_n of k r; H2 from
// the recurrence H2 = (2 / x) H1 - H0 (exact for the Hankel functions)
__device__ __forceinline__ void green2d_acc(Cx acc[6], double dx, double dy, double k, double omega, double eps, double mu,
                                            double imp, int c0, Cx f0) {
  const double r = sqrt(dx * dx + dy * dy), rx = dx / r, ry = dy / r, kr = k * r;
  const Cx h0 = {j0(kr), y0(kr)}, h1 = {j1(kr), y1(kr)};
  const Cx H0 = cmul(h0, f0), H1 = cmul(h1, f0);
  const Cx ikH1 = {-0.25 * k * H1.im, 0.25 * k * H1.re};
  const bool electric = c0 < 5;
  const int pd = ff_dir(c0);
  if (pd == 2) {
    const Cx own = cscale(H0, -0.25 * omega * (electric ? mu : eps));
    if (electric) {
      acc[2].re += own.re, acc[2].im += own.im;
      acc[3].re -= ry * ikH1.re, acc[3].im -= ry * ikH1.im;
      acc[4].re += rx * ikH1.re, acc[4].im += rx * ikH1.im;
    }
    else {
      acc[0].re += ry * ikH1.re, acc[0].im += ry * ikH1.im;
      acc[1].re -= rx * ikH1.re, acc[1].im -= rx * ikH1.im;
      acc[5].re += own.re, acc[5].im += own.im;
    }
    return;
  }
  const double two_x = 2.0 / kr;
  const Cx h2 = {two_x * h1.re - h0.re, two_x * h1.im - h0.im};
  const Cx H2 = cmul(h2, f0);
  const Cx D = {H0.re - H2.re, H0.im - H2.im};
  const double px = pd == 0 ? 1.0 : 0.0, py = pd == 1 ? 1.0 : 0.0;
  const double prh = px * rx + py * ry, rxp = rx * py - ry * px;
  const double radial = prh / r * 0.25 * (electric ? imp : 1.0 / imp), across = rxp * omega * (electric ? mu : eps) * 0.125;
  const int own = electric ? 0 : 3;
  acc[own + 0].re += -(rx * radial) * H1.re + (ry * across) * D.re;
  acc[own + 0].im += -(rx * radial) * H1.im + (ry * across) * D.im;
  acc[own + 1].re += -(ry * radial) * H1.re - (rx * across) * D.re;
  acc[own + 1].im += -(ry * radial) * H1.im - (rx * across) * D.im;
  if (electric)
    acc[5].re -= rxp * ikH1.re, acc[5].im -= rxp * ikH1.im;
  else
    acc[2].re += rxp * ikH1.re, acc[2].im += rxp * ikH1.im;
}

// per-block constants of a frequency group; cyc = k / (2 pi), the wavenumber in cycles
template <int FB> struct FarGroup {
  double k[FB], omega[FB], inv_k[FB], kn4[FB], cyc[FB];
  double dcyc, eps, mu, imp;
  int nq;
  bool recur;
};

// (sin, cos) of 2 pi t: t less its nearest integer (exact), then sincospi of twice the rest
__device__ __forceinline__ void phase(double t, double &sn, double &cs) { sincospi(2.0 * (t - rint(t)), &sn, &cs); }

// The currents [jlo, jhi) of one class (direction PD, electric or magnetic): the thread's
// share, strided over the block, added to acc for each frequency of the group.
template <int DIM, int FB, int PD, bool ELECTRIC>
__device__ __forceinline__ void far_class(Cx (&acc)[FB][6], const FarArgs &a, const FarGroup<FB> &g, int jlo, int jhi,
                                          int f_first, double X, double Y, double Zc) {
  constexpr int c0 = (ELECTRIC ? 0 : 5) + (PD == 2 ? 4 : PD);
  const double zf = ELECTRIC ? 1.0 / g.imp : -g.imp, inv_em = 1.0 / (ELECTRIC ? g.eps : g.mu);
  for (int j = jlo + threadIdx.x; j < jhi; j += FF_BLOCK) {
    Cx f0[FB];
#pragma unroll
    for (int q = 0; q < FB; ++q) {
      const double2 v = q < g.nq ? a.f0[size_t(f_first + q) * a.nsrc + j] : make_double2(0.0, 0.0);
      f0[q] = {v.x, v.y};
    }
    if (DIM == 3) {
      const double dx = X - a.sx[j], dy = Y - a.sy[j], dz = Zc - a.sz[j];
      const double r2 = dx * dx + dy * dy + dz * dz, ir = rsqrt(r2), r = r2 * ir;
      const double rh[3] = {dx * ir, dy * ir, dz * ir};
      double cs = 0, sn = 0, dcs = 1, dsn = 0;
      if (g.recur) {
        phase(g.cyc[0] * r, sn, cs);
        phase(g.dcyc * r, dsn, dcs);
      }
#pragma unroll
      for (int q = 0; q < FB; ++q) {
        if (q >= g.nq) break;
        if (!g.recur)
          phase(g.cyc[q] * r, sn, cs);
        else if (q > 0) {
          const double c2 = cs * dcs - sn * dsn;
          sn = sn * dcs + cs * dsn;
          cs = c2;
        }
        green3d_term<PD, ELECTRIC>(acc[q], rh, ir * g.inv_k[q], cs, sn, g.kn4[q] * ir * inv_em, zf, f0[q]);
      }
    }
    else {
#pragma unroll
      for (int q = 0; q < FB; ++q)
        if (q < g.nq) green2d_acc(acc[q], X - a.sx[j], Y - a.sy[j], g.k[q], g.omega[q], g.eps, g.mu, g.imp, c0, f0[q]);
    }
  }
}

// Block b: the (far point, frequency group) pair b / segs and the current segment b % segs.
// A group is FB consecutive frequencies (fewer in the last group): the geometry of a current
// (distance, direction, 1/r) is computed once for the group.  In 3D, when the frequencies are
// uniformly spaced (a.uniform), e^{ikr} of the group's later frequencies follows from the
// first by repeated multiplication with e^{i dk r}, two phase evaluations per current and
// group.  The
// currents come sorted by class (Ex Ey Ez Hx Hy Hz, a.cls_lo), so each class runs a loop of
// its own with its direction and kind known at compile time.
template <int DIM, int FB> __global__ __launch_bounds__(FF_BLOCK) void k_farfield(FarArgs a, double *partial) {
  const int b = blockIdx.x;
  const int ngroup = (a.nfreq + FB - 1) / FB;
  const int pair = b / a.segs, seg = b - pair * a.segs;
  const int p = pair / ngroup, f_first = (pair - p * ngroup) * FB;
  const int per = (a.nsrc + a.segs - 1) / a.segs;
  const int lo = seg * per, hi = min(a.nsrc, lo + per);
  const double X = a.px[p], Y = a.py[p], Zc = a.pz[p];
  FarGroup<FB> g;
  g.nq = min(FB, a.nfreq - f_first);
  const double n = sqrt(a.eps * a.mu);
  g.imp = sqrt(a.mu / a.eps), g.eps = a.eps, g.mu = a.mu;
#pragma unroll
  for (int q = 0; q < FB; ++q) {
    const double fr = a.freq[f_first + min(q, g.nq - 1)];
    g.omega[q] = 2 * FF_PI * fr;
    g.k[q] = g.omega[q] * n;
    g.inv_k[q] = 1.0 / g.k[q];
    g.kn4[q] = g.k[q] * n * (0.25 / FF_PI);
    g.cyc[q] = fr * n;
  }
  g.recur = DIM == 3 && a.uniform && g.nq > 1;
  g.dcyc = g.recur ? g.cyc[FB > 1 ? 1 : 0] - g.cyc[0] : 0.0;
  Cx acc[FB][6];
#pragma unroll
  for (int q = 0; q < FB; ++q)
#pragma unroll
    for (int c = 0; c < 6; ++c) acc[q][c] = {0.0, 0.0};
  for (int cls = 0; cls < 6; ++cls) {
    const int jlo = max(lo, a.cls_lo[cls]), jhi = min(hi, a.cls_lo[cls + 1]);
    if (jlo >= jhi) continue;
    switch (cls) {
      case 0: far_class<DIM, FB, 0, true>(acc, a, g, jlo, jhi, f_first, X, Y, Zc); break;
      case 1: far_class<DIM, FB, 1, true>(acc, a, g, jlo, jhi, f_first, X, Y, Zc); break;
      case 2: far_class<DIM, FB, 2, true>(acc, a, g, jlo, jhi, f_first, X, Y, Zc); break;
      case 3: far_class<DIM, FB, 0, false>(acc, a, g, jlo, jhi, f_first, X, Y, Zc); break;
      case 4: far_class<DIM, FB, 1, false>(acc, a, g, jlo, jhi, f_first, X, Y, Zc); break;
      default: far_class<DIM, FB, 2, false>(acc, a, g, jlo, jhi, f_first, X, Y, Zc); break;
    }
  }
  const int nq = g.nq;
  // wave reduction, then across the block's waves through LDS; one partial per frequency
  __shared__ double red[FF_BLOCK / 64][12];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
#pragma unroll
  for (int q = 0; q < FB; ++q) {
    double v[12];
#pragma unroll
    for (int c = 0; c < 6; ++c) v[2 * c] = acc[q][c].re, v[2 * c + 1] = acc[q][c].im;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
      for (int c = 0; c < 12; ++c) v[c] += __shfl_down(v[c], off, 64);
    if (lane == 0)
#pragma unroll
      for (int c = 0; c < 12; ++c) red[wave][c] = v[c];
    __syncthreads();
    if (threadIdx.x < 12 && q < nq) {
      double sum = 0;
#pragma unroll
      for (int w = 0; w < FF_BLOCK / 64; ++w) sum += red[w][threadIdx.x];
      // partial layout: [point][frequency][segment][12]
      partial[((size_t(p) * a.nfreq + f_first + q) * a.segs + seg) * 12 + threadIdx.x] = sum;
    }
    __syncthreads();
  }
}

// out[pair * 12 + q] = the sum over the pair's segments, in segment order
__global__ void k_farfield_sum(const double *partial, int npair, int segs, double *out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npair * 12) return;
  const int pair = t / 12, q = t - pair * 12;
  double s = 0;
  for (int g = 0; g < segs; ++g) s += partial[(size_t(pair) * segs + g) * 12 + q];
  out[t] = s;
}

void launch_farfield(const FarArgs &a, double *partial, double *out, hipStream_t s) {
  const int fb = farfield_group(a.dim);
  const size_t nblk = size_t(a.npts) * ((a.nfreq + fb - 1) / fb) * a.segs;
  if (a.dim == 3)
    hipLaunchKernelGGL((k_farfield<3, FF_GROUP3>), dim3(unsigned(nblk)), dim3(FF_BLOCK), 0, s, a, partial);
  else
    hipLaunchKernelGGL((k_farfield<2, 1>), dim3(unsigned(nblk)), dim3(FF_BLOCK), 0, s, a, partial);
  const int npair = a.npts * a.nfreq;
  hipLaunchKernelGGL(k_farfield_sum, dim3((npair * 12 + 255) / 256), dim3(256), 0, s, partial, npair, a.segs, out);
}

} // namespace mx
