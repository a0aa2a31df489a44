// near2far.hip — the near-to-far transformation's field sums on gfx950.
//
// dft_near2far::farfield_lowlevel (src/near2far.cpp:350-397) sums, for every far point and
// frequency, the homogeneous-medium Green's function of every equivalent surface current (a
// DFT accumulator of a tangential field, at its grid location and its periodic copies).  The
// work is a dense (far points x frequencies) by (surface currents) product of a transcendental
// kernel: FP64 vector arithmetic (sincos, or Bessel functions in 2D), no contraction to feed
// MFMA.  Each block takes one (far point, frequency) pair and one segment of the current list;
// its 256 threads stride over the segment, keep the six complex fields in registers, and
// reduce them through the wave and LDS into one partial per block.  A second pass adds the
// partials of each pair in segment order, so the result does not depend on timing.
//
// The currents come as SoA lists prepared on the host: position x, y, z, the current's
// component (the reference's component numbering), and one amplitude per frequency
// (frequency-major, so a block's reads of one frequency are contiguous).  The periodic copies
// are separate entries whose amplitude carries their Bloch phase.
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

// green3d (src/near2far.cpp:133-186) for one current: the fields of the current's kind are
// pre (t1 p + t2 r), the other kind pre t3 (r x p) over (J) or times (magnetic current,
// negated) the impedance, with 1/(ikr) = -i/(kr) and 1/(ikr)^2 = -1/(kr)^2 expanded.
__device__ __forceinline__ void green3d_acc(Cx acc[6], double dx, double dy, double dz, double k, double kn, double eps,
                                            double mu, double imp, int c0, Cx f0) {
  const double r = sqrt(dx * dx + dy * dy + dz * dz), ir = 1.0 / r;
  const double rh[3] = {dx * ir, dy * ir, dz * ir};
  const double kr = k * r, ikr = 1.0 / kr, ikr2 = ikr * ikr;
  double sn, cs;
  sincos(kr, &sn, &cs);
  const bool electric = c0 < 5;
  // i e^{ikr} k n / (4 pi r) f0 / (eps or mu)
  const double amp = kn * ir * (0.25 / FF_PI) / (electric ? eps : mu);
  const Cx pre = cmul(f0, Cx{-sn * amp, cs * amp});
  const int pd = ff_dir(c0);
  const double prh = rh[pd];
  const Cx t1 = {1.0 - ikr2, ikr}, t2 = {(-1.0 + 3.0 * ikr2) * prh, -3.0 * ikr * prh}, t3 = {1.0, ikr};
  const Cx pt1 = cmul(pre, t1), pt2 = cmul(pre, t2), pt3 = cmul(pre, t3);
  // r x p for p the unit vector of direction pd
  double rxp[3];
  rxp[0] = pd == 2 ? rh[1] : pd == 1 ? -rh[2] : 0.0;
  rxp[1] = pd == 0 ? rh[2] : pd == 2 ? -rh[0] : 0.0;
  rxp[2] = pd == 1 ? rh[0] : pd == 0 ? -rh[1] : 0.0;
  const int same = electric ? 0 : 3, other = electric ? 3 : 0;
  const double zf = electric ? 1.0 / imp : -imp;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    Cx v = cscale(pt2, rh[d]);
    if (d == pd) v.re += pt1.re, v.im += pt1.im;
    acc[same + d].re += v.re;
    acc[same + d].im += v.im;
    acc[other + d].re += pt3.re * rxp[d] * zf;
    acc[other + d].im += pt3.im * rxp[d] * zf;
  }
}

// green2d (src/near2far.cpp:207-267) for one current, with H_n = J_n + i Y_n of k r
__device__ __forceinline__ void green2d_acc(Cx acc[6], double dx, double dy, double k, double omega, double eps, double mu,
                                            double imp, int c0, Cx f0) {
  const double r = sqrt(dx * dx + dy * dy), rx = dx / r, ry = dy / r, kr = k * r;
  const Cx H0 = cmul(Cx{j0(kr), y0(kr)}, f0), H1 = cmul(Cx{j1(kr), y1(kr)}, f0);
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
  const Cx H2 = cmul(Cx{jn(2, kr), yn(2, kr)}, f0);
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

template <int DIM> __global__ __launch_bounds__(FF_BLOCK) void k_farfield(FarArgs a, double *partial) {
  const int b = blockIdx.x;
  const int pair = b / a.segs, seg = b - pair * a.segs;
  const int p = pair / a.nfreq, fi = pair - p * a.nfreq;
  const int per = (a.nsrc + a.segs - 1) / a.segs;
  const int lo = seg * per, hi = min(a.nsrc, lo + per);
  const double X = a.px[p], Y = a.py[p], Zc = a.pz[p];
  const double freq = a.freq[fi], omega = 2 * FF_PI * freq, n = sqrt(a.eps * a.mu), k = omega * n;
  const double imp = sqrt(a.mu / a.eps);
  const double2 *f0 = a.f0 + size_t(fi) * a.nsrc;
  Cx acc[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) acc[q] = {0.0, 0.0};
  for (int j = lo + threadIdx.x; j < hi; j += FF_BLOCK) {
    const double2 v = f0[j];
    const Cx amp = {v.x, v.y};
    if (DIM == 3)
      green3d_acc(acc, X - a.sx[j], Y - a.sy[j], Zc - a.sz[j], k, k * n, a.eps, a.mu, imp, a.comp[j], amp);
    else
      green2d_acc(acc, X - a.sx[j], Y - a.sy[j], k, omega, a.eps, a.mu, imp, a.comp[j], amp);
  }
  // wave reduction, then across the block's waves through LDS
  double v[12];
#pragma unroll
  for (int q = 0; q < 6; ++q) v[2 * q] = acc[q].re, v[2 * q + 1] = acc[q].im;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int q = 0; q < 12; ++q) v[q] += __shfl_down(v[q], off, 64);
  __shared__ double red[FF_BLOCK / 64][12];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < 12; ++q) red[wave][q] = v[q];
  __syncthreads();
  if (threadIdx.x < 12) {
    double s = 0;
#pragma unroll
    for (int w = 0; w < FF_BLOCK / 64; ++w) s += red[w][threadIdx.x];
    partial[size_t(b) * 12 + threadIdx.x] = s;
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
  const size_t nblk = size_t(a.npts) * a.nfreq * a.segs;
  if (a.dim == 3)
    hipLaunchKernelGGL(k_farfield<3>, dim3(unsigned(nblk)), dim3(FF_BLOCK), 0, s, a, partial);
  else
    hipLaunchKernelGGL(k_farfield<2>, dim3(unsigned(nblk)), dim3(FF_BLOCK), 0, s, a, partial);
  const int npair = a.npts * a.nfreq;
  hipLaunchKernelGGL(k_farfield_sum, dim3((npair * 12 + 255) / 256), dim3(256), 0, s, partial, npair, a.segs, out);
}

} // namespace mx
