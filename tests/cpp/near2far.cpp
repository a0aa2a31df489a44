// The near-to-far check of the reference's tests/near2far.cpp (check_2d_3d, its NEAR2FAR part)
// written against the C++ API, with the continuous-wave solve replaced by DFTs at one
// frequency: the transform of a closed box of near2far surfaces around a point current equals
// the fields the run itself accumulates at points outside the box, within the reference's
// tolerance 0.05 * 30 / resolution.  Usage: near2far [Ez|Hz]; exit status 0 when it passes.
#include "meep_amd.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>
using namespace meep;

static double two(const vec &) { return 2.0; }

int main(int argc, char **argv) {
  const component c0 = argc > 1 && std::strcmp(argv[1], "Hz") == 0 ? Hz : Ez;
  const double a = 12, L = 0.8, c = 3.0, w = 0.5;
  structure s(voltwo(6.0, 6.0, a), two, pml(1.0));
  fields f(&s);
  gaussian_src_time src(w, 0.4);
  src.is_integrated = false;
  f.add_point_source(c0, src, vec(c, c));
  volume_list vl(volume(vec(c - L, c + L), vec(c + L, c + L)), Sy, 1.0,
                 new volume_list(volume(vec(c + L, c + L), vec(c + L, c - L)), Sx, 1.0,
                                 new volume_list(volume(vec(c - L, c - L), vec(c + L, c - L)), Sy, -1.0,
                                                 new volume_list(volume(vec(c - L, c - L), vec(c - L, c + L)), Sx, -1.0))));
  dft_near2far n2f = f.add_dft_near2far(&vl, w, w, 1, 1);
  const component cs[6] = {Ex, Ey, Ez, Hx, Hy, Hz};
  std::vector<vec> pts;
  std::vector<dft_fields> mons;
  for (int i = 0; i < 6; ++i) {
    const double r = L + 0.3 + 0.1 * i;
    pts.push_back(vec(c + r * 2 / std::sqrt(5.0), c + r / std::sqrt(5.0)));
    mons.push_back(f.add_dft_fields(cs, 6, volume(pts.back()), w, w, 1, true, 1));
  }
  while (f.time() < 40.0) f.step();
  double diff = 0, dot = 0;
  for (int i = 0; i < 6; ++i) {
    std::complex<double> EH[6];
    n2f.farfield_lowlevel(EH, pts[i]);
    for (int k = 0; k < 6; ++k) {
      int rank;
      size_t dims[3];
      std::complex<realnum> *v = f.get_dft_array(mons[i], cs[k], 0, &rank, dims);
      const std::complex<double> direct = v ? std::complex<double>(v[0]) : 0.0;
      delete[] v;
      diff += std::norm(EH[k] - direct);
      dot += std::norm(direct);
    }
  }
  const double relerr = std::sqrt(diff / dot), tol = 0.05 * 30 / a;
  master_printf("NEAR2FAR %s: relerr = %g (tolerance %g)\n", c0 == Ez ? "Ez" : "Hz", relerr, tol);
  return relerr <= tol ? 0 : 1;
}
