// The known-answer cases of the reference (tests/known_results.cpp:27-157) as a program of
// our own against the C++ API layer include/meep_amd.hpp, with a case filter and a --single
// switch (float realnum at run time).  It builds against either implementation of the
// C-ABI: the MI355X engine (libmeep_amd.so) or the CPU oracle (libmeep_oracle.so).  The
// reference's own file, compiled unmodified against include/meep.hpp, is run by
// tests/test_abi.py (oracle) and tests/test_gpu_cpp_api.py (engine).
//
// Usage: known_results [case-name-substring]   (runs every case when no argument is given)
// Exit status 0 when every selected case matches its golden value.
#include "meep_amd.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>

using namespace meep;

static double dpml = 1.0;

static double one(const vec &) { return 1.0; }

static double rods(const vec &r) { // rods (tests/known_results.cpp:29-41): period-1 lattice
  double px = r.x() - std::floor(r.x() + 0.5), py = r.y() - std::floor(r.y() + 0.5);
  return px * px + py * py < 0.3 ? 12.0 : 1.0;
}

static double sigma_polariton(const vec &) { return 7.63; }

static const cdouble src_amp(0, -2 * pi * 0.2);

static double run_ez(fields &f, const grid_volume &gv, double ttot) {
  f.add_point_source(Ez, 0.2, 3.0, 0.0, 2.0, gv.center(), src_amp);
  while (f.round_time() < ttot) f.step();
  return real(f.get_field(Ez, gv.center()));
}

static double using_pml_ez(const grid_volume &gv, double eps(const vec &)) {
  structure s(gv, eps, pml(dpml));
  fields f(&s);
  return run_ez(f, gv, 30.0);
}

static double x_periodic_y_pml(const grid_volume &gv, double eps(const vec &)) {
  structure s(gv, eps, pml(dpml, Y));
  fields f(&s);
  f.use_bloch(X, 0.1);
  return run_ez(f, gv, 30.0);
}

static double x_periodic(const grid_volume &gv, double eps(const vec &)) {
  structure s(gv, eps);
  fields f(&s);
  f.use_bloch(X, 0.1);
  return run_ez(f, gv, 30.0);
}

static double periodic_ez(const grid_volume &gv, double eps(const vec &)) {
  structure s(gv, eps);
  fields f(&s);
  f.use_bloch(gv.dim == D1 ? vec(0.3) : gv.dim == D2 ? vec(0.3, 0.4) : vec(0.3, 0.5, 0.8));
  return run_ez(f, gv, 30.0);
}

static double metallic_ez(const grid_volume &gv, double eps(const vec &)) {
  structure s(gv, eps);
  fields f(&s);
  return run_ez(f, gv, 10.0);
}

static double polariton_ex(const grid_volume &gv, double eps(const vec &)) {
  structure s(gv, eps);
  s.add_susceptibility(sigma_polariton, E_stuff, lorentzian_susceptibility(0.3, 0.1));
  fields f(&s);
  f.add_point_source(Ex, 0.2, 3.0, 0.0, 2.0, gv.center(), src_amp);
  while (f.round_time() < 10.0) f.step();
  return real(f.get_field(Ex, gv.center()));
}

static double polariton_energy(const grid_volume &gv, double eps(const vec &)) {
  structure s(gv, eps);
  s.add_susceptibility(sigma_polariton, E_stuff, lorentzian_susceptibility(0.3, 0.1));
  fields f(&s);
  f.add_point_source(Ex, 0.2, 3.0, 0.0, 2.0, gv.center(), src_amp);
  while (f.round_time() < 10.0) f.step();
  return f.field_energy();
}

struct known_case {
  const char *name;
  double golden;
  double (*run)(const grid_volume &, double (*)(const vec &));
  grid_volume gv;
  double (*eps)(const vec &);
};

int main(int argc, char **argv) {
  const char *filter = nullptr;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--single"))
      realnum_is_float() = true;
    else
      filter = argv[i];
  }
  const double a = 10.0;
  // golden values: tests/known_results.cpp:143-156
  const known_case cases[] = {
      {"1D polariton", -0.0894851, polariton_ex, volone(1.0, a), one},
      {"1D polariton energy", 0.0863443, polariton_energy, volone(1.0, a), one},
      {"1x1 metallic 2D TM", 5.20605, metallic_ez, voltwo(1.0, 1.0, a), one},
      {"1x1 PML 2D TM", 0.883776, using_pml_ez, voltwo(1.0 + 2 * dpml, 1.0 + 2 * dpml, a), one},
      {"1x1 X periodic 2D TM", 0.110425, x_periodic, voltwo(1.0, 1.0, a), one},
      {"1x1 fully periodic 2D TM rods", -4.78767, periodic_ez, voltwo(1.0, 3.0, a), rods},
      {"1x1 fully periodic 2D TM", 1.12502, periodic_ez, voltwo(1.0, 3.0, a), one},
      {"1x1 X periodic Y PML 2D TM", 0.608815, x_periodic_y_pml, voltwo(1.0, 1.0 + 2 * dpml, a), one},
      {"1x1x1 metallic 3D", -41.8057, metallic_ez, vol3d(1.0, 1.0, 1.0, a), one},
      {"1x1x1 X periodic 3D", -100.758, x_periodic, vol3d(1.0, 1.0, 1.0, a), one},
      {"1x1x1 X periodic Y PML 3D", -101.398, x_periodic_y_pml, vol3d(1.0, 1.0 + 2 * dpml, 1.0, a), one},
      {"1x1x1 fully periodic 3D rods", -103.844, periodic_ez, vol3d(1.0, 1.0, 1.0, a), rods},
      {"1x1x1 fully periodic 3D", -99.1618, periodic_ez, vol3d(1.0, 1.0, 1.0, a), one},
  };
  // tests/known_results.cpp:43-49
  const double thresh = realnum_is_float() ? 1e-4 : 1e-5;
  int failed = 0, ran = 0;
  std::printf("engine %s, %s precision\n", mp_engine_name(), realnum_is_float() ? "single" : "double");
  for (const known_case &k : cases) {
    if (filter && !std::strstr(k.name, filter)) continue;
    ++ran;
    try {
      double v = k.run(k.gv, k.eps);
      bool ok = std::fabs(v - k.golden) <= std::fabs(k.golden) * thresh && v == v;
      std::printf("%s %s (%.7g, golden %g, relerr %.2g)\n", ok ? "Passed" : "Failed", k.name, v, k.golden,
                  std::fabs(v - k.golden) / std::fabs(k.golden));
      failed += !ok;
    } catch (const std::exception &e) {
      std::printf("Failed %s: %s\n", k.name, e.what());
      ++failed;
    }
    std::fflush(stdout);
  }
  if (!ran) {
    std::printf("no case matches %s\n", filter);
    return 2;
  }
  return failed ? 1 : 0;
}
