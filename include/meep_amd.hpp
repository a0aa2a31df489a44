// meep_amd.hpp — the subset of Meep's C++ API (namespace meep) whose fields::step() runs on
// the MI355X engine.  Header-only: every method forwards to the C-ABI in meep_amd.h.
//
// A program written against the reference's classes for this path — grid_volume from
// vol1d/vol2d/vol3d, structure(gv, eps, pml(...)), fields(&s), add_point_source,
// add_volume_source, use_bloch, step(), get_field, add_dft_flux / flux() — compiles against
// this header and links -lmeep_amd.  Names, argument meaning and the error behaviour
// (meep::abort -> std::runtime_error("meep: ...")) follow src/meep.hpp of NanoComp/meep; the
// line cited on each member is the declaration it mirrors.
#pragma once

#include "meep_amd.h"

#include <math.h>

#include <algorithm>
#include <chrono>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace meep {

typedef std::complex<double> cdouble;

// realnum (src/meep.hpp:43-46): float when built with MEEP_AMD_SINGLE (the reference's
// --enable-single), which also makes every structure single precision (realnum_is_float)
#ifdef MEEP_AMD_SINGLE
typedef float realnum;
#else
typedef double realnum;
#endif
const double pi = 3.141592653589793238462643383276; // src/meep.hpp:53
const double infinity = HUGE_VAL;                    // src/meep.hpp:55
// src/meep.hpp:50-51: 0 suppresses the library's non-error messages (this layer prints none)
inline int verbosity = 1;

// ---- process helpers (src/meep/mympi.hpp:29-120) for a single process: rank 0 of 1 ----
inline double wall_time() { // src/mympi.cpp:146-152
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
class initialize { // src/meep/mympi.hpp:31-39
public:
  initialize(int &argc, char **&argv) : t_start(wall_time()) { (void)argc, (void)argv; }
  double elapsed_time() { return wall_time() - t_start; }

private:
  double t_start;
};
inline void all_wait() {}
inline int count_processors() { return 1; }
inline int my_rank() { return 0; }
inline bool am_really_master() { return true; }
inline int am_master() { return 1; }
inline bool with_mpi() { return false; }
inline void master_printf(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
inline void master_printf(const char *fmt, ...) { // src/mympi.cpp:523-533
  va_list ap;
  va_start(ap, fmt);
  std::vfprintf(stdout, fmt, ap);
  va_end(ap);
  std::fflush(stdout);
}
inline void master_printf_stderr(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
inline void master_printf_stderr(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vfprintf(stderr, fmt, ap);
  va_end(ap);
}

// src/meep/vec.hpp:31-79 and src/meep.hpp:1594 (same numeric values)
enum component { Ex = 0, Ey, Er, Ep, Ez, Hx, Hy, Hr, Hp, Hz, Dx, Dy, Dr, Dp, Dz, Bx, By, Br, Bp, Bz };
enum direction { X = 0, Y, Z, R, P, NO_DIRECTION };
enum boundary_side { High = 0, Low };
enum boundary_condition { Periodic = 0, Metallic, Magnetic, None };
enum ndim { D1 = 0, D2, D3, Dcyl };

[[noreturn]] inline void abort(const char *fmt, ...) { // src/mympi.cpp:238-256
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw std::runtime_error(std::string("meep: ") + buf);
}
inline void check(int rc) {
  if (rc < 0) throw std::runtime_error(mp_last_error());
}

// vec (src/meep/vec.hpp:621-750); cylindrical vectors exist as values (veccyl) but no GPU
// path accepts them
class vec {
public:
  vec() : dim(D1) { t[0] = t[1] = t[2] = 0; }
  explicit vec(double zz) : dim(D1) { t[0] = t[1] = 0, t[2] = zz; }
  vec(double xx, double yy) : dim(D2) { t[0] = xx, t[1] = yy, t[2] = 0; }
  vec(double xx, double yy, double zz) : dim(D3) { t[0] = xx, t[1] = yy, t[2] = zz; }
  friend vec veccyl(double rr, double zz);
  double x() const { return t[0]; }
  double y() const { return t[1]; }
  double z() const { return t[2]; }
  double in_direction(direction d) const { return t[d]; }
  void set_direction(direction d, double v) { t[d] = v; }
  const double *data() const { return t; }
  ndim dim;

private:
  double t[3];
};
inline vec veccyl(double rr, double zz) { // src/meep/vec.hpp:744-749 (r kept in x)
  vec v(rr, 0.0, zz);
  v.dim = Dcyl;
  return v;
}

// grid_volume (src/meep/vec.hpp:1014-1177): size and resolution of a cell with its low
// corner at the origin
class grid_volume {
public:
  ndim dim;
  double a;
  double size[3];
  vec center() const { // grid_volume::center (src/vec.cpp:1089-1103)
    double c[3];
    for (int d = 0; d < 3; ++d) {
      int n = int(size[d] * a + 0.5);
      c[d] = (n - n % 2) * 0.5 / a;
    }
    if (dim == D1) return vec(c[2]);
    if (dim == D2) return vec(c[0], c[1]);
    return vec(c[0], c[1], c[2]);
  }
  double inva() const { return 1.0 / a; }
};
inline grid_volume volone(double zsize, double a) { return grid_volume{D1, a, {0, 0, zsize}}; }
inline grid_volume vol1d(double zsize, double a) { return volone(zsize, a); }
inline grid_volume voltwo(double xsize, double ysize, double a) { return grid_volume{D2, a, {xsize, ysize, 0}}; }
inline grid_volume vol2d(double xsize, double ysize, double a) { return voltwo(xsize, ysize, a); }
inline grid_volume vol3d(double xsize, double ysize, double zsize, double a) {
  return grid_volume{D3, a, {xsize, ysize, zsize}};
}

// volume (src/meep/vec.hpp:948-1004): a box between two corners, ordered per direction as
// volume::volume does (src/vec.cpp:163-167)
class volume {
public:
  volume(const vec &a, const vec &b) : lo(a), hi(b) {
    for (int d = 0; d < 3; ++d) {
      const direction dd = direction(d);
      lo.set_direction(dd, std::min(a.in_direction(dd), b.in_direction(dd)));
      hi.set_direction(dd, std::max(a.in_direction(dd), b.in_direction(dd)));
    }
  }
  explicit volume(const vec &p) : lo(p), hi(p) {}
  vec get_min_corner() const { return lo; }
  vec get_max_corner() const { return hi; }

private:
  vec lo, hi;
};

// boundary_region / pml() (src/meep.hpp:660-736, src/structure.cpp:252-273)
typedef double (*pml_profile_func)(double u, void *func_data);
inline double pml_quadratic_profile(double u, void *) { return u * u; }
class boundary_region {
public:
  enum boundary_region_kind { NOTHING_SPECIAL, PML };
  struct term {
    int d, side;
    double thickness, R, mean_stretch;
    pml_profile_func profile;
    void *profile_data;
    double integral, integral_u;
  };
  std::vector<term> terms;
  boundary_region() {}
  boundary_region(boundary_region_kind kind, double thickness, double Rasymptotic, double mean_stretch,
                  pml_profile_func pml_profile, void *pml_profile_data, double pml_profile_integral,
                  double pml_profile_integral_u, direction d, boundary_side side) {
    if (kind == PML)
      terms.push_back({int(d), int(side), thickness, Rasymptotic, mean_stretch, pml_profile, pml_profile_data,
                       pml_profile_integral, pml_profile_integral_u});
  }
  boundary_region operator+(const boundary_region &o) const {
    boundary_region r(*this);
    r.terms.insert(r.terms.end(), o.terms.begin(), o.terms.end());
    return r;
  }
};
inline boundary_region pml(double thickness, direction d, boundary_side side, double Rasymptotic = 1e-15,
                           double mean_stretch = 1.0) {
  return boundary_region(boundary_region::PML, thickness, Rasymptotic, mean_stretch, pml_quadratic_profile, nullptr,
                         1. / 3., 1. / 4., d, side);
}
inline boundary_region pml(double thickness, direction d, double Rasymptotic = 1e-15, double mean_stretch = 1.0) {
  return pml(thickness, d, Low, Rasymptotic, mean_stretch) + pml(thickness, d, High, Rasymptotic, mean_stretch);
}
inline boundary_region pml(double thickness, double Rasymptotic = 1e-15, double mean_stretch = 1.0) {
  return pml(thickness, X, Rasymptotic, mean_stretch) + pml(thickness, Y, Rasymptotic, mean_stretch) +
         pml(thickness, Z, Rasymptotic, mean_stretch);
}
namespace detail {
inline double profile_trampoline(const double *r, void *user) {
  const boundary_region::term *t = static_cast<const boundary_region::term *>(user);
  return t->profile(r[0], t->profile_data);
}
} // namespace detail

// lorentzian_susceptibility (src/meep.hpp:246-280)
class lorentzian_susceptibility {
public:
  lorentzian_susceptibility(double omega_0, double gamma, bool no_omega_0_denominator = false)
      : omega_0(omega_0), gamma(gamma), drude(no_omega_0_denominator) {}
  double omega_0, gamma;
  bool drude;
};

// src_time classes (src/meep.hpp:943-1080); is_integrated defaults to true as in C++
class src_time {
public:
  mp_src_time d{};
  bool is_integrated = true;
  mp_src_time desc() const {
    mp_src_time r = d;
    r.is_integrated = is_integrated ? 1 : 0;
    return r;
  }
};
class gaussian_src_time : public src_time {
public:
  gaussian_src_time(double f, double fwidth, double s = 5.0) { // src/sources.cpp:72-83
    d.kind = 0;
    d.freq = f;
    d.width = 1.0 / fwidth;
    d.start_time = 0;
    d.end_time = 2 * d.width * s;
  }
  gaussian_src_time(double f, double w, double st, double et) { // src/sources.cpp:85-96
    d.kind = 0;
    d.freq = f;
    d.width = w;
    d.start_time = st;
    d.end_time = et;
  }
};
class continuous_src_time : public src_time {
public:
  continuous_src_time(double f, double w = 0.0, double st = 0.0, double et = 1e300, double s = 3.0) {
    d.kind = 1;
    d.freq = f;
    d.width = w;
    d.start_time = st;
    d.end_time = et;
    d.slowness = s;
  }
};

namespace detail {
inline double eps_trampoline(const double *r, void *u) {
  typedef double (*fn)(const vec &);
  return reinterpret_cast<fn>(u)(vec(r[0], r[1], r[2]));
}
} // namespace detail

// realnum (src/meep.hpp:43-46) is float when the reference is configured --enable-single.
// Here the choice is made per structure at run time; this flag is the default that every
// structure created afterwards takes (initialised from MEEP_AMD_SINGLE at compile time).
inline bool &realnum_is_float() {
#ifdef MEEP_AMD_SINGLE
  static bool f = true;
#else
  static bool f = false;
#endif
  return f;
}

// symmetry (src/meep.hpp:745-812): only the identity has a GPU path.
class symmetry {
public:
  bool is_identity = true;
};
inline symmetry identity() { return symmetry(); }

// structure (src/meep.hpp:815-930)
class structure {
public:
  // structure(gv, double eps(const vec&), boundary_region, symmetry, num_chunks, Courant,
  // use_anisotropic_averaging, tol, maxeval) (src/meep.hpp:835-839); eps sees (x, y, z)
  // with absent coordinates 0.  One chunk always: num_chunks is accepted and ignored.
  structure(const grid_volume &gv, double eps(const vec &), const boundary_region &br = boundary_region(),
            const symmetry &sym = identity(), int num_chunks = 0, double Courant = 0.5,
            bool use_anisotropic_averaging = false, double tol = 1e-4, int maxeval = 100000)
      : gv_(gv) {
    (void)num_chunks, (void)tol, (void)maxeval;
    if (!sym.is_identity) meep::abort("only the identity symmetry is supported by the GPU engine");
    if (use_anisotropic_averaging) meep::abort("subpixel averaging is not supported by the GPU engine");
    std::vector<int> dirs, sides;
    std::vector<double> th, R, stretch, integ, integ_u;
    std::vector<mp_scalar_func> prof;
    std::vector<void *> user;
    for (const auto &t : br.terms) {
      dirs.push_back(t.d);
      sides.push_back(t.side);
      th.push_back(t.thickness);
      R.push_back(t.R);
      stretch.push_back(t.mean_stretch);
      prof.push_back(t.profile ? detail::profile_trampoline : nullptr);
      user.push_back(const_cast<boundary_region::term *>(&t));
      integ.push_back(t.integral);
      integ_u.push_back(t.integral_u);
    }
    int dim = gv.dim == D1 ? 1 : gv.dim == D2 ? 2 : 3;
    p = mp_structure_new_pml(dim, gv.size[0], gv.size[1], gv.size[2], gv.a, eps ? detail::eps_trampoline : nullptr,
                             reinterpret_cast<void *>(eps), int(dirs.size()), dirs.data(), sides.data(), th.data(),
                             R.data(), stretch.data(), prof.data(), user.data(), integ.data(), integ_u.data(), Courant,
                             realnum_is_float() ? 1 : 0);
    if (!p) throw std::runtime_error(mp_last_error());
  }
  ~structure() { mp_structure_delete(p); }
  structure(const structure &) = delete;
  structure &operator=(const structure &) = delete;
  // add_susceptibility(sigma, ft, sus) (src/meep.hpp:864), ft E_stuff (0) or H_stuff (1)
  void add_susceptibility(double sigma(const vec &), int ft, const lorentzian_susceptibility &sus) {
    check(mp_structure_add_susceptibility_ft(p, ft, sus.omega_0, sus.gamma, sus.drude, detail::eps_trampoline,
                                             reinterpret_cast<void *>(sigma)));
  }
  // set_conductivity(component, double C(const vec &)) (src/meep.hpp:856, src/structure.cpp:435-438)
  void set_conductivity(component c, double C(const vec &)) {
    check(mp_structure_set_conductivity(p, int(c), detail::eps_trampoline, reinterpret_cast<void *>(C)));
  }
  const grid_volume &gv() const { return gv_; }
  mp_structure *handle() { return p; }

private:
  grid_volume gv_;
  mp_structure *p;
};
constexpr int E_stuff = 0;
constexpr int H_stuff = 1;

// linspace (src/dft.cpp:30-40)
inline std::vector<double> linspace(double freq_min, double freq_max, size_t Nfreq) {
  std::vector<double> fr(Nfreq);
  if (Nfreq <= 1) {
    if (Nfreq == 1) fr[0] = (freq_min + freq_max) * 0.5;
    return fr;
  }
  const double df = (freq_max - freq_min) / (Nfreq - 1);
  for (size_t i = 0; i < Nfreq; ++i) fr[i] = freq_min + double(i) * df;
  return fr;
}

// dft_flux (src/meep.hpp:1209-1255)
class dft_flux {
public:
  dft_flux(mp_dft_flux *h, int n) : h(h), n(n) {}
  std::vector<double> flux_vec() const {
    std::vector<double> out(n);
    check(mp_dft_flux_flux(h, out.data()));
    return out;
  }
  // flux() returns a new[]-allocated array as the reference does (src/dft.cpp:542-556)
  double *flux() const {
    double *out = new double[n];
    check(mp_dft_flux_flux(h, out));
    return out;
  }
  // scale_dfts (src/dft.cpp:597-600)
  void scale_dfts(std::complex<double> scale) { check(mp_dft_flux_scale_dfts(h, scale.real(), scale.imag())); }
  int Nfreq() const { return n; }
  mp_dft_flux *handle() const { return h; }

private:
  mp_dft_flux *h;
  int n;
};

// the new[] array of fields::get_dft_array (src/dft.cpp:1259-1290), null where the reference's
// is (no chunk of the component, or a monitor on one grid point)
template <class H>
std::complex<realnum> *get_dft_array_of(long (*fn)(H *, int, int, int *, size_t *, double *), H *h, component c,
                                        int num_freq, int *rank, size_t dims[3]) {
  long n = fn(h, int(c), num_freq, rank, dims, nullptr);
  if (n < 0) throw std::runtime_error(mp_last_error());
  if (n == 0) return nullptr;
  std::vector<double> raw(size_t(2 * n));
  if (fn(h, int(c), num_freq, rank, dims, raw.data()) < 0) throw std::runtime_error(mp_last_error());
  std::complex<realnum> *out = new std::complex<realnum>[size_t(n)];
  for (long q = 0; q < n; ++q) out[q] = std::complex<realnum>(realnum(raw[2 * q]), realnum(raw[2 * q + 1]));
  return out;
}

// dft_fields (src/meep.hpp:1420-1441): accumulators per chunk, as the engine holds them
class dft_fields {
public:
  dft_fields(mp_dft_fields *h, int n) : h(h), n(n) {}
  int num_chunks() const { return mp_dft_fields_num_chunks(h); }
  component chunk_component(int i) const {
    int c = mp_dft_fields_chunk_component(h, i);
    if (c < 0) throw std::runtime_error(mp_last_error());
    return component(c);
  }
  // dft_chunk::dft of chunk i: chunk_points(i) x Nfreq complex values, frequency fastest
  std::vector<cdouble> chunk_dft(int i) const {
    long np = mp_dft_fields_chunk_points(h, i);
    std::vector<double> raw(size_t(2 * np * n));
    check(mp_dft_fields_get_dft(h, i, raw.data()));
    std::vector<cdouble> out(size_t(np * n));
    for (size_t q = 0; q < out.size(); ++q) out[q] = cdouble(raw[2 * q], raw[2 * q + 1]);
    return out;
  }
  int Nfreq() const { return n; }
  mp_dft_fields *handle() const { return h; }

private:
  mp_dft_fields *h;
  int n;
};

// sourcedata (src/meep.hpp:1323-1328): points of one chunk's component array and amplitudes
struct sourcedata {
  component near_fd_comp;
  std::vector<ptrdiff_t> idx_arr;
  int fc_idx;
  std::vector<cdouble> amp_arr;
};

// flux_vol (src/meep.hpp:2322-2352): the flux of E x H through a volume, taken by every step;
// owned by its fields as in the reference
class flux_vol {
public:
  explicit flux_vol(mp_flux_vol *h) : h(h) {}
  ~flux_vol() { mp_flux_vol_delete(h); }
  flux_vol(const flux_vol &) = delete;
  flux_vol &operator=(const flux_vol &) = delete;
  double flux() {
    double v = 0;
    check(mp_flux_vol_flux(h, &v));
    return v;
  }

private:
  mp_flux_vol *h;
};

// derived_component (src/meep/vec.hpp:57-66): the flux components name a surface's normal
enum derived_component { Sx = 100, Sy, Sr, Sp, Sz, EnergyDensity, D_EnergyDensity, H_EnergyDensity };

// volume_list (src/meep/vec.hpp:1233-1255): surfaces with a component (Sx, Sy, Sz or a field
// component naming the normal direction) and a complex weight; owns the rest of its list
class volume_list {
public:
  volume_list(const volume &v, int c, std::complex<double> weight = 1.0, volume_list *next = 0)
      : v(v), c(c), weight(weight), next(next) {}
  ~volume_list() { delete next; }
  volume_list(const volume_list &) = delete;
  volume_list &operator=(const volume_list &) = delete;
  volume v;
  int c;
  std::complex<double> weight;
  volume_list *next;
};

// dft_energy (src/meep.hpp:1258-1290): electric / magnetic / total energy spectra, new[]
// arrays of Nfreq values
class dft_energy {
public:
  dft_energy(mp_dft_energy *h, int n) : h(h), n(n) {}
  double *electric() { return get(1); }
  double *magnetic() { return get(2); }
  double *total() { return get(0); }
  int Nfreq() const { return n; }

private:
  double *get(int which) {
    double *out = new double[size_t(n)];
    if (mp_dft_energy_get(h, which, out)) {
      delete[] out;
      throw std::runtime_error(mp_last_error());
    }
    return out;
  }
  mp_dft_energy *h;
  int n;
};

// dft_force (src/meep.hpp:1292-1320): force() is a new[] array of Nfreq values
class dft_force {
public:
  dft_force(mp_dft_force *h, int n) : h(h), n(n) {}
  double *force() {
    double *out = new double[size_t(n)];
    if (mp_dft_force_force(h, out)) {
      delete[] out;
      throw std::runtime_error(mp_last_error());
    }
    return out;
  }
  int Nfreq() const { return n; }

private:
  mp_dft_force *h;
  int n;
};

// dft_near2far (src/meep.hpp:1331-1395): owned by its fields as the other monitors are
class dft_near2far {
public:
  dft_near2far(mp_near2far *h, int n) : h(h), n(n) {}
  // farfield(x): a new[] array of 6 * Nfreq values, Ex Ey Ez Hx Hy Hz per frequency
  std::complex<double> *farfield(const vec &x) {
    std::vector<double> raw(size_t(12 * n));
    check(mp_near2far_farfield(h, x.data(), raw.data()));
    std::complex<double> *out = new std::complex<double>[size_t(6 * n)];
    for (int q = 0; q < 6 * n; ++q) out[q] = std::complex<double>(raw[2 * q], raw[2 * q + 1]);
    return out;
  }
  // farfield_lowlevel(EH, x): the same into EH[6 * Nfreq]
  void farfield_lowlevel(std::complex<double> *EH, const vec &x) {
    std::complex<double> *v = farfield(x);
    std::copy(v, v + 6 * n, EH);
    delete[] v;
  }
  // get_farfields_array(where, rank, dims, N, resolution): a new[] 6 x 2 x N x Nfreq array
  double *get_farfields_array(const volume &where, int &rank, size_t *dims, size_t &N, double resolution) {
    long np = mp_near2far_get_farfields_array(h, where.get_min_corner().data(), where.get_max_corner().data(),
                                              resolution, &rank, dims, nullptr);
    if (np < 0) throw std::runtime_error(mp_last_error());
    N = size_t(np);
    double *out = new double[12 * N * size_t(n)];
    if (mp_near2far_get_farfields_array(h, where.get_min_corner().data(), where.get_max_corner().data(), resolution,
                                        &rank, dims, out) < 0) {
      delete[] out;
      throw std::runtime_error(mp_last_error());
    }
    return out;
  }
  // flux(direction, where, resolution): a new[] array of Nfreq values
  double *flux(direction df, const volume &where, double resolution) {
    double *out = new double[size_t(n)];
    if (mp_near2far_flux(h, int(df), where.get_min_corner().data(), where.get_max_corner().data(), resolution, out)) {
      delete[] out;
      throw std::runtime_error(mp_last_error());
    }
    return out;
  }
  // scale_dfts (src/near2far.cpp:124-126)
  void scale_dfts(std::complex<double> scale) { check(mp_near2far_scale_dfts(h, scale.real(), scale.imag())); }
  int Nfreq() const { return n; }
  mp_near2far *handle() const { return h; }

private:
  mp_near2far *h;
  int n;
};

// green2d / green3d (src/meep.hpp:2389-2392): EH[6] of a point current c0 of amplitude f0 at x0
inline void green_of(int dim, std::complex<double> *EH, const vec &x, double freq, double eps, double mu, const vec &x0,
                     component c0, std::complex<double> f0) {
  double amp[2] = {f0.real(), f0.imag()}, raw[12];
  check(mp_near2far_green(dim, x.data(), freq, eps, mu, x0.data(), int(c0), amp, raw));
  for (int q = 0; q < 6; ++q) EH[q] = std::complex<double>(raw[2 * q], raw[2 * q + 1]);
}
inline void green2d(std::complex<double> *EH, const vec &x, double freq, double eps, double mu, const vec &x0,
                    component c0, std::complex<double> f0) {
  green_of(2, EH, x, freq, eps, mu, x0, c0, f0);
}
inline void green3d(std::complex<double> *EH, const vec &x, double freq, double eps, double mu, const vec &x0,
                    component c0, std::complex<double> f0) {
  green_of(3, EH, x, freq, eps, mu, x0, c0, f0);
}

// fields (src/meep.hpp:1716-2320)
class fields {
public:
  // fields(structure*, m, beta, zero_fields_near_cylorigin, loop_tile_base_db,
  // loop_tile_base_eh, bfast_scaled_k) (src/meep.hpp:1740-1742): the loop tile bases tune the
  // reference's CPU loops and have no meaning here; m, beta and bfast_scaled_k select
  // formulations without a GPU path
  explicit fields(structure *s, double m = 0, double beta = 0, bool zero_fields_near_cylorigin = true,
                  int loop_tile_base_db = 0, int loop_tile_base_eh = 0,
                  std::vector<double> bfast_scaled_k = {0, 0, 0})
      : gv(s->gv()) {
    (void)zero_fields_near_cylorigin, (void)loop_tile_base_db, (void)loop_tile_base_eh;
    if (m != 0 || beta != 0) meep::abort("only m = 0, beta = 0 are supported by the GPU engine");
    for (double kq : bfast_scaled_k)
      if (kq != 0) meep::abort("the bfast formulation (bfast_scaled_k != 0) is not supported by the GPU engine");
    p = mp_fields_new(s->handle());
    if (!p) throw std::runtime_error(mp_last_error());
    dt = mp_fields_dt(p);
  }
  double dt; // the time step Courant / a (fields::dt, src/meep.hpp:1725)
  ~fields() {
    for (flux_vol *v : flux_vols) delete v;
    for (mp_dft_flux *h : fluxes) mp_dft_flux_delete(h);
    for (mp_dft_fields *h : dft_fields_list) mp_dft_fields_delete(h);
    for (mp_near2far *h : near2far_list) mp_near2far_delete(h);
    for (mp_dft_energy *h : energy_list) mp_dft_energy_delete(h);
    for (mp_dft_force *h : force_list) mp_dft_force_delete(h);
    mp_fields_delete(p);
  }
  fields(const fields &) = delete;
  fields &operator=(const fields &) = delete;

  void use_real_fields() { check(mp_fields_use_real_fields(p)); }                       // :1746
  void use_bloch(direction d, cdouble k) { check(mp_fields_use_bloch_dir(p, d, k.real(), k.imag())); } // :1762
  void use_bloch(const vec &k) { // :1763
    double kk[3];
    int j = 0;
    if (gv.dim == D1)
      kk[j++] = k.z();
    else {
      kk[j++] = k.x();
      kk[j++] = k.y();
      if (gv.dim == D3) kk[j++] = k.z();
    }
    check(mp_fields_use_bloch(p, kk));
  }
  void set_boundary(boundary_side b, direction d, boundary_condition c) { check(mp_fields_set_boundary(p, b, d, c)); } // :1761
  // deprecated add_point_source (src/meep.hpp:1888-1889)
  void add_point_source(component c, double freq, double width, double peaktime, double cutoff, const vec &pos,
                        cdouble amp = 1.0, int is_continuous = 0) {
    check(mp_fields_add_point_source_legacy(p, c, freq, width, peaktime, cutoff, pos.data(), amp.real(), amp.imag(),
                                            is_continuous));
  }
  void add_point_source(component c, const src_time &src, const vec &pos, cdouble amp = 1.0) { // :1890
    mp_src_time d = src.desc();
    check(mp_fields_add_volume_source(p, c, &d, pos.data(), pos.data(), amp.real(), amp.imag(), nullptr, nullptr));
  }
  void add_volume_source(component c, const src_time &src, const volume &where, cdouble amp = 1.0) { // :1893
    mp_src_time d = src.desc();
    check(mp_fields_add_volume_source(p, c, &d, where.get_min_corner().data(), where.get_max_corner().data(), amp.real(),
                                      amp.imag(), nullptr, nullptr));
  }
  // add_volume_source with an amplitude array (:1895-1896, src/sources.cpp:394-417)
  void add_volume_source(component c, const src_time &src, const volume &where, cdouble *arr, size_t dim1, size_t dim2,
                         size_t dim3, cdouble amp) {
    mp_src_time d = src.desc();
    std::vector<double> flat(2 * dim1 * dim2 * dim3);
    for (size_t q = 0; q < dim1 * dim2 * dim3; ++q) flat[2 * q] = arr[q].real(), flat[2 * q + 1] = arr[q].imag();
    check(mp_fields_add_volume_source_array(p, c, &d, where.get_min_corner().data(), where.get_max_corner().data(),
                                            flat.data(), dim1, dim2, dim3, amp.real(), amp.imag()));
  }
  // add_srcdata (:1907-1908, src/sources.cpp:314-334); one chunk, so fc_idx must be 0
  void add_srcdata(sourcedata cur_data, src_time *src, size_t n, cdouble *amp_arr, bool needs_boundary_fix) {
    if (cur_data.fc_idx != 0) meep::abort("fields chunk index out of range");
    if (n == 0) {
      n = cur_data.idx_arr.size();
      amp_arr = cur_data.amp_arr.data();
    }
    if (n != cur_data.idx_arr.size()) meep::abort("add_srcdata: %zu amplitudes for %zu points", n, cur_data.idx_arr.size());
    std::vector<long> idx(cur_data.idx_arr.begin(), cur_data.idx_arr.end());
    std::vector<double> amp(2 * n);
    for (size_t q = 0; q < n; ++q) amp[2 * q] = amp_arr[q].real(), amp[2 * q + 1] = amp_arr[q].imag();
    mp_src_time d = src->desc();
    check(mp_fields_add_srcdata(p, cur_data.near_fd_comp, &d, long(n), idx.data(), amp.data(), needs_boundary_fix));
  }
  void require_source_components() { check(mp_fields_require_source_components(p)); } // :1907
  void step() { check(mp_fields_step(p)); }                // :1871
  double time() const { return mp_fields_time(p); }        // :1877
  double round_time() const { return mp_fields_round_time(p); } // :1876
  double last_source_time() { return mp_fields_last_source_time(p); } // :1887
  cdouble get_field(component c, const vec &loc) const {   // :2183
    double re = 0, im = 0;
    check(mp_fields_get_field(p, c, loc.data(), &re, &im));
    return cdouble(re, im);
  }
  // add_dft_flux(direction, volume, freq, Nfreq, use_symmetry, centered_grid,
  // decimation_factor) and its overloads (src/meep.hpp:2044-2059); with the identity symmetry
  // use_symmetry changes nothing
  dft_flux add_dft_flux(direction d, const volume &where, const double *freq, size_t Nfreq, bool use_symmetry = true,
                        bool centered_grid = true, int decimation_factor = 0) {
    mp_dft_flux *h = mp_fields_add_dft_flux(p, d, where.get_min_corner().data(), where.get_max_corner().data(), freq,
                                            int(Nfreq), use_symmetry, centered_grid, decimation_factor);
    if (!h) throw std::runtime_error(mp_last_error());
    fluxes.push_back(h);
    return dft_flux(h, int(Nfreq));
  }
  dft_flux add_dft_flux(direction d, const volume &where, const std::vector<double> &freq, bool use_symmetry = true,
                        bool centered_grid = true, int decimation_factor = 0) {
    return add_dft_flux(d, where, freq.data(), freq.size(), use_symmetry, centered_grid, decimation_factor);
  }
  dft_flux add_dft_flux(direction d, const volume &where, double freq_min, double freq_max, int Nfreq,
                        bool use_symmetry = true, bool centered_grid = true, int decimation_factor = 0) {
    return add_dft_flux(d, where, linspace(freq_min, freq_max, Nfreq), use_symmetry, centered_grid, decimation_factor);
  }
  // add_dft_fields(components, num_components, volume, freq, Nfreq, use_centered_grid,
  // decimation_factor, persist) and overloads (src/meep.hpp:2078-2093); the accumulators live
  // as long as the fields object, so persist changes nothing
  dft_fields add_dft_fields(const component *cs, int num_components, const volume &where, const double *freq, size_t Nfreq,
                            bool use_centered_grid = true, int decimation_factor = 0, bool persist = false) {
    (void)persist;
    std::vector<int> ci(cs, cs + num_components);
    mp_dft_fields *h = mp_fields_add_dft_fields(p, ci.data(), num_components, where.get_min_corner().data(),
                                                where.get_max_corner().data(), freq, int(Nfreq), use_centered_grid,
                                                decimation_factor);
    if (!h) throw std::runtime_error(mp_last_error());
    dft_fields_list.push_back(h);
    return dft_fields(h, int(Nfreq));
  }
  dft_fields add_dft_fields(const component *cs, int num_components, const volume &where, const std::vector<double> &freq,
                            bool use_centered_grid = true, int decimation_factor = 0, bool persist = false) {
    return add_dft_fields(cs, num_components, where, freq.data(), freq.size(), use_centered_grid, decimation_factor,
                          persist);
  }
  dft_fields add_dft_fields(const component *cs, int num_components, const volume &where, double freq_min, double freq_max,
                            int Nfreq, bool use_centered_grid = true, int decimation_factor = 0, bool persist = false) {
    return add_dft_fields(cs, num_components, where, linspace(freq_min, freq_max, Nfreq), use_centered_grid,
                          decimation_factor, persist);
  }
  // get_dft_array(dft_flux | dft_fields, component, num_freq, rank, dims) (src/meep.hpp:2121-2124):
  // a new[] array of prod(dims) values, or null
  std::complex<realnum> *get_dft_array(dft_flux flux, component c, int num_freq, int *rank, size_t dims[3]) {
    return get_dft_array_of(mp_dft_flux_get_dft_array, flux.handle(), c, num_freq, rank, dims);
  }
  std::complex<realnum> *get_dft_array(dft_fields fdft, component c, int num_freq, int *rank, size_t dims[3]) {
    return get_dft_array_of(mp_dft_fields_get_dft_array, fdft.handle(), c, num_freq, rank, dims);
  }
  std::complex<realnum> *get_dft_array(dft_near2far n2f, component c, int num_freq, int *rank, size_t dims[3]) {
    return get_dft_array_of(mp_near2far_get_dft_array, n2f.handle(), c, num_freq, rank, dims);
  }
  // add_dft_force(where, freq, Nfreq, decimation_factor) (src/stress.cpp:153-198): each
  // volume's component names the force direction (Sx, Sy, Sz or an E component)
  dft_force add_dft_force(const volume_list *where, const double *freq, size_t Nfreq, int decimation_factor = 0) {
    std::vector<int> comps;
    std::vector<double> mins, maxs, weights;
    for (const volume_list *w = where; w; w = w->next) {
      const vec lo = w->v.get_min_corner(), hi = w->v.get_max_corner();
      comps.push_back(w->c);
      mins.insert(mins.end(), lo.data(), lo.data() + 3);
      maxs.insert(maxs.end(), hi.data(), hi.data() + 3);
      weights.push_back(w->weight.real());
      weights.push_back(w->weight.imag());
    }
    mp_dft_force *h = mp_fields_add_dft_force(p, int(comps.size()), comps.data(), mins.data(), maxs.data(), weights.data(),
                                              freq, int(Nfreq), decimation_factor);
    if (!h) throw std::runtime_error(mp_last_error());
    force_list.push_back(h);
    return dft_force(h, int(Nfreq));
  }
  dft_force add_dft_force(const volume_list *where, double freq_min, double freq_max, int Nfreq, int decimation_factor = 0) {
    const std::vector<double> fr = linspace(freq_min, freq_max, Nfreq);
    return add_dft_force(where, fr.data(), fr.size(), decimation_factor);
  }
  // add_dft_energy(where, freq, Nfreq, decimation_factor) (src/dft.cpp:725-750)
  dft_energy add_dft_energy(const volume_list *where, const double *freq, size_t Nfreq, int decimation_factor = 0) {
    std::vector<double> mins, maxs;
    for (const volume_list *w = where; w; w = w->next) {
      const vec lo = w->v.get_min_corner(), hi = w->v.get_max_corner();
      mins.insert(mins.end(), lo.data(), lo.data() + 3);
      maxs.insert(maxs.end(), hi.data(), hi.data() + 3);
    }
    mp_dft_energy *h = mp_fields_add_dft_energy(p, int(mins.size() / 3), mins.data(), maxs.data(), freq, int(Nfreq),
                                                decimation_factor);
    if (!h) throw std::runtime_error(mp_last_error());
    energy_list.push_back(h);
    return dft_energy(h, int(Nfreq));
  }
  dft_energy add_dft_energy(const volume_list *where, double freq_min, double freq_max, int Nfreq,
                            int decimation_factor = 0) {
    const std::vector<double> fr = linspace(freq_min, freq_max, Nfreq);
    return add_dft_energy(where, fr.data(), fr.size(), decimation_factor);
  }
  // add_dft_near2far(where, freq, Nfreq, decimation_factor, Nperiods) (src/meep.hpp:2161-2168):
  // each surface's normal is the direction of its component (Sx, Sy, Sz or a field component),
  // else the volume's normal
  dft_near2far add_dft_near2far(const volume_list *where, const double *freq, size_t Nfreq, int decimation_factor = 0,
                                int Nperiods = 1) {
    std::vector<int> dirs;
    std::vector<double> mins, maxs, weights;
    for (const volume_list *w = where; w; w = w->next) {
      int d = -1;
      if (w->c == Sx || w->c == Ex || w->c == Hx || w->c == Dx || w->c == Bx) d = X;
      if (w->c == Sy || w->c == Ey || w->c == Hy || w->c == Dy || w->c == By) d = Y;
      if (w->c == Sz || w->c == Ez || w->c == Hz || w->c == Dz || w->c == Bz) d = Z;
      dirs.push_back(d);
      const vec lo = w->v.get_min_corner(), hi = w->v.get_max_corner();
      mins.insert(mins.end(), lo.data(), lo.data() + 3);
      maxs.insert(maxs.end(), hi.data(), hi.data() + 3);
      weights.push_back(w->weight.real());
      weights.push_back(w->weight.imag());
    }
    mp_near2far *h = mp_fields_add_dft_near2far(p, int(dirs.size()), dirs.data(), mins.data(), maxs.data(), weights.data(),
                                                freq, int(Nfreq), decimation_factor, Nperiods);
    if (!h) throw std::runtime_error(mp_last_error());
    near2far_list.push_back(h);
    return dft_near2far(h, int(Nfreq));
  }
  dft_near2far add_dft_near2far(const volume_list *where, const std::vector<double> &freq, int decimation_factor = 0,
                                int Nperiods = 1) {
    return add_dft_near2far(where, freq.data(), freq.size(), decimation_factor, Nperiods);
  }
  dft_near2far add_dft_near2far(const volume_list *where, double freq_min, double freq_max, int Nfreq,
                                int decimation_factor = 0, int Nperiods = 1) {
    return add_dft_near2far(where, linspace(freq_min, freq_max, Nfreq), decimation_factor, Nperiods);
  }
  double field_energy() { // :2196
    double e = 0;
    check(mp_fields_field_energy(p, &e));
    return e;
  }
  // the energy in a box (:2196-2201, src/energy_and_flux.cpp:54-95)
  double field_energy_in_box(const volume &where) { return energy_in_box(0, where); }
  double electric_energy_in_box(const volume &where) { return energy_in_box(1, where); }
  double magnetic_energy_in_box(const volume &where) { return energy_in_box(2, where); }
  // time-domain flux (:2203-2207, src/energy_and_flux.cpp:186-245)
  double flux_in_box(direction d, const volume &where) {
    double v = 0;
    check(mp_fields_flux_in_box(p, d, where.get_min_corner().data(), where.get_max_corner().data(), &v));
    return v;
  }
  double flux_in_box_wrongH(direction d, const volume &where) {
    double v = 0;
    check(mp_fields_flux_in_box_wrongH(p, d, where.get_min_corner().data(), where.get_max_corner().data(), &v));
    return v;
  }
  flux_vol *add_flux_vol(direction d, const volume &where) { return add_flux(int(d), where); }
  flux_vol *add_flux_plane(const volume &where) { return add_flux(-1, where); }
  flux_vol *add_flux_plane(const vec &p1, const vec &p2) { return add_flux(-1, volume(p1, p2)); }
  void print_times() { // :1879; the phases' wall times as the engine accounts them
    double t7[7];
    check(mp_fields_get_times(p, t7));
    static const char *names[7] = {"stepping B", "", "stepping D", "", "boundaries", "dft", "total"};
    for (int q = 0; q < 7; ++q)
      if (names[q][0]) std::printf("    %-12s %g s\n", names[q], t7[q]);
  }
  mp_fields *handle() { return p; }

private:
  double energy_in_box(int which, const volume &where) {
    double v = 0;
    check(mp_fields_energy_in_box(p, which, where.get_min_corner().data(), where.get_max_corner().data(), &v));
    return v;
  }
  flux_vol *add_flux(int d, const volume &where) {
    mp_flux_vol *h = mp_fields_add_flux_vol(p, d, where.get_min_corner().data(), where.get_max_corner().data());
    if (!h) throw std::runtime_error(mp_last_error());
    flux_vols.push_back(new flux_vol(h));
    return flux_vols.back();
  }
  grid_volume gv;
  mp_fields *p;
  std::vector<mp_dft_flux *> fluxes;
  std::vector<mp_dft_fields *> dft_fields_list;
  std::vector<mp_near2far *> near2far_list;
  std::vector<mp_dft_energy *> energy_list;
  std::vector<mp_dft_force *> force_list;
  std::vector<flux_vol *> flux_vols;
};

} // namespace meep
