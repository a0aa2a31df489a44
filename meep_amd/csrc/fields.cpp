// fields.cpp — host side of the MI355X engine: structure setup, the per-step launch plan and
// the C-ABI of include/meep_amd.h.
//
// One fields object is one chunk on one GPU with one HIP stream.  State arrays live in HBM in
// the reference layout; the host keeps only the plan (component descriptors, halo tables,
// sparse sources, DFT point lists) and never touches field values during a step.  There is
// no CPU path: every entry point that advances or reads fields goes through the kernels in
// kernels.hip and fails with an error when no GPU is usable.
#include "../../include/meep_amd.h"
#include "engine.hpp"
#include "grid.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <chrono>
#include <complex>
#include <cstdarg>
#include <cfloat>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

namespace mx {

typedef std::complex<double> cdouble;
const double pi = 3.141592653589793238462643383279502884197;
enum { E_ft = 0, H_ft = 1, D_ft = 2, B_ft = 3 };
enum { Periodic = 0, Metallic = 1, Magnetic = 2, None = 3 };
enum { High = 0, Low = 1 };

thread_local std::string g_last_error;

struct mx_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};
[[noreturn]] static void fail(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw mx_error(std::string("meep: ") + buf);
}
#define HIPCHECK(x)                                                                                \
  do {                                                                                             \
    hipError_t e_ = (x);                                                                           \
    if (e_ != hipSuccess) fail("HIP error %s at %s:%d", hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

bool Geom::interpolate(int c, const V3 &pc, IV locs[8], double w[8]) const {
  const double SMALL = 1e-13;
  V3 p;
  for (int d = 0; d < 3; ++d)
    if (has(d)) p.t[d] = (pc.t[d] - ysh(c, d) * 0.5 * inva) * a;
  IV mid;
  for (int d = 0; d < 3; ++d)
    if (has(d)) mid.t[d] = int(std::floor(p.t[d])) * 2 + 1 + ysh(c, d);
  V3 midv = loc(mid), dv;
  for (int d = 0; d < 3; ++d)
    if (has(d)) dv.t[d] = (pc.t[d] - midv.t[d]) * (2 * a);
  int have = 1;
  for (int i = 0; i < 8; ++i) {
    locs[i] = round_vec(midv);
    w[i] = 1.0;
  }
  for (int d = 0; d < 3; ++d) {
    if (!has(d)) continue;
    for (int i = 0; i < have; ++i) {
      locs[have + i] = locs[i];
      w[have + i] = w[i];
      locs[i].t[d] = mid.t[d] - 1;
      w[i] *= 0.5 * (1.0 - dv.t[d]);
      locs[have + i].t[d] = mid.t[d] + 1;
      w[have + i] *= 0.5 * (1.0 + dv.t[d]);
    }
    have *= 2;
  }
  for (int i = have; i < 8; ++i) w[i] = 0.0;
  double tw = 0;
  for (int i = 0; i < have; ++i) tw += w[i];
  for (int i = 0; i < have; ++i) w[i] += (1.0 - tw) * (1.0 / have);
  for (int i = 0; i < have; ++i) {
    if (w[i] < 0.0) {
      if (-w[i] >= SMALL * 1e5) fail("large negative interpolation weight");
      w[i] = 0.0;
    }
    else if (w[i] < SMALL)
      w[i] = 0.0;
  }
  // compact the non-zero weights to the front (src/vec.cpp:511-526)
  {
    IV *l = locs;
    double *ww = w;
    int n = have;
    while (n) {
      if (std::fabs(ww[0]) < 2e-15) {
        ww[0] = ww[n - 1];
        l[0] = l[n - 1];
        ww[n - 1] = 0.0;
        l[n - 1] = IV();
      }
      else {
        ++ww;
        ++l;
      }
      --n;
    }
  }
  bool same = true;
  for (int i = 0; i < 8 && w[i]; ++i)
    if (w[i] != w[0]) same = false;
  if (same) {
    int nw = 0;
    for (int i = 0; i < 8 && w[i]; ++i) ++nw;
    for (int i = 0; i < 8 && w[i]; ++i) w[i] = 1.0 / nw;
  }
  return true;
}

// ---- host structure -------------------------------------------------------------------
struct HostSus {
  double omega_0, gamma; // stored as realnum by the reference (src/meep.hpp:276)
  bool drude;
  int ft = 0; // the field type it polarises: 0 E_stuff, 1 H_stuff
  std::vector<double> sigma[NC]; // diagonal sigma[c][dir(c)], empty when trivial
};

struct HostStructure {
  Geom gv;
  bool single = false;
  double courant = 0.5, dt = 0.5;
  std::vector<double> chi1inv[NC]; // diagonal, empty when trivial
  // uPML profile in double; each side's profile written into one array (one chunk holds both)
  std::vector<double> sig[3], kap[3];
  bool pml_dir[3] = {false, false, false};
  std::vector<HostSus> sus;
  // conductivity[c][dir(c)] by D or B component c (src/meep.hpp:599), in realnum precision;
  // empty when trivial
  std::vector<double> cond[NC];
  bool any_cond() const {
    for (int c = 0; c < NC; ++c)
      if (!cond[c].empty()) return true;
    return false;
  }
  // condinv = 1 / (1 + conductivity dt / 2) in realnum (structure_chunk::update_condinv,
  // src/structure.cpp:693-707), reference layout; empty when c has no conductivity
  std::vector<double> condinv(int c) const {
    std::vector<double> v;
    if (cond[c].empty()) return v;
    v.resize(cond[c].size());
    for (size_t i = 0; i < v.size(); ++i) {
      const double q = 1 / (1 + cond[c][i] * dt * 0.5);
      v[i] = single ? double(float(q)) : q;
    }
    return v;
  }
};

// structure_chunk::use_pml (src/structure.cpp:625-691) with the quadratic profile of pml()
// a uPML term's mean stretch and profile (boundary_region, src/meep.hpp:665-672); the default
// profile is u^2 (pml_quadratic_profile, src/structure.cpp:252-255)
struct PmlShape {
  double mean_stretch = 1.0;
  mp_scalar_func profile = nullptr;
  void *user = nullptr;
  double integral = 1. / 3., integral_u = 1. / 4.;
  double at(double u) const {
    if (!profile) return u * u;
    const double r[3] = {u, 0.0, 0.0};
    return profile(r, user);
  }
};
static void host_use_pml(HostStructure &s, int d, double dx, int side, double Rasym, const PmlShape &sh) {
  if (dx <= 0.0) return;
  const Geom &gv = s.gv;
  const double a = gv.a;
  const double prefac = (-std::log(Rasym)) / (4 * dx * sh.integral);
  const double kprefac = (sh.mean_stretch - 1) / sh.integral_u;
  const double bloc = side == High ? gv.num[d] * gv.inva : 0.0;
  auto pml_x = [&](int i) {
    double here = i * 0.5 / a;
    return 0.5 / a * ((int)(dx * (2 * a) + 0.5) - (int)(std::fabs(bloc - here) * (2 * a) + 0.5));
  };
  bool found = false;
  for (int i = 0; i <= 2 * gv.num[d] + 1; ++i)
    if (pml_x(i) > 0) found = true;
  if (!found) return;
  if (!s.pml_dir[d]) {
    s.pml_dir[d] = true;
    s.sig[d].assign(2 * gv.num[d] + 2, 0.0);
    s.kap[d].assign(2 * gv.num[d] + 2, 1.0);
  }
  for (int i = 0; i <= 2 * gv.num[d] + 1; ++i) {
    double x = pml_x(i);
    if (x > 0) {
      const double sp = sh.at(x / dx);
      s.sig[d][i] = 0.5 * s.dt * prefac * sp;
      s.kap[d][i] = 1 + kprefac * sp * (x / dx);
    }
  }
}

// ---- sources ----------------------------------------------------------------------------
struct SrcTime {
  int kind = 0;
  double freq = 0, width = 0, peak_time = 0, cutoff = 0, fwidth = 0;
  double start_time = 0, end_time = 0, slowness = 3.0;
  bool is_integrated = true;

  static SrcTime gaussian(double f, double w, double st, double et) {
    // gaussian_src_time(f, w, start, end) — src/sources.cpp:85-96
    SrcTime s;
    s.freq = f;
    s.width = w;
    s.peak_time = 0.5 * (st + et);
    s.cutoff = (et - st) * 0.5;
    s.fwidth = std::sqrt(-2.0 * std::log(1e-7)) / (w * pi);
    while (std::exp(-s.cutoff * s.cutoff / (2 * w * w)) < 1e-100) s.cutoff *= 0.9;
    s.cutoff = float(s.cutoff);
    return s;
  }
  static SrcTime continuous(double f, double w, double st, double et, double slow) {
    SrcTime s;
    s.kind = 1;
    s.freq = f;
    s.width = w;
    s.start_time = float(st);
    s.end_time = float(et);
    s.slowness = slow;
    return s;
  }
  cdouble dipole(double time) const {
    if (kind == 0) { // src/sources.cpp:98-108
      double tt = time - peak_time;
      if (float(std::fabs(tt)) > cutoff) return 0.0;
      cdouble amp = 1.0 / cdouble(0, -2 * pi * freq);
      return std::exp(-tt * tt / (2 * width * width)) * std::polar(1.0, -2 * pi * freq * tt) * amp;
    }
    float rtime = float(time); // src/sources.cpp:128-145
    if (rtime < start_time || rtime > end_time) return 0.0;
    cdouble amp = 1.0 / (cdouble(0, -1.0) * (2 * pi) * freq);
    if (width == 0.0) return std::exp(cdouble(0, -1.0) * (2 * pi) * freq * time) * amp;
    double ts = (time - start_time) / width - slowness;
    double te = (end_time - time) / width - slowness;
    return std::exp(cdouble(0, -1.0) * (2 * pi) * freq * time) * amp * (1.0 + std::tanh(ts)) *
           (1.0 + std::tanh(te)) * 0.25;
  }
  // src_time::current (src/meep.hpp:989-991)
  cdouble current(double time, double dt) const { return (dipole(time + dt) - dipole(time)) / dt; }
  double last_time() const { return kind == 0 ? float(peak_time + cutoff) : end_time; }
  double get_fwidth() const { return kind == 0 ? fwidth : 0.0; }
  bool equal(const SrcTime &o) const {
    if (kind != o.kind || is_integrated != o.is_integrated) return false;
    if (kind == 0) return freq == o.freq && width == o.width && peak_time == o.peak_time && cutoff == o.cutoff;
    return freq == o.freq && width == o.width && start_time == o.start_time && end_time == o.end_time &&
           slowness == o.slowness;
  }
};

struct SrcVol { // src_vol (src/meep_internals.hpp:49-82)
  int c;        // E or H component
  int src;      // index into the source-time list
  std::vector<int64_t> index;
  std::vector<cdouble> amp;
  std::vector<int> jpos; // array indices X, Y, Z of each point (0 for absent directions)
  // device copies for lists applied by k_src, made on first use and refreshed when the
  // amplitudes change
  int64_t *d_idx = nullptr;
  double2 *d_amp = nullptr; // list order
  bool dirty = true;
  // folded periodic images of the list's points (point k, halo index, sign), for the plan
  // version img_version
  int n_img = 0;
  long img_version = -1;
  int *d_img_k = nullptr;
  int64_t *d_img_dst = nullptr;
  int8_t *d_img_sgn = nullptr;
};

// ---- device buffers -----------------------------------------------------------------------
// Set for the duration of mp_fields_step_n_noalloc: any device allocation or release on this
// thread then fails instead of reaching hipMalloc / hipFree.
thread_local bool g_no_alloc = false;
static void alloc_allowed(const char *what) {
  if (g_no_alloc)
    fail("mp_fields_step_n_noalloc: the step would allocate or free device memory (%s); step the fields "
         "with mp_fields_step_n first",
         what);
}
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void alloc(size_t n) {
    alloc_allowed("a state, plan or staging buffer");
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = n;
    if (n) HIPCHECK(hipMalloc(&p, n));
  }
};

static void ensure_gpu() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0)
    fail("no usable GPU: the MI355X engine has no CPU path (hipGetDeviceCount: %s)", hipGetErrorString(e));
}

// how a step computes one value of its source row: a table entry (kind 0) is the real (cmp 0)
// or imaginary part of (amp * s(t)) * mult, a box entry (kind 1) the part cmp of s(t), with
// s the current (or dipole, for integrated sources, toff 1) of source-time src at
// t * dt + toff * dt
struct SrcRecipe {
  int src;
  cdouble amp;
  double mult;
  int cmp;
  double toff;
  int kind;
};

struct FieldsBase {
  virtual ~FieldsBase() {}
  HostStructure *s = nullptr;
  virtual void use_real_fields() = 0;
  virtual void use_bloch(int d, cdouble k) = 0;
  virtual void set_boundary(int side, int d, int cond) = 0;
  virtual void add_volume_source(int c, const SrcTime &st, V3 wmin, V3 wmax, cdouble amp, mp_complex_func A, void *Au) = 0;
  virtual void add_srcdata(int c, const SrcTime &st, const long *ref_index, long n, const double *amp, bool fix) = 0;
  virtual void require_source_components() = 0;
  virtual int dimension() const = 0;
  virtual void step_n(long n, bool check_each) = 0;
  // step_n with the source rows staged in `rows` (device memory of rows_bytes) and no device
  // allocation or release
  virtual void step_n_noalloc(long n, double *rows, size_t rows_bytes) = 0;
  virtual int device_id() const = 0;
  virtual long t_now() const = 0;
  virtual double dt_now() const = 0;
  virtual double last_source_time() const = 0;
  virtual cdouble get_field(int c, const V3 &p) = 0;
  virtual double field_energy() = 0;
  virtual void adj_begin(void *monitor, int c, V3 mn, V3 mx, int interval, int *box_lo, int *box_n) = 0;
  virtual void adj_end() = 0;
  virtual double adj_objective() = 0;
  virtual void adj_gradient(double *out) = 0;
  virtual int get_array(int c, int cmp, int kind, int ipol, double *out) = 0;
  virtual void set_array(int c, int cmp, int kind, int ipol, const double *in) = 0;
  virtual bool have_component(int c) const = 0;
  virtual void times(double *t7) const = 0;
  virtual void set_dft_acc(int kind) = 0;
  virtual void set_ktiming(bool on) = 0;
  virtual void get_ktiming(int which, double *ms, long *n) = 0;
  virtual void reset_ktiming() = 0;
  virtual void *add_dft_flux(int d, V3 mn, V3 mx, const std::vector<double> &freq, bool centered, int dec, int *dec_out) = 0;
  virtual void *add_dft_fields(const int *comps, int ncomp, V3 mn, V3 mx, const std::vector<double> &freq, bool centered, int dec,
                               int *dec_out) = 0;
  virtual int flux_chunk_component(void *fl, int which, int i) = 0;
  virtual void flux(void *fl, double *out) = 0;
  virtual int flux_chunks(void *fl, int which) = 0;
  virtual long flux_points(void *fl, int which, int i) = 0;
  virtual void flux_dft(void *fl, int which, int i, double *out) = 0;
  // replace chunk i's accumulators with in (re, im pairs); multiply every accumulator of the
  // monitor (both lists) by s
  virtual void flux_set_dft(void *fl, int which, int i, const double *in) = 0;
  virtual void flux_scale(void *fl, cdouble s) = 0;
  // fields::get_dft_array over the flux (both lists) or dft_fields (E list) monitor fl;
  // returns the number of complex values
  virtual long dft_array(void *fl, bool both, int c, int num_freq, int *rank, size_t *dims, double *out) = 0;
  // near2far: the surfaces (normal direction or -1, box, weight), and the six fields of every
  // frequency at np far points (out: np * nfreq * 12 doubles, point-major)
  virtual void *add_dft_near2far(int nsurf, const int *dirs, const double *vmins, const double *vmaxs, const double *weights,
                                 const std::vector<double> &freq, int dec, int nperiods, int *dec_out) = 0;
  virtual void n2f_info(void *n, double *eps_mu, int *periodic, long *npoints) = 0;
  // dft_energy: the volumes, and the total (which 0), electric (1) or magnetic (2) spectrum
  virtual void *add_dft_energy(int nvol, const double *vmins, const double *vmaxs, const std::vector<double> &freq, int dec,
                               int *dec_out) = 0;
  virtual void energy_spectrum(void *e, int which, double *out) = 0;
  // dft_force: the volumes (force direction class per volume, complex weights) and the force
  virtual void *add_dft_force(int nvol, const int *fdirs, const double *vmins, const double *vmaxs, const double *weights,
                              const std::vector<double> &freq, int dec, int *dec_out) = 0;
  virtual void force_spectrum(void *fo, double *out) = 0;
  virtual void n2f_eval(void *n, const double *pts, long np, double *out) = 0;
  // time-domain flux: flux_vol planes, flux_in_box (sync) and flux_in_box_wrongH
  virtual void *add_flux_vol(int d, V3 mn, V3 mx) = 0;
  virtual double flux_vol_value(void *fv) = 0;
  virtual double flux_in_box(int d, V3 mn, V3 mx, bool sync) = 0;
  // field_energy_in_box (which 0), electric_ (1), magnetic_energy_in_box (2)
  virtual double energy_in_box(int which, V3 mn, V3 mx) = 0;
  virtual void flux_delete(void *fl) = 0;
};

template <class R> struct DftChunk {
  int c;
  int npts = 0, nfreq = 0;
  int decimation = 1;
  cdouble scale;
  int64_t avg1 = 0, avg2 = 0;
  int acc_kind = MP_DFT_REALNUM;
  DevBuf idx, w, omega, acc, comp;
  // deferred accumulation (DftDefer): up to `defer` steps of point values wait in snap
  int defer = 1, nsnap = 0;
  double snap_time[DFT_DEFER] = {};
  DevBuf snap;
  // where the points sit, for get_dft_array: the box in half-pixel coordinates, the periodic
  // image it stands for, its integration weights, and the monitor's empty directions
  IV is, ie, shift;
  V3 s0, s1, e0, e1;
  double dV0 = 1;
  bool weighted = false;
  cdouble stored_weight = 1.0;
  bool empty[3] = {};
  int vc = -1; // near2far: the equivalent current's component
  cdouble extra_weight = 1.0; // dft_force: the readout's factor (src/dft.cpp:97,224)
  bool sqrt_w = false;        // accumulated with the square root of the integration weight
};
template <class R> struct Flux {
  virtual ~Flux() = default;
  std::vector<DftChunk<R> *> E, H;
  int nfreq;
  V3 wmin, wmax; // the monitor volume
};
// dft_near2far (src/meep.hpp:1455-1520): the accumulators of the tangential fields (the E list)
// with the medium and the periodic copies, and the equivalent-current list the far-field
// kernel reads, kept on the device until the fields step again
// dft_force (src/stress.cpp:27-114): offdiag1 and offdiag2 as the base's E and H lists, diag
template <class R> struct Force : Flux<R> {
  std::vector<DftChunk<R> *> D;
};
// dft_energy (src/dft.cpp:654-750): the E and H lists of the base, with D and B beside them
template <class R> struct Energy : Flux<R> {
  std::vector<DftChunk<R> *> D, B;
};
template <class R> struct N2F : Flux<R> {
  std::vector<double> freq;
  double eps = 0, mu = 0;
  int periodic_d[2] = {-1, -1}, periodic_n[2] = {0, 0};
  double periodic_k[2] = {0, 0}, period[2] = {0, 0};
  long built_t = -1;
  unsigned built_version = 0, built_acc = 0;
  int nsrc = 0, cls_lo[7] = {};
  DevBuf sx, sy, sz, f0, d_freq;
};

template <class R> struct Fields : FieldsBase {
  Geom gv;
  double a, dt, courant;
  long t = 0;
  bool is_real = false;
  bool components_allocated = false;
  hipStream_t stream = nullptr;
  int device = 0; // the GPU current when the object was created (its stream's device)

  R *f[NC][2] = {};
  // the uPML and conductivity auxiliaries are FP64 arrays whatever realnum is
  double *f_u[NC][2] = {}, *f_w[NC][2] = {}, *f_cond[NC][2] = {};
  std::vector<std::unique_ptr<DevBuf>> bufs;
  // materials and uPML on the device
  R *d_chi1inv[NC] = {};
  // conductivity of the D and B components and its condinv (device layout), and condinv on
  // the host (reference layout) for the current sources; empty when absent
  R *d_cond[NC] = {}, *d_condinv[NC] = {};
  std::vector<double> h_condinv[NC];
  bool any_cond = false;
  R *d_sig[3] = {}, *d_kap[3] = {}, *d_siginv[3] = {};
  struct Pol {
    int isus;
    bool allocated = false;
    R *P[NC][2] = {}, *Pp[NC][2] = {};
    R *sigma[NC] = {};
    bool sig_uni[NC] = {}; // sigma of the component is sig_val over the whole pole box
    R sig_val[NC] = {};
    PoleCoef<R> coef;
  };
  std::vector<Pol> pols; // the reference's list order: newest susceptibility first

  int bc[2][3];
  cdouble k[3], eikna[3];
  std::vector<SrcTime> srcs;
  std::vector<SrcVol> sources[4];
  // curl plan (figure_out_step_plan, src/fields.cpp:428-456)
  int plus_c[NC], minus_c[NC], plus_d[NC], minus_d[NC];
  bool have_p[NC], have_m[NC];

  // halo tables, rebuilt when allocations or boundary conditions change
  bool plan_valid = false;
  // two merged halo lists: 0 = B, H and their polarisations (after the B phase),
  // 1 = D, E and P_E (after the D phase)
  DevBuf d_conn[2], d_phase[2];
  int n_conn[2] = {0, 0};
  DevBuf d_slots;
  std::vector<R *> h_slots;
  // finiteness probe at the cell centre (src/step.cpp:137-138)
  DevBuf d_chk_ptrs, d_chk_idx, d_chk_w, d_flag;
  int n_chk = 0;
  DevBuf d_scratch; // saved f_minus_p points for long integrated-source lists

  std::vector<std::unique_ptr<DftChunk<R>>> dft_chunks;
  std::vector<std::unique_ptr<Flux<R>>> fluxes;
  int dft_acc = MP_DFT_REALNUM;
  double timers[7] = {0, 0, 0, 0, 0, 0, 0};

  // vectorised kernel on the interior box unless MEEP_AMD_GENERIC_KERNEL is set (both kernels
  // are held to the oracle by the tests)
  bool use_fast = getenv("MEEP_AMD_GENERIC_KERNEL") == nullptr;
  // Whole rows: the launch boxes span whole rows, with the stride-1 direction's uPML taken
  // point by point inside the interior box (rows_pml), instead of thin end slabs whose short
  // row pieces coalesce poorly and cost two more launches per phase.  Measured on MI355X:
  // faster for a chunk stepped alone (384x384x192 with 16-cell uPML: 2.05 against 2.38 ms per
  // step; 600x400 2D with 50-cell uPML: 0.068 against 0.098 ms), slower inside a batch, whose
  // per-point-uPML group then carries most of the work (64 squares of 128-512 cells with
  // 10-cell uPML: 0.220 against 0.186 ms); a batch sets batch_split on its chunks.
  // MEEP_AMD_WHOLE_ROWS=0 or 1 overrides the choice.
  int whole_rows_env = getenv("MEEP_AMD_WHOLE_ROWS") ? atoi(getenv("MEEP_AMD_WHOLE_ROWS")) : -1;
  bool batch_split = false;
  bool rows_pml = false;
  // kernel timing: event pairs recorded around launches, resolved after each synchronisation
  bool ktiming = false;
  std::vector<hipEvent_t> ev_pool;
  struct Pending {
    int kind;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  static constexpr int NKIND = 8; // the MP_K_* classes
  double k_ms[NKIND] = {0, 0, 0, 0, 0, 0, 0, 0};
  long k_n[NKIND] = {0, 0, 0, 0, 0, 0, 0, 0};
  int open_kind = -1;
  hipEvent_t open_ev = nullptr;
  hipEvent_t get_event() {
    if (!ev_pool.empty()) {
      hipEvent_t e = ev_pool.back();
      ev_pool.pop_back();
      return e;
    }
    hipEvent_t e;
    alloc_allowed("a timing event");
    HIPCHECK(hipEventCreate(&e));
    return e;
  }
  void kbegin(int kind) {
    if (!ktiming || collector) return;
    open_kind = kind;
    open_ev = get_event();
    HIPCHECK(hipEventRecord(open_ev, stream));
  }
  void kend() {
    if (!ktiming || open_kind < 0 || collector) return;
    hipEvent_t b = get_event();
    HIPCHECK(hipEventRecord(b, stream));
    pending.push_back({open_kind, open_ev, b});
    open_kind = -1;
  }
  void resolve_events() {
    for (const Pending &p : pending) {
      float ms = 0;
      HIPCHECK(hipEventElapsedTime(&ms, p.a, p.b));
      k_ms[p.kind] += ms;
      k_n[p.kind] += 1;
      ev_pool.push_back(p.a);
      ev_pool.push_back(p.b);
    }
    pending.clear();
  }
  void set_ktiming(bool on) override { ktiming = on; }
  void get_ktiming(int which, double *ms, long *n) override {
    if (which < 0 || which >= NKIND) fail("unknown kernel class %d", which);
    HIPCHECK(hipStreamSynchronize(stream));
    resolve_events();
    *ms = k_ms[which];
    *n = k_n[which];
  }
  void reset_ktiming() override {
    HIPCHECK(hipStreamSynchronize(stream));
    resolve_events();
    for (int q = 0; q < NKIND; ++q) k_ms[q] = 0, k_n[q] = 0;
  }

  R *dalloc(const R *copy_from) {
    bufs.emplace_back(new DevBuf);
    bufs.back()->alloc(sizeof(R) * gv.ntot);
    R *p = static_cast<R *>(bufs.back()->p);
    if (copy_from)
      HIPCHECK(hipMemcpyAsync(p, copy_from, sizeof(R) * gv.ntot, hipMemcpyDeviceToDevice, stream));
    else
      HIPCHECK(hipMemsetAsync(p, 0, sizeof(R) * gv.ntot, stream));
    return p;
  }
  // an FP64 auxiliary array: a copy of the realnum array copy_from (f_u and f_w start as
  // copies of f, src/step_db.cpp:61-65, src/update_eh.cpp:174-178), or zero
  double *daux(const R *copy_from) {
    bufs.emplace_back(new DevBuf);
    bufs.back()->alloc(sizeof(double) * gv.ntot);
    double *p = static_cast<double *>(bufs.back()->p);
    if (copy_from)
      launch_widen<R>(p, copy_from, gv.ntot, stream);
    else
      HIPCHECK(hipMemsetAsync(p, 0, sizeof(double) * gv.ntot, stream));
    return p;
  }
  template <class T> T *upload(const std::vector<T> &v) {
    bufs.emplace_back(new DevBuf);
    bufs.back()->alloc(sizeof(T) * v.size());
    HIPCHECK(hipMemcpy(bufs.back()->p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return static_cast<T *>(bufs.back()->p);
  }
  int ncmp() const { return is_real ? 1 : 2; }
  // a reference-layout host array in the device layout (padding filled with pad)
  std::vector<R> to_device(const std::vector<double> &ref, R pad) const {
    std::vector<R> v(gv.ntot, pad);
    gv.for_each_point([&](int64_t r, int64_t d) { v[d] = R(ref[r]); });
    return v;
  }

  explicit Fields(HostStructure *st) {
    ensure_gpu();
    s = st;
    gv = st->gv;
    gv.pad_rows(4); // 16-byte aligned rows for both FP32 (4 points) and FP64 (2 points)
    a = gv.a;
    dt = st->dt;
    courant = st->courant;
    if (gv.ntot >= (size_t(1) << 32)) fail("chunk too large for 32-bit cell indexing (%zu points)", gv.ntot);
    HIPCHECK(hipGetDevice(&device));
    HIPCHECK(hipStreamCreate(&stream));
    for (int d = 0; d < 3; ++d) {
      k[d] = 0.0;
      eikna[d] = 1.0;
      for (int b = 0; b < 2; ++b) bc[b][d] = gv.has(d) ? Metallic : None;
    }
    for (int c = 0; c < NC; ++c) have_p[c] = have_m[c] = false;
    // materials
    for (int c = 0; c < 5; ++c)
      if (!s->chi1inv[c].empty()) d_chi1inv[c] = upload(to_device(s->chi1inv[c], R(1)));
    for (int c = 10; c < NC; ++c)
      if (!s->cond[c].empty()) {
        h_condinv[c] = s->condinv(c);
        d_cond[c] = upload(to_device(s->cond[c], R(0)));
        d_condinv[c] = upload(to_device(h_condinv[c], R(1)));
        any_cond = true;
      }
    for (int d = 0; d < 3; ++d)
      if (s->pml_dir[d]) {
        size_t n = s->sig[d].size();
        std::vector<R> sg(n), kp(n), si(n);
        for (size_t i = 0; i < n; ++i) {
          sg[i] = R(s->sig[d][i]);
          kp[i] = R(s->kap[d][i]);
          si[i] = R(1) / (kp[i] + sg[i]);
        }
        d_sig[d] = upload(sg);
        d_kap[d] = upload(kp);
        d_siginv[d] = upload(si);
      }
    if (int(s->sus.size()) > MAXPOLE) fail("at most %d susceptibilities are supported", MAXPOLE);
    for (int i = int(s->sus.size()) - 1; i >= 0; --i) {
      Pol p;
      p.isus = i;
      const HostSus &su = s->sus[i];
      // lorentzian_susceptibility::update_P coefficients in realnum (src/susceptibility.cpp:192-196)
      const R om = R(su.omega_0), ga = R(su.gamma), rdt = R(dt);
      const R omega2pi = R(2 * pi * om), g2pi = R(ga * 2 * pi);
      p.coef.omega0dtsqr = omega2pi * omega2pi * rdt * rdt;
      p.coef.gamma1inv = R(1) / (R(1) + g2pi * rdt / R(2));
      p.coef.gamma1 = R(1) - g2pi * rdt / R(2);
      p.coef.omega0dtsqr_denom = su.drude ? R(0) : p.coef.omega0dtsqr;
      // where sigma is non-zero, by loop axis
      int blo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, bhi[3] = {-1, -1, -1};
      for (int c = 0; c < 10; ++c) {
        if (su.sigma[c].empty()) continue;
        int j[3], hi[3];
        for (int d = 0; d < 3; ++d) hi[d] = gv.has(d) ? gv.num[d] : 0;
        for (j[0] = 0; j[0] <= hi[0]; ++j[0])
          for (j[1] = 0; j[1] <= hi[1]; ++j[1])
            for (j[2] = 0; j[2] <= hi[2]; ++j[2])
              if (su.sigma[c][gv.ref_of(j)] != 0.0)
                for (int d = 0; d < 3; ++d) {
                  blo[d] = std::min(blo[d], j[d]);
                  bhi[d] = std::max(bhi[d], j[d]);
                }
      }
      // components whose sigma (in realnum) takes one value over the whole box
      if (bhi[0] >= 0 || bhi[1] >= 0 || bhi[2] >= 0)
        for (int c = 0; c < 10; ++c) {
          if (su.sigma[c].empty()) continue;
          int j[3], lo[3], hi[3];
          for (int d = 0; d < 3; ++d) {
            lo[d] = gv.has(d) ? blo[d] : 0;
            hi[d] = gv.has(d) ? bhi[d] : 0;
          }
          const R v0 = R(su.sigma[c][gv.ref_of(lo)]);
          bool uni = true;
          for (j[0] = lo[0]; uni && j[0] <= hi[0]; ++j[0])
            for (j[1] = lo[1]; uni && j[1] <= hi[1]; ++j[1])
              for (j[2] = lo[2]; uni && j[2] <= hi[2]; ++j[2])
                if (R(su.sigma[c][gv.ref_of(j)]) != v0) uni = false;
          p.sig_uni[c] = uni;
          p.sig_val[c] = v0;
        }
      for (int k = 0; k < 3; ++k) {
        const int d = gv.ydir(k);
        const bool has = gv.has(d);
        p.coef.lo[k] = has ? blo[d] : 0;
        p.coef.hi[k] = has ? bhi[d] : 0;
        if (bhi[d] < 0) p.coef.lo[k] = 1, p.coef.hi[k] = 0; // nowhere
      }
      pols.push_back(p);
    }
    // sigma of every pole for one component in one block of consecutive arrays (the row
    // kernel addresses a component's poles from one base pointer)
    for (int c = 0; c < 10; ++c) {
      std::vector<size_t> with;
      for (size_t q = 0; q < pols.size(); ++q)
        if (!s->sus[pols[q].isus].sigma[c].empty()) with.push_back(q);
      if (with.empty()) continue;
      std::vector<R> host(gv.ntot * with.size());
      for (size_t j = 0; j < with.size(); ++j) {
        const std::vector<R> v = to_device(s->sus[pols[with[j]].isus].sigma[c], R(0));
        std::copy(v.begin(), v.end(), host.begin() + j * gv.ntot);
      }
      R *base = upload(host);
      for (size_t j = 0; j < with.size(); ++j) pols[with[j]].sigma[c] = base + j * gv.ntot;
    }
    d_flag.alloc(sizeof(int));
    HIPCHECK(hipMemset(d_flag.p, 0, sizeof(int)));
    // unit directions are periodic (src/fields.cpp:80-85)
    for (int d = 0; d < 3; ++d)
      if (gv.has(d) && gv.num[d] == 1) use_bloch(d, 0.0);
  }
  ~Fields() override {
    if (stream) {
      (void)hipStreamSynchronize(stream);
      for (int ft = 0; ft < 4; ++ft)
        for (auto &sv : sources[ft]) {
          if (sv.d_idx) (void)hipFree(sv.d_idx);
          if (sv.d_amp) (void)hipFree(sv.d_amp);
          if (sv.d_img_k) (void)hipFree(sv.d_img_k);
          if (sv.d_img_dst) (void)hipFree(sv.d_img_dst);
          if (sv.d_img_sgn) (void)hipFree(sv.d_img_sgn);
        }
      for (const Pending &p : pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
      }
      for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
      (void)hipStreamDestroy(stream);
    }
  }

  bool has_field(int c) const {
    if (gv.dim == 1) return c == 0 || c == 6 || c == 10 || c == 16;
    return comp_dir(c) != NODIR;
  }
  static bool is_tm(int c) { return c == 5 || c == 6 || c == 15 || c == 16 || c == 4 || c == 14; }

  void figure_out_step_plan() {
    auto cross = [](int a, int b) { return (3 + 2 * a - b) % 3; };
    auto cross_negative = [](int a, int b) { return ((3 + b - a) % 3) == 2; };
    for (int c = 0; c < NC; ++c) have_p[c] = have_m[c] = false;
    for (int c1 = 0; c1 < NC; ++c1) {
      if (!f[c1][0]) continue;
      int dc1 = comp_dir(c1);
      for (int c2 = 0; c2 < NC; ++c2) {
        bool pair = ((is_electric(c1) || is_D(c1)) && is_magnetic(c2)) || ((is_magnetic(c1) || is_B(c1)) && is_electric(c2));
        if (!pair) continue;
        int dc2 = comp_dir(c2);
        if (dc2 == NODIR || dc1 == dc2 || !has_field(c1) || !has_field(c2)) continue;
        int dd = cross(dc1, dc2);
        if (!gv.has(dd)) continue;
        if (cross_negative(dc2, dc1)) {
          minus_c[c1] = c2;
          minus_d[c1] = dd;
          have_m[c1] = true;
        }
        else {
          plus_c[c1] = c2;
          plus_d[c1] = dd;
          have_p[c1] = true;
        }
      }
    }
  }
  // require_component / alloc_f (src/fields.cpp:480-504,548-574)
  void require_component(int c) {
    if (!has_field(c)) fail("cannot require a component absent from this grid");
    components_allocated = true;
    bool changed = false;
    for (int ca = 0; ca < NC; ++ca) {
      if (!has_field(ca)) continue;
      if (gv.dim == 2 && (is_tm(c) ^ is_tm(ca))) continue;
      for (int cmp = 0; cmp < ncmp(); ++cmp)
        if (!f[ca][cmp]) {
          changed = true;
          if (is_magnetic(ca)) {
            int bcmp = dir_comp(15, comp_dir(ca));
            if (!f[bcmp][cmp]) f[bcmp][cmp] = dalloc(nullptr);
            f[ca][cmp] = f[bcmp][cmp];
          }
          else
            f[ca][cmp] = dalloc(nullptr);
        }
    }
    if (changed) {
      figure_out_step_plan();
      plan_valid = false;
    }
  }

  void use_real_fields() override {
    for (int d = 0; d < 3; ++d)
      if (bc[High][d] == Periodic && k[d] != 0.0) fail("Can't use real fields with bloch boundary conditions!");
    is_real = true;
    for (int c = 0; c < NC; ++c) {
      f[c][1] = nullptr;
      f_u[c][1] = f_w[c][1] = nullptr;
    }
    plan_valid = false;
  }
  void set_boundary(int b, int d, int cond) override {
    if (bc[b][d] != cond) {
      bc[b][d] = cond;
      if (cond == Periodic || bc[1 - b][d] == Periodic) bc[1 - b][d] = cond;
      plan_valid = false;
    }
  }
  void use_bloch(int d, cdouble kk) override {
    k[d] = kk;
    for (int b = 0; b < 2; ++b) set_boundary(b, d, Periodic);
    if (std::real(kk) * gv.num[d] == 0.5 * a)
      eikna[d] = -std::exp(-std::imag(kk) * ((2 * pi / a) * gv.num[d]));
    else
      eikna[d] = std::exp(cdouble(0, 1) * kk * ((2 * pi / a) * gv.num[d]));
    if (is_real && kk != 0.0) fail("Can't use real fields with bloch boundary conditions!");
    plan_valid = false;
  }

  // ---- halo plan (find_metals + connect_the_chunks for one chunk) ----
  bool on_metal_boundary(const IV &here) const {
    for (int d = 0; d < 3; ++d) {
      if (!gv.has(d)) continue;
      if (here.t[d] == 2 * gv.num[d] && bc[High][d] == Metallic) return true;
      if (bc[Low][d] == Magnetic && here.t[d] == 1) return true;
      if (bc[Low][d] == Metallic && here.t[d] == 0) return true;
    }
    return false;
  }
  bool locate(IV *there, cdouble *ph) const {
    if (!gv.owns(*there))
      for (int d = 0; d < 3; ++d) {
        if (!gv.has(d) || bc[High][d] != Periodic) continue;
        const int L = 2 * gv.num[d];
        if (there->t[d] <= 0)
          while (there->t[d] <= 0) {
            there->t[d] += L;
            *ph *= std::conj(eikna[d]);
          }
        else
          while (there->t[d] - L > 0) {
            there->t[d] -= L;
            *ph *= eikna[d];
          }
      }
    return gv.owns(*there);
  }
  // Periodic wrap along the stride-1 direction is folded into the update kernels when its
  // Bloch phase is real (+1 or -1); those halo points then leave the connection lists.
  // per direction: +1 / -1 when its periodic halo is folded into the kernels (real Bloch
  // phase on both sides), 0 when the halo lists carry it
  int wrap_dir[3] = {0, 0, 0};
  bool p_conn = false; // some halo connection moves P (its pointers then stay in place)
  // slots: 0..19 the field arrays, 20 + 20*p + c the polarisation P of pole p
  void rebuild_plan() {
    HIPCHECK(hipStreamSynchronize(stream));
    ++plan_version;
    const int nslot = NC + NC * int(pols.size());
    h_slots.assign(2 * nslot, nullptr);
    for (int c = 0; c < NC; ++c)
      for (int cmp = 0; cmp < 2; ++cmp) h_slots[2 * c + cmp] = f[c][cmp];
    for (size_t p = 0; p < pols.size(); ++p)
      for (int c = 0; c < NC; ++c)
        for (int cmp = 0; cmp < 2; ++cmp) h_slots[2 * (NC + NC * p + c) + cmp] = pols[p].P[c][cmp];
    d_slots.alloc(sizeof(R *) * h_slots.size());
    HIPCHECK(hipMemcpy(d_slots.p, h_slots.data(), sizeof(R *) * h_slots.size(), hipMemcpyHostToDevice));

    std::vector<Conn> lists[2];
    std::vector<double2> phs[2];
    bool B_redundant[5];
    for (int q = 0; q < 5; ++q) B_redundant[q] = f[5 + q][0] == f[15 + q][0];
    auto phase_class = [](cdouble ph) {
      auto close = [](cdouble p, double r) { return std::fabs(p.imag()) < 1e-13 && std::fabs(p.real() - r) < 1e-13; };
      return close(ph, 1.0) ? 2 : close(ph, -1.0) ? 1 : 0;
    };
    for (int d = 0; d < 3; ++d) {
      const cdouble e = eikna[d];
      wrap_dir[d] = 0;
      if (gv.has(d) && bc[High][d] == Periodic && bc[Low][d] == Periodic && std::fabs(e.imag()) < 1e-13) {
        if (std::fabs(e.real() - 1.0) < 1e-13) wrap_dir[d] = 1;
        if (std::fabs(e.real() + 1.0) < 1e-13) wrap_dir[d] = -1;
      }
    }
    // halo planes of directions the lists carry come first, so the folded planes' boxes hold
    // only points that are halo points along folded directions alone
    int order[3], no = 0;
    for (int d = 0; d < 3; ++d)
      if (!wrap_dir[d]) order[no++] = d;
    for (int d = 0; d < 3; ++d)
      if (wrap_dir[d]) order[no++] = d;
    for (int corig = 0; corig < NC; ++corig) {
      if (!f[corig][0]) continue;
      if (is_B(corig) && B_redundant[corig - 15]) continue;
      // the non-owned planes: index 0 (no Yee shift) or n (Yee shift) along each direction
      // (get_boundary_icorners, src/vec.cpp:395-430), as disjoint boxes
      int lo[3], hi[3];
      for (int d = 0; d < 3; ++d) {
        lo[d] = 0;
        hi[d] = gv.has(d) ? gv.num[d] : 0;
      }
      for (int od = 0; od < 3; ++od) {
        const int d = order[od];
        if (!gv.has(d)) continue;
        const int plane = gv.ysh(corig, d) ? gv.num[d] : 0;
        int blo[3], bhi[3];
        for (int q = 0; q < 3; ++q) {
          blo[q] = lo[q];
          bhi[q] = hi[q];
        }
        blo[d] = bhi[d] = plane;
        // folded planes are written by the kernels themselves
        const bool folded = wrap_dir[d] != 0;
        for (int j0 = blo[0]; j0 <= (folded ? blo[0] - 1 : bhi[0]); ++j0)
          for (int j1 = blo[1]; j1 <= bhi[1]; ++j1)
            for (int j2 = blo[2]; j2 <= bhi[2]; ++j2) {
              int j[3] = {j0, j1, j2};
              IV here = gv.iloc(corig, j);
              const int64_t n = gv.index(corig, here);
              cdouble ph = 1.0;
              if (!locate(&here, &ph) || on_metal_boundary(here)) continue;
              Conn cn;
              cn.dst = uint32_t(n);
              cn.src = uint32_t(gv.index(corig, here));
              cn.slot = uint16_t(corig);
              cn.cls = uint16_t(phase_class(ph));
              const double2 phv = make_double2(ph.real(), ph.imag());
              const int ft = field_type(corig);
              const int grp = (ft == 1 || ft == 3) ? 0 : 1;
              lists[grp].push_back(cn);
              phs[grp].push_back(phv);
              if (is_electric(corig) || is_magnetic(corig))
                for (size_t p = 0; p < pols.size(); ++p)
                  if (pols[p].allocated && pols[p].P[corig][0]) {
                    Conn cp = cn;
                    cp.slot = uint16_t(NC + NC * int(p) + corig);
                    lists[grp].push_back(cp);
                    phs[grp].push_back(phv);
                  }
            }
        // later boxes exclude this plane (the boxes are disjoint)
        if (plane == 0)
          lo[d] = 1;
        else
          hi[d] = plane - 1;
      }
    }
    p_conn = false; // halo lists that hold P slot pointers pin the P arrays
    for (int q = 0; q < 2; ++q)
      for (const Conn &cn : lists[q])
        if (cn.slot >= NC) p_conn = true;
    for (int q = 0; q < 2; ++q)
      for (const Conn &cn : lists[q]) {
        if (size_t(cn.dst) >= gv.ntot || size_t(cn.src) >= gv.ntot || cn.slot >= nslot)
          fail("internal: halo connection outside the array");
        if (cn.cls == 0 && (!h_slots[2 * cn.slot + 1] || ncmp() < 2)) fail("internal: Bloch phase on real fields");
        if (!h_slots[2 * cn.slot]) fail("internal: halo connection to an unallocated array");
      }
    for (int q = 0; q < 2; ++q) {
      n_conn[q] = int(lists[q].size());
      d_conn[q].alloc(sizeof(Conn) * lists[q].size());
      bool any_phase = false;
      for (const Conn &cn : lists[q]) any_phase = any_phase || cn.cls == 0;
      d_phase[q].alloc(any_phase ? sizeof(double2) * phs[q].size() : 0);
      if (n_conn[q]) HIPCHECK(hipMemcpy(d_conn[q].p, lists[q].data(), sizeof(Conn) * lists[q].size(), hipMemcpyHostToDevice));
      if (any_phase) HIPCHECK(hipMemcpy(d_phase[q].p, phs[q].data(), sizeof(double2) * phs[q].size(), hipMemcpyHostToDevice));
    }
    // centre probe: E and D components at the interpolation points of the centre
    std::vector<const R *> cp;
    std::vector<int64_t> ci;
    std::vector<double> cw;
    V3 cen = gv.center();
    for (int c = 0; c < 5; ++c) {
      if (!has_field(c) || !f[c][0]) continue;
      for (int which = 0; which < 2; ++which) {
        int cc = which ? c + 10 : c;
        IV locs[8];
        double w[8];
        gv.interpolate(cc, cen, locs, w);
        for (int q = 0; q < 8 && w[q]; ++q) {
          IV il = locs[q];
          cdouble ph = 1.0;
          locate(&il, &ph);
          if (!gv.owns(il)) continue;
          for (int cmp = 0; cmp < ncmp(); ++cmp)
            if (f[cc][cmp]) {
              cp.push_back(f[cc][cmp]);
              ci.push_back(gv.index(cc, il));
              cw.push_back(w[q]);
            }
        }
      }
    }
    n_chk = int(cp.size());
    d_chk_ptrs.alloc(sizeof(R *) * cp.size());
    d_chk_idx.alloc(sizeof(int64_t) * ci.size());
    d_chk_w.alloc(sizeof(double) * cw.size());
    if (n_chk) {
      HIPCHECK(hipMemcpy(d_chk_ptrs.p, cp.data(), sizeof(R *) * cp.size(), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(d_chk_idx.p, ci.data(), sizeof(int64_t) * ci.size(), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(d_chk_w.p, cw.data(), sizeof(double) * cw.size(), hipMemcpyHostToDevice));
    }
    plan_valid = true;
  }
  // step_boundaries for a merged group of field types (see d_conn)
  void boundaries(int grp) {
    if (!n_conn[grp]) return;
    if (collector) {
      collector->conn[grp].push_back({static_cast<const Conn *>(d_conn[grp].p), static_cast<const double2 *>(d_phase[grp].p),
                                      n_conn[grp], static_cast<void *const *>(d_slots.p), ncmp()});
      return;
    }
    kbegin(MP_K_HALO);
    launch_conn<R>(static_cast<const Conn *>(d_conn[grp].p), static_cast<const double2 *>(d_phase[grp].p), n_conn[grp],
                   static_cast<R *const *>(d_slots.p), ncmp(), stream);
    kend();
  }
  int cur_cmp = 0; // real (0) or imaginary (1) part being stepped
  // Every pointer a step kernel can dereference for the slots it will visit is checked on
  // the host first, so a planning mistake raises an error instead of faulting the GPU.
  void check_args(const StepArgs<R> &a, bool curl, bool upd, bool fast) const {
    for (int c = 0; c < a.nc; ++c) {
      if (curl && a.cc[c].f) {
        const CurlComp<R> &k = a.cc[c];
        if (k.g2 && !k.g1) fail("internal: curl slot %d has g2 without g1", c);
        if (k.dsigu >= 0 && !k.fu) fail("internal: curl slot %d lacks its PML auxiliary", c);
        if ((k.dsig >= 0 || k.dsigu >= 0) && (!a.pml.sig[k.dsig >= 0 ? k.dsig : k.dsigu])) fail("internal: missing PML profile");
        if (k.cnd && (!k.cndinv || (k.dsig >= 0 && !k.fcnd))) fail("internal: curl slot %d lacks its conductivity arrays", c);
        if (k.cnd && fast) fail("internal: conductive slot %d in the vectorised kernel", c);
        if (fast) {
          if (a.hasp[c] && !a.F[(c + 2) % 3]) fail("internal: fast slot %d lacks its plus partner", c);
          if (a.hasm[c] && !a.F[(c + 1) % 3]) fail("internal: fast slot %d lacks its minus partner", c);
        }
      }
      if (upd && a.uc[c].f) {
        const UpdComp<R> &u = a.uc[c];
        if (!u.g) fail("internal: update slot %d has no source array", c);
        if (u.dsigw >= 0 && (!u.fw || !a.pml.sig[u.dsigw])) fail("internal: update slot %d lacks its PML auxiliary", c);
        for (int p = 0; p < MAXPOLE; ++p)
          if (p < a.npole && u.P[p] && !u.Pp[p]) fail("internal: polarisation without its previous value");
      }
    }
    // sources: one of this chunk's device source sets (validated when built), with its rows
    if (a.src) {
      const SrcSet *base = splan.dev;
      if (!base || a.src < base || a.src >= base + 4) fail("internal: bad source set");
      for (int k = 0; k < 3; ++k)
        if (a.sbox_lo[k] > a.sbox_hi[k]) fail("internal: empty source box");
      if (!collector && !a.vrow) fail("internal: source set without its value row");
    }
    // loop-axis view: launch box inside the grid, slots' ranges inside the arrays
    const Grid &g = a.g;
    for (int k = 0; k < 3; ++k) {
      const int d = gv.ydir(k);
      const int n = gv.has(d) ? gv.num[d] : 0;
      if (g.alo[k] < 0 || g.an[k] < 1 || g.alo[k] + g.an[k] - 1 > n) fail("internal: launch box outside the grid");
      if (g.ast[k] != (gv.has(d) ? gv.st[d] : 0)) fail("internal: launch strides do not match the layout");
      if (g.wrap[k] < -1 || g.wrap[k] > 1 || (g.wrap[k] && (!gv.has(d) || g.nax[k] != n))) fail("internal: bad folded wrap");
      for (int c = 0; c < a.nc; ++c) {
        const AxisSlot &x = curl ? a.cc[c].ax : a.uc[c].ax;
        if ((curl ? a.cc[c].f : a.uc[c].f) && (x.lo[k] < 0 || x.hi[k] > n)) fail("internal: slot %d range outside the grid", c);
      }
    }
    for (int c = 0; c < a.nc; ++c) {
      // the vectorised kernel derives owned ranges and metal planes from the Yee-shift bits
      const AxisSlot &x = curl ? a.cc[c].ax : a.uc[c].ax;
      if (curl ? a.cc[c].f : a.uc[c].f)
        for (int k = 0; k < 3; ++k) {
          const int n = g.nax[k], y = (x.ybits >> k) & 1;
          const int lo = n ? 1 - y : 0, hi = n ? n - y : 0;
          const int mhi = ((g.mhi_bits >> k) & 1) && y == 0 && n ? n : -1;
          const int mlo = ((g.mlo_bits >> k) & 1) && y == 1 && n ? 0 : -1;
          if (lo != x.lo[k] || hi != x.hi[k] || mhi != x.mhi[k] || mlo != x.mlo[k]) fail("internal: slot %d axis %d bits", c, k);
        }
      if (curl && a.cc[c].f) {
        const CurlComp<R> &k = a.cc[c];
        if ((k.ksig >= 0) != (k.sig != nullptr) || (k.ksigu >= 0) != (k.sigu != nullptr)) fail("internal: slot %d PML axes", c);
      }
      if (upd && a.uc[c].f && (a.uc[c].ksigw >= 0) != (a.uc[c].sigw != nullptr)) fail("internal: slot %d PML W axis", c);
    }
  }
  // Loop-axis view of an argument block (the only form the kernels read): the launch box and
  // each slot's owned range, metal planes and Yee shift by loop axis, and the loop axis, Yee
  // shift and profiles of each slot's uPML directions.
  void axis_slot(AxisSlot &ax, const int *lo, const int *hi, const int *ysh) const {
    for (int k = 0; k < 3; ++k) {
      const int d = gv.ydir(k);
      ax.mhi[k] = ax.mlo[k] = -1;
      if (!gv.has(d)) {
        ax.lo[k] = ax.hi[k] = 0;
        continue;
      }
      ax.lo[k] = lo[d];
      ax.hi[k] = hi[d];
      if (bc[High][d] == Metallic && ysh[d] == 0) ax.mhi[k] = gv.num[d];
      if (bc[Low][d] == Magnetic && ysh[d] == 1) ax.mlo[k] = 0;
    }
    ax.ybits = 0;
    for (int k = 0; k < 3; ++k) {
      ax.ysh[k] = gv.has(gv.ydir(k)) ? ysh[gv.ydir(k)] : 0;
      ax.ybits |= ax.ysh[k] << k;
    }
  }
  int axis_of(int d) const {
    if (d < 0) return -1;
    for (int k = 0; k < 3; ++k)
      if (gv.ydir(k) == d && gv.has(d)) return k;
    fail("internal: direction %d has no loop axis", d);
  }
  void finalize_grid(Grid &g) const {
    for (int k = 0; k < 3; ++k) {
      const int d = g.axdir[k];
      g.alo[k] = d >= 0 ? g.blo[d] : 0;
      g.an[k] = d >= 0 ? g.bhi[d] - g.blo[d] + 1 : 1;
      g.ast[k] = d >= 0 ? g.st[d] : 0;
    }
    g.mhi_bits = g.mlo_bits = 0;
    for (int k = 0; k < 3; ++k) {
      const int d = g.axdir[k];
      g.nax[k] = d >= 0 ? gv.num[d] : 0;
      g.wrap[k] = d >= 0 ? wrap_dir[d] : 0;
      if (d >= 0 && ((g.hi_metal >> d) & 1)) g.mhi_bits |= 1 << k;
      if (d >= 0 && ((g.lo_magnetic >> d) & 1)) g.mlo_bits |= 1 << k;
      g.qlo[k] = d >= 0 && boxes_valid ? quiet_lo[d] : 1;
      g.qhi[k] = d >= 0 && boxes_valid ? quiet_hi[d] : 0;
    }
  }
  void finalize_args(StepArgs<R> &a) const {
    finalize_grid(a.g);
    for (int c = 0; c < 3; ++c) {
      CurlComp<R> &k = a.cc[c];
      axis_slot(k.ax, k.lo, k.hi, k.ysh);
      // a slot the launch does not move may hold zeroed indices (direction 0, absent in 1D)
      k.ksig = k.f ? axis_of(k.dsig) : -1;
      k.ksigu = k.f ? axis_of(k.dsigu) : -1;
      k.ysig = k.dsig >= 0 ? k.ysh[k.dsig] : 0;
      k.ysigu = k.dsigu >= 0 ? k.ysh[k.dsigu] : 0;
      k.sig = k.dsig >= 0 ? a.pml.sig[k.dsig] : nullptr;
      k.kap = k.dsig >= 0 ? a.pml.kap[k.dsig] : nullptr;
      k.siginv = k.dsig >= 0 ? a.pml.siginv[k.dsig] : nullptr;
      k.sigu = k.dsigu >= 0 ? a.pml.sig[k.dsigu] : nullptr;
      k.kapu = k.dsigu >= 0 ? a.pml.kap[k.dsigu] : nullptr;
      k.siginvu = k.dsigu >= 0 ? a.pml.siginv[k.dsigu] : nullptr;
      UpdComp<R> &u = a.uc[c];
      axis_slot(u.ax, u.lo, u.hi, u.ysh);
      u.ksigw = u.f ? axis_of(u.dsigw) : -1;
      u.ysigw = u.dsigw >= 0 ? u.ysh[u.dsigw] : 0;
      u.sigw = u.dsigw >= 0 ? a.pml.sig[u.dsigw] : nullptr;
      u.kapw = u.dsigw >= 0 ? a.pml.kap[u.dsigw] : nullptr;
    }
  }
  void launch_box(StepArgs<R> &b, const int *lo, const int *hi, bool curl, bool upd, bool fast, bool pml) {
    for (int d = 0; d < 3; ++d) {
      b.g.blo[d] = lo[d];
      b.g.bhi[d] = hi[d];
    }
    finalize_args(b);
    check_args(b, curl, upd, fast);
    if (collector) {
      BatchItem<R> it;
      it.a = b;
      it.a.rev = 0;
      if (fast) {
        constexpr int V = 16 / sizeof(R);
        it.g0 = b.g.alo[2] / V;
        it.ngroups = (b.g.alo[2] + b.g.an[2] - 1) / V - it.g0 + 1;
        it.count = uint32_t(it.ngroups) * uint32_t(b.g.an[0]) * uint32_t(b.g.an[1]);
      }
      else {
        it.g0 = it.ngroups = 0;
        it.count = uint32_t(b.g.an[0]) * uint32_t(b.g.an[1]) * uint32_t(b.g.an[2]);
      }
      const int kind = fast ? (pml ? 2 : 1) + (upd && b.npole > 0 ? 2 : 0) : 0;
      collector->items[cur_phase][kind * 4 + (curl ? 2 : 0) + (upd ? 1 : 0)].push_back(it);
      return;
    }
    if (fast) {
      if (!launch_row3<R>(b, curl, upd, stream)) launch_fast<R>(b, curl, upd, pml, stream);
    }
    else
      launch_step<R>(b, curl, upd, stream);
  }
  // ---- batched stepping: one step's launches recorded as items ---------------------------
  struct Collect {
    std::map<int, std::vector<BatchItem<R>>> items[2]; // phase (B, D) -> variant -> items
    std::vector<BatchConn> conn[2];
    std::vector<BatchDft<R>> dft;
    int64_t vofs = 0; // this chunk's offset in the batch's source row
  };
  Collect *collector = nullptr;
  int cur_phase = 0;
  // Records the launches of the next step; false when that step would allocate arrays or
  // needs the unfused source path (then it must run on its own first).
  bool collect_step(Collect &C) {
    for (auto &sv : sources[B_ft])
      if (srcs[sv.src].is_integrated) fail("integrated magnetic sources are not supported by the GPU engine");
    if (have_fmp_H()) fail("magnetic susceptibilities are not supported in batched stepping");
    bool have_fmp_E = false;
    for (auto &sv : sources[D_ft])
      if (srcs[sv.src].is_integrated) have_fmp_E = true;
    for (auto &p : pols)
      for (int c = 0; c < 5; ++c)
        for (int cmp = 0; cmp < ncmp(); ++cmp)
          if (needs_P(p, c, cmp)) have_fmp_E = true;
    if (!plan_valid) rebuild_plan();
    struct Guard {
      Fields *F;
      ~Guard() { F->collector = nullptr; }
    } guard{this};
    collector = &C;
    StepArgs<R> A = base_args();
    cur_phase = 0;
    for (int cmp = 0; cmp < ncmp(); ++cmp) {
      cur_cmp = cmp;
      int which[3];
      A.nc = curl_comps(B_ft, cmp, A.cc, which);
      A.cmp = cmp;
      A.npole = 0;
      set_batch_sources(A, 0, cmp, C.vofs);
      if (!splan.L[0][cmp].rest_cur.empty() || upd_pending(cmp, which, A.nc, false)) return false;
      const int nu = upd_comps(H_ft, cmp, which, A.nc, A.uc, false);
      step_kernel(A, true, nu > 0, MP_K_STEP_BH);
    }
    if (!plan_valid) return false;
    boundaries(0);
    alloc_pols();
    if (!plan_valid) return false;
    cur_phase = 1;
    for (int cmp = 0; cmp < ncmp(); ++cmp) {
      cur_cmp = cmp;
      int which[3];
      A.nc = curl_comps(D_ft, cmp, A.cc, which);
      A.cmp = cmp;
      upd_comps(E_ft, cmp, which, A.nc, A.uc, have_fmp_E);
      if (!plan_valid) return false;
      set_batch_sources(A, 1, cmp, C.vofs);
      A.npole = int(pols.size());
      for (int p = 0; p < A.npole; ++p) A.pole[p] = pols[p].coef;
      if (!splan.L[1][cmp].rest_cur.empty() || !splan.L[1][cmp].rest_dip.empty()) return false;
      step_kernel(A, true, true, MP_K_STEP_DE);
    }
    boundaries(1);
    for (auto &cp : dft_chunks) {
      DftChunk<R> &ch = *cp;
      if (!f[ch.c][0]) continue;
      if (ch.nfreq > 512) return false; // beyond the batched kernel's phase table
      BatchDft<R> bd;
      memset(&bd, 0, sizeof bd);
      bd.a.f0 = f[ch.c][0];
      bd.a.f1 = f[ch.c][1];
      bd.a.idx = static_cast<const int64_t *>(ch.idx.p);
      bd.a.w = static_cast<const double *>(ch.w.p);
      bd.a.omega = static_cast<const double *>(ch.omega.p);
      bd.a.scale_re = ch.scale.real();
      bd.a.scale_im = ch.scale.imag();
      bd.a.avg1 = ch.avg1;
      bd.a.avg2 = ch.avg2;
      bd.a.npts = ch.npts;
      bd.a.nfreq = ch.nfreq;
      bd.a.acc = ch.acc.p;
      bd.a.comp = ch.comp.p;
      bd.a.acc_kind = ch.acc_kind;
      bd.dt = dt;
      bd.half = (is_magnetic(ch.c) || is_B(ch.c)) ? 1 : 0;
      bd.decimation = ch.decimation;
      bd.t_start = t;
      if (ch.nsnap) fail("internal: deferred DFT steps pending at batch planning");
      const size_t need = sizeof(R) * size_t(ch.npts) * DFT_DEFER * (f[ch.c][1] ? 2 : 1);
      if (ch.snap.bytes < need) ch.snap.alloc(need);
      bd.snap0 = static_cast<R *>(ch.snap.p);
      bd.snap1 = f[ch.c][1] ? bd.snap0 + size_t(DFT_DEFER) * size_t(ch.npts) : nullptr;
      C.dft.push_back(bd);
    }
    return plan_valid;
  }
  // the finiteness probe of the centre points on stream s, and its verdict
  void probe_on(hipStream_t s) {
    if (n_chk)
      launch_check<R>(static_cast<const R *const *>(d_chk_ptrs.p), static_cast<const int64_t *>(d_chk_idx.p),
                      static_cast<const double *>(d_chk_w.p), n_chk, static_cast<int *>(d_flag.p), s);
  }
  void check_flag() {
    int flag = 0;
    HIPCHECK(hipMemcpy(&flag, d_flag.p, sizeof(int), hipMemcpyDeviceToHost));
    if (flag) fail("simulation fields are NaN or Inf");
  }
  void step_kernel(const StepArgs<R> &a, bool curl, bool upd, int kind) {
    if (!boxes_valid) compute_boxes();
    cur_phase = kind == MP_K_STEP_BH ? 0 : 1;
    kbegin(kind);
    StepArgs<R> b = a;
    // conductive curls run the generic kernel (the vectorised one carries no conductivity)
    bool cond_curl = false;
    if (curl)
      for (int d = 0; d < 3; ++d) cond_curl = cond_curl || (b.cc[d].f && b.cc[d].cnd);
    const bool fast = use_fast && !cond_curl && (!curl || fast_partners(kind == MP_K_STEP_BH ? B_ft : D_ft, cur_cmp, b));
    if (!interior_empty) {
      // backward differences read the previous plane: walk the planes downwards so it was
      // fetched just before (forward differences prefetch their neighbour naturally)
      b.rev = fast && kind == MP_K_STEP_DE ? 1 : 0;
      // the interior launch alone is timed too (the class of the phase's largest kernel)
      const bool mark = ktiming && !collector;
      hipEvent_t e0 = nullptr;
      if (mark) {
        e0 = get_event();
        HIPCHECK(hipEventRecord(e0, stream));
      }
      launch_box(b, interior.lo, interior.hi, curl, upd, fast, rows_pml);
      if (mark) {
        hipEvent_t e1 = get_event();
        HIPCHECK(hipEventRecord(e1, stream));
        pending.push_back({kind == MP_K_STEP_BH ? MP_K_BH_INTERIOR : MP_K_DE_INTERIOR, e0, e1});
      }
    }
    b.rev = 0;
    for (const Box &bx : slabs) launch_box(b, bx.lo, bx.hi, curl, upd, fast, true);
    kend();
  }
  // ---- per-step argument blocks ----
  Grid grid() const {
    Grid g;
    memset(&g, 0, sizeof g);
    for (int kx = 0; kx < 3; ++kx) {
      const int d = gv.ydir(kx);
      g.axdir[kx] = gv.has(d) ? d : -1;
    }
    for (int d = 0; d < 3; ++d) {
      g.st[d] = gv.st[d];
      g.num[d] = gv.num[d];
      g.blo[d] = 0;
      g.bhi[d] = gv.has(d) ? gv.num[d] : 0;
    }
    g.hi_metal = g.lo_magnetic = 0;
    for (int d = 0; d < 3; ++d)
      if (gv.has(d)) {
        if (bc[High][d] == Metallic) g.hi_metal |= 1 << d;
        if (bc[Low][d] == Magnetic) g.lo_magnetic |= 1 << d;
      }
    return g;
  }
  // The interior box: array indices j along each direction where no uPML profile is active
  // for either Yee shift (so every component there updates with the plain formulas), and
  // the slabs that cover the rest of the grid without overlap.
  struct Box {
    int lo[3], hi[3];
  };
  bool boxes_valid = false;
  int quiet_lo[3] = {1, 1, 1}, quiet_hi[3] = {0, 0, 0}; // per direction: no uPML profile active
  Box interior;
  bool interior_empty = false;
  std::vector<Box> slabs;
  // The launch boxes: the grid is split along every direction with uPML into an interior
  // box, where no uPML profile is active (the plain kernels), and slabs, where the kernels
  // take the uPML point by point.  Boxes along the stride-1 direction start and end inside
  // 16-byte groups; a group shared by two boxes is updated lane by lane by each.
  void compute_boxes() {
    int ilo[3], ihi[3], full_hi[3];
    interior_empty = false;
    rows_pml = false;
    for (int d = 0; d < 3; ++d) {
      full_hi[d] = gv.has(d) ? gv.num[d] : 0;
      ilo[d] = 0;
      ihi[d] = full_hi[d];
      if (!gv.has(d) || !s->pml_dir[d]) continue;
      auto quiet = [&](int j) {
        for (int k = 2 * j; k <= 2 * j + 1; ++k)
          if (s->sig[d][k] != 0.0 || s->kap[d][k] != 1.0) return false;
        return true;
      };
      while (ilo[d] <= full_hi[d] && !quiet(ilo[d])) ++ilo[d];
      while (ihi[d] >= ilo[d] && !quiet(ihi[d])) --ihi[d];
      for (int j = ilo[d]; j <= ihi[d]; ++j)
        if (!quiet(j)) fail("uPML profile active inside the interior of direction %d", d);
      if (ilo[d] > ihi[d]) interior_empty = true;
    }
    for (int d = 0; d < 3; ++d) {
      quiet_lo[d] = ilo[d];
      quiet_hi[d] = ihi[d];
    }
    const bool whole_rows = whole_rows_env >= 0 ? whole_rows_env != 0 : !batch_split;
    if (whole_rows) { // the stride-1 direction's uPML stays inside whole-row boxes
      const int d = gv.fast;
      rows_pml = gv.has(d) && (ilo[d] > 0 || ihi[d] < full_hi[d]);
      ilo[d] = 0;
      ihi[d] = full_hi[d];
    }
    slabs.clear();
    if (interior_empty) {
      Box b;
      for (int d = 0; d < 3; ++d) b.lo[d] = 0, b.hi[d] = full_hi[d];
      slabs.push_back(b);
    }
    else {
      for (int d = 0; d < 3; ++d) interior.lo[d] = ilo[d], interior.hi[d] = ihi[d];
      // slabs: along X over everything, then Y within the X-interior, then Z within both
      int clo[3], chi[3];
      for (int d = 0; d < 3; ++d) clo[d] = 0, chi[d] = full_hi[d];
      for (int d = 0; d < 3; ++d) {
        if (!gv.has(d)) continue;
        if (ilo[d] > 0) {
          Box b;
          for (int q = 0; q < 3; ++q) b.lo[q] = clo[q], b.hi[q] = chi[q];
          b.lo[d] = 0;
          b.hi[d] = ilo[d] - 1;
          slabs.push_back(b);
        }
        if (ihi[d] < full_hi[d]) {
          Box b;
          for (int q = 0; q < 3; ++q) b.lo[q] = clo[q], b.hi[q] = chi[q];
          b.lo[d] = ihi[d] + 1;
          b.hi[d] = full_hi[d];
          slabs.push_back(b);
        }
        clo[d] = ilo[d];
        chi[d] = ihi[d];
      }
    }
    boxes_valid = true;
  }
  bool pml_in(int d) const { return d >= 0 && d < 3 && s->pml_dir[d]; }

  // step_db's per-component setup (src/step_db.cpp:47-84); allocates f_u lazily as a copy of f
  int curl_comps(int ft, int cmp, CurlComp<R> *cc, int *which) {
    const int c0 = ft == B_ft ? 15 : 10;
    for (int d = 0; d < 3; ++d) {
      CurlComp<R> &k = cc[d];
      memset(&k, 0, sizeof k);
      k.dsig = k.dsigu = -1;
      which[d] = -1;
      const int cc_ = dir_comp(c0, d);
      if (!has_field(cc_) || !f[cc_][cmp]) continue;
      const int dc = comp_dir(cc_);
      const int dsig0 = cycle_dir(dc, 1), dsigu0 = cycle_dir(dc, 2);
      k.dsig = pml_in(dsig0) ? dsig0 : -1;
      k.dsigu = pml_in(dsigu0) ? dsigu0 : -1;
      if (k.dsigu >= 0 && !f_u[cc_][cmp]) f_u[cc_][cmp] = daux(f[cc_][cmp]);
      // f_cond where uPML acts on f and the component conducts (src/step_db.cpp:66-69)
      if (k.dsig >= 0 && d_cond[cc_] && !f_cond[cc_][cmp]) f_cond[cc_][cmp] = daux(nullptr);
      k.f = f[cc_][cmp];
      k.fu = f_u[cc_][cmp];
      k.cnd = d_cond[cc_];
      k.cndinv = d_condinv[cc_];
      k.fcnd = f_cond[cc_][cmp];
      k.dt2 = R(dt * 0.5);
      const R *g1 = have_p[cc_] ? f[plus_c[cc_]][cmp] : nullptr;
      const R *g2 = have_m[cc_] ? f[minus_c[cc_]][cmp] : nullptr;
      int64_t s1 = have_p[cc_] ? gv.st[plus_d[cc_]] : 0, s2 = have_m[cc_] ? gv.st[minus_d[cc_]] : 0;
      if (ft == D_ft) {
        s1 = -s1;
        s2 = -s2;
      }
      R dtdx = R(courant);
      if (!g1) {
        std::swap(g1, g2);
        std::swap(s1, s2);
        dtdx = -dtdx;
      }
      k.g1 = g1;
      k.g2 = g2;
      k.s1 = s1;
      k.s2 = s2;
      k.dtdx = dtdx;
      for (int q = 0; q < 3; ++q) {
        k.lo[q] = gv.own_lo(cc_, q);
        k.hi[q] = gv.own_hi(cc_, q);
        k.ysh[q] = gv.ysh(cc_, q);
      }
      // the owned range keeps i + s1 and i + s2 inside the array (checked, not assumed)
      for (int which_s = 0; which_s < 2; ++which_s) {
        const bool used = which_s == 0 ? (g1 != nullptr) : (g2 != nullptr);
        if (!used) continue;
        const int64_t sv = which_s == 0 ? s1 : s2;
        int dd = -1;
        for (int q = 0; q < 3; ++q)
          if (gv.has(q) && (gv.st[q] == sv || gv.st[q] == -sv)) dd = q;
        if (dd < 0) fail("internal: curl stride matches no direction");
        if ((sv > 0 && k.hi[dd] + 1 > gv.num[dd]) || (sv < 0 && k.lo[dd] - 1 < 0))
          fail("internal: curl neighbour outside the array");
      }
      which[d] = cc_;
    }
    return 3;
  }
  // Partner description for the vectorised kernel: F[d] is the partner component along d and
  // slot d's plus / minus partners must be F[(d+2)%3] along d+1 and F[(d+1)%3] along d+2.
  // Returns false when the plan has another shape (then the generic kernel runs everywhere).
  bool fast_partners(int ft, int cmp, StepArgs<R> &a) const {
    const int p0 = ft == B_ft ? 0 : 5, c0 = ft == B_ft ? 15 : 10;
    for (int d = 0; d < 3; ++d) a.F[d] = has_field(dir_comp(p0, d)) ? f[dir_comp(p0, d)][cmp] : nullptr;
    for (int d = 0; d < 3; ++d) {
      a.hasp[d] = a.hasm[d] = 0;
      const int cc_ = dir_comp(c0, d);
      if (!has_field(cc_) || !f[cc_][cmp]) continue;
      if (have_p[cc_] && f[plus_c[cc_]][cmp]) {
        if (comp_dir(plus_c[cc_]) != (d + 2) % 3 || plus_d[cc_] != (d + 1) % 3) return false;
        a.hasp[d] = 1;
      }
      if (have_m[cc_] && f[minus_c[cc_]][cmp]) {
        if (comp_dir(minus_c[cc_]) != (d + 1) % 3 || minus_d[cc_] != (d + 2) % 3) return false;
        a.hasm[d] = 1;
      }
    }
    a.sgn = ft == B_ft ? 1 : -1;
    a.dtdx = R(courant);
    return true;
  }
  // update_eh's per-component setup (src/update_eh.cpp:148-195), aligned with the curl
  // components in src (entry q updates the E/H partner of src[q]).  Allocates H (a copy of B)
  // and W (a copy of E/H) lazily, as the reference does on its first call; an entry whose
  // E/H still aliases D/B has f == nullptr and is skipped.  Returns how many entries update.
  int upd_comps(int ft, int cmp, const int *src, int n, UpdComp<R> *uc, bool have_fmp) {
    int active = 0;
    for (int q = 0; q < n; ++q) {
      UpdComp<R> &e = uc[q];
      memset(&e, 0, sizeof e);
      e.dsigw = -1;
      if (src[q] < 0) continue;
      const int dc = src[q], ec = dc - 10;
      for (int d = 0; d < 3; ++d) {
        e.lo[d] = gv.own_lo(ec, d);
        e.hi[d] = gv.own_hi(ec, d);
        e.ysh[d] = gv.ysh(ec, d);
      }
      e.dsigw = -1;
      if (!f[ec][cmp]) continue;
      const int dsigw = pml_in(comp_dir(ec)) ? comp_dir(ec) : -1;
      const R *u = d_chi1inv[ec];
      if (f[ec][cmp] == f[dc][cmp] && (u || have_fmp || dsigw >= 0)) {
        f[ec][cmp] = dalloc(f[dc][cmp]);
        plan_valid = false;
      }
      if (dsigw >= 0 && !f_w[ec][cmp]) f_w[ec][cmp] = daux(f[ec][cmp]);
      if (f[ec][cmp] == f[dc][cmp]) continue;
      e.f = f[ec][cmp];
      e.g = f[dc][cmp];
      e.u = u;
      e.fw = f_w[ec][cmp];
      e.dsigw = dsigw;
      for (int p = 0; p < MAXPOLE; ++p)
        if (p < int(pols.size()) && pols[p].allocated) {
          e.P[p] = pols[p].P[ec][cmp];
          e.Pp[p] = pols[p].Pp[ec][cmp];
          e.sigma[p] = pols[p].sigma[ec];
          if (e.sigma[p] && pols[p].sig_uni[ec]) {
            e.sigone |= 1 << p;
            e.sigval[p] = pols[p].sig_val[ec];
          }
        }
      ++active;
    }
    return active;
  }
  // whether update_eh(ft) would allocate something now (then the curl runs on its own first)
  bool upd_pending(int cmp, const int *src, int n, bool have_fmp) const {
    for (int q = 0; q < n; ++q) {
      if (src[q] < 0) continue;
      const int dc = src[q], ec = dc - 10;
      if (!f[ec][cmp]) continue;
      const int dsigw = pml_in(comp_dir(ec)) ? comp_dir(ec) : -1;
      if (f[ec][cmp] == f[dc][cmp] && (d_chi1inv[ec] || have_fmp || dsigw >= 0)) return true;
      if (dsigw >= 0 && !f_w[ec][cmp]) return true;
    }
    return false;
  }

  bool needs_P(const Pol &p, int c, int cmp) const {
    return (is_electric(c) || is_magnetic(c)) && p.sigma[c] && f[c][cmp];
  }
  // some magnetic susceptibility polarises an H component (then H is its own array and the
  // B phase subtracts and updates P)
  bool have_fmp_H() const {
    for (auto &p : pols)
      for (int c = 5; c < 10; ++c)
        for (int cmp = 0; cmp < ncmp(); ++cmp)
          if (needs_P(p, c, cmp)) return true;
    return false;
  }
  // zeroed arrays of n * ntot values (n consecutive state arrays)
  R *dalloc_n(size_t n) {
    bufs.emplace_back(new DevBuf);
    bufs.back()->alloc(sizeof(R) * gv.ntot * n);
    R *p = static_cast<R *>(bufs.back()->p);
    HIPCHECK(hipMemsetAsync(p, 0, sizeof(R) * gv.ntot * n, stream));
    return p;
  }
  // lorentzian_data (src/susceptibility.cpp:98-139): P and P_prev of the poles allocated
  // now, for each component, in one block of consecutive arrays (P, P_prev of the first
  // pole, then of the next), so the row kernel addresses them from one base pointer
  void alloc_pols() {
    bool any = false;
    for (auto &p : pols) any = any || !p.allocated;
    if (!any) return;
    for (int c = 0; c < 10; ++c)
      for (int cmp = 0; cmp < ncmp(); ++cmp) {
        std::vector<Pol *> need;
        for (auto &p : pols)
          if (!p.allocated && needs_P(p, c, cmp)) need.push_back(&p);
        if (need.empty()) continue;
        R *base = dalloc_n(2 * need.size());
        for (size_t j = 0; j < need.size(); ++j) {
          need[j]->P[c][cmp] = base + (2 * j) * gv.ntot;
          need[j]->Pp[c][cmp] = base + (2 * j + 1) * gv.ntot;
        }
      }
    for (auto &p : pols) p.allocated = true;
    plan_valid = false;
  }

  // ---- source plan ----
  // Every launch's sources sit in a device-resident SrcSet; the values they subtract come
  // from a row of doubles per step, formed on the host in FP64 for all the steps of a call
  // (SrcRecipe: amp * current * dt or amp * dipole at the step's time) and uploaded once.
  // Long lists that fit no box are applied after the curl with k_src (LongSrc).
  struct LongSrc {
    SrcVol *sv;
    int slot;
    double mult;
    double toff; // time of the list within the step, in units of dt
    bool integrated;
  };
  struct SrcLaunch { // the sources of one (phase, cmp) launch
    SrcSet set;      // host copy (vbase relative to the chunk's row)
    int64_t lo = INT64_MAX, hi = -1;  // flat index range of the source points
    int blo[3] = {0, 0, 0}, bhi[3] = {-1, -1, -1}; // their loop-axis box
    std::vector<LongSrc> rest_cur, rest_dip;
  };
  struct SrcPlan {
    long pv = -1, sv = -1;
    SrcLaunch L[2][2]; // [B phase, D phase][cmp]
    const SrcSet *dev = nullptr; // the four SrcSets, index 2 * phase + cmp
    std::vector<SrcRecipe> rec;  // layout of the per-step row
  } splan;
  // The SrcSets and the rows live in set_buf and rows_buf.  A buffer replaced while launches
  // that read it may still be queued moves to `keep`, which step_n empties after it has
  // synchronised the stream; rows_free says that no queued launch reads rows_buf.
  std::unique_ptr<DevBuf> set_buf, rows_buf;
  // source rows staged in the caller-independent scratch array by step_n_noalloc
  const double *ext_rows = nullptr;
  long ext_t0 = 0, ext_n = 0;
  std::vector<std::unique_ptr<DevBuf>> keep;
  bool rows_free = true;
  long rows_t0 = 0, rows_n = 0, rows_hint = 1;

  // the scalar of a long list at step tt: current(time) (times dt in mult) or dipole(time)
  cdouble long_scal(const LongSrc &ls, long tt) const {
    const SrcTime &st = srcs[ls.sv->src];
    const double time = double(tt) * dt + ls.toff * dt;
    return ls.integrated ? st.dipole(time) : st.current(time, dt);
  }
  // the step's source row value of recipe r at step tt
  double recipe_value(const SrcRecipe &r, long tt) const {
    const SrcTime &src = srcs[r.src];
    const double time = double(tt) * dt + r.toff * dt;
    const cdouble scal = r.toff == 1.0 ? src.dipole(time) : src.current(time, dt);
    if (r.kind == 0) {
      const cdouble A = (r.amp * scal) * r.mult;
      return r.cmp ? A.imag() : A.real();
    }
    return r.cmp ? scal.imag() : scal.real();
  }
  // which[d]: the component of slot d of field type ft (-1 when absent)
  void slot_comps(int ft, int cmp, int *which) const {
    const int c0 = ft == B_ft ? 15 : 10;
    for (int d = 0; d < 3; ++d) {
      const int c = dir_comp(c0, d);
      which[d] = has_field(c) && f[c][cmp] ? c : -1;
    }
  }
  void build_srcplan() {
    splan.rec.clear();
    for (int ph = 0; ph < 2; ++ph)
      for (int cmp = 0; cmp < 2; ++cmp) {
        SrcLaunch &L = splan.L[ph][cmp];
        memset(&L.set, 0, sizeof L.set);
        L.lo = INT64_MAX;
        L.hi = -1;
        for (int k = 0; k < 3; ++k) {
          L.blo[k] = INT32_MAX;
          L.bhi[k] = INT32_MIN;
        }
        L.rest_cur.clear();
        L.rest_dip.clear();
        if (cmp >= ncmp()) continue;
        int which[3], whiche[3];
        slot_comps(ph ? D_ft : B_ft, cmp, which);
        fill_tab(ph ? D_ft : B_ft, cmp, which, 3, false, L.set.cur, L.set.curb, L.rest_cur, ph ? 0.5 : 0.0);
        if (ph) {
          for (int q = 0; q < 3; ++q) whiche[q] = which[q] < 0 ? -1 : which[q] - 10;
          fill_tab(D_ft, cmp, whiche, 3, true, L.set.dip, L.set.dipb, L.rest_dip, 1.0);
        }
        for (const SrcTab *tb : {&L.set.cur, &L.set.dip})
          if (tb->n) {
            L.lo = std::min(L.lo, tb->lo);
            L.hi = std::max(L.hi, tb->hi);
            for (int e = 0; e < tb->n; ++e) {
              int j[3];
              index_to_axes(tb->idx[e], j);
              for (int k = 0; k < 3; ++k) {
                L.blo[k] = std::min(L.blo[k], j[k]);
                L.bhi[k] = std::max(L.bhi[k], j[k]);
              }
            }
          }
        for (const SrcBoxes *bx : {&L.set.curb, &L.set.dipb}) {
          if (bx->n < 0 || bx->n > SRC_BOX) fail("internal: bad source box count");
          for (int q = 0; q < bx->n; ++q) {
            if (!bx->amp[q] || bx->comp[q] < 0 || bx->comp[q] >= 3 || which[bx->comp[q]] < 0) fail("internal: bad source box %d", q);
            int64_t first = 0, last = 0;
            for (int k = 0; k < 3; ++k) {
              const int d = gv.ydir(k);
              const int hi = gv.has(d) ? gv.num[d] : 0;
              if (bx->lo[q][k] < 0 || bx->hi[q][k] > hi || bx->lo[q][k] > bx->hi[q][k]) fail("internal: source box outside the grid");
              const int64_t st = gv.has(d) ? gv.st[d] : 0;
              first += bx->lo[q][k] * st;
              last += bx->hi[q][k] * st;
            }
            L.lo = std::min(L.lo, first);
            L.hi = std::max(L.hi, last);
            for (int k = 0; k < 3; ++k) {
              L.blo[k] = std::min(L.blo[k], bx->lo[q][k]);
              L.bhi[k] = std::max(L.bhi[k], bx->hi[q][k]);
            }
          }
        }
        if (L.lo > L.hi)
          for (int k = 0; k < 3; ++k) L.blo[k] = 1, L.bhi[k] = 0;
      }
    std::vector<SrcSet> sets(4);
    for (int ph = 0; ph < 2; ++ph)
      for (int cmp = 0; cmp < 2; ++cmp) sets[2 * ph + cmp] = splan.L[ph][cmp].set;
    if (set_buf) keep.push_back(std::move(set_buf));
    set_buf.reset(new DevBuf);
    set_buf->alloc(sizeof(SrcSet) * 4);
    HIPCHECK(hipMemcpy(set_buf->p, sets.data(), sizeof(SrcSet) * 4, hipMemcpyHostToDevice));
    splan.dev = static_cast<const SrcSet *>(set_buf->p);
    splan.pv = plan_version;
    splan.sv = src_version;
    rows_n = 0;
  }
  bool srcplan_stale() const { return splan.pv != plan_version || splan.sv != src_version; }
  // loop-axis indices of array index i (0 along absent axes)
  void index_to_axes(int64_t i, int *j) const {
    for (int k = 0; k < 3; ++k) {
      const int d = gv.ydir(k);
      if (!gv.has(d)) {
        j[k] = 0;
        continue;
      }
      j[k] = int(i / gv.st[d]);
      i -= int64_t(j[k]) * gv.st[d];
    }
  }
  void ensure_srcplan() {
    if (!plan_valid) rebuild_plan();
    if (srcplan_stale()) build_srcplan();
  }
  // the source row of step tt (rows for rows_hint steps from tt are formed and uploaded when
  // tt is not covered); null when no launch has sources
  const double *src_row(long tt) {
    if (srcplan_stale()) build_srcplan();
    if (splan.rec.empty()) return nullptr;
    const size_t rl = splan.rec.size();
    if (ext_rows) { // rows staged by step_n_noalloc
      if (tt < ext_t0 || tt >= ext_t0 + ext_n) fail("internal: step outside the staged source rows");
      return ext_rows + size_t(tt - ext_t0) * rl;
    }
    if (tt < rows_t0 || tt >= rows_t0 + rows_n) {
      const long n = std::max<long>(rows_hint, 1);
      std::vector<double> rows(rl * size_t(n));
      for (long st = 0; st < n; ++st)
        for (size_t q = 0; q < rl; ++q) rows[size_t(st) * rl + q] = recipe_value(splan.rec[q], tt + st);
      const size_t need = sizeof(double) * rows.size();
      if (!rows_buf || !rows_free || rows_buf->bytes < need) {
        if (rows_buf) keep.push_back(std::move(rows_buf));
        rows_buf.reset(new DevBuf);
        rows_buf->alloc(need);
      }
      HIPCHECK(hipMemcpy(rows_buf->p, rows.data(), need, hipMemcpyHostToDevice));
      rows_free = false;
      rows_t0 = tt;
      rows_n = n;
    }
    return static_cast<const double *>(rows_buf->p) + size_t(tt - rows_t0) * rl;
  }
  // the sources of launch (phase, cmp) of a batched step (values from the batch's row at vofs)
  void set_batch_sources(StepArgs<R> &A, int ph, int cmp, int64_t vofs) {
    if (srcplan_stale()) build_srcplan();
    const SrcLaunch &L = splan.L[ph][cmp];
    const bool any = L.lo <= L.hi;
    A.src = any ? splan.dev + (2 * ph + cmp) : nullptr;
    for (int k = 0; k < 3; ++k) {
      A.sbox_lo[k] = any ? L.blo[k] : 1;
      A.sbox_hi[k] = any ? L.bhi[k] : 0;
    }
    A.vrow = nullptr;
    A.vofs = vofs;
  }
  // the sources of launch (phase, cmp) at step tt into A
  void set_sources(StepArgs<R> &A, int ph, int cmp, long tt) {
    const double *row = src_row(tt);
    const SrcLaunch &L = splan.L[ph][cmp];
    const bool any = L.lo <= L.hi;
    A.src = any ? splan.dev + (2 * ph + cmp) : nullptr;
    for (int k = 0; k < 3; ++k) {
      A.sbox_lo[k] = any ? L.blo[k] : 1;
      A.sbox_hi[k] = any ? L.bhi[k] : 0;
    }
    A.vrow = any ? row : nullptr;
    A.vofs = 0;
  }

  // Sources of field type ft2 whose integrated flag is `integrated`, for the components in
  // slots `which`: short lists go into the table, long ones into dense index boxes while they
  // fit (SRC_BOX, and a box at most 8x its points; the long lists of one source time and
  // component share one box), the rest into `rest`, which the caller applies with k_src.
  // Each table entry and box appends its recipes to the row layout.  Values are
  // (amp * current(t)) * dt for currents (src/step.cpp:307-315) and amp * dipole(t) for
  // integrated sources (src/update_eh.cpp:128-138), t = (step + toff) dt.
  void fill_tab(int ft2, int cmp, const int *which, int nc, bool integrated, SrcTab &tab, SrcBoxes &bx,
                std::vector<LongSrc> &rest, double toff) {
    tab.n = 0;
    tab.lo = INT64_MAX;
    tab.hi = -1;
    bx.n = 0;
    std::vector<SrcRecipe> trec, brec;
    std::vector<LongSrc> longs;
    for (SrcVol &sv : sources[ft2]) {
      const SrcTime &st = srcs[sv.src];
      if (st.is_integrated != integrated) continue;
      const double mult = integrated ? 1.0 : dt;
      const int c2 = dir_comp(ft2 == D_ft ? 10 : 15, comp_dir(sv.c));
      int slot = -1;
      for (int q = 0; q < nc; ++q)
        if (which[q] == (integrated ? sv.c : c2)) slot = q;
      if (slot < 0) continue;
      if (tab.n + int(sv.index.size()) <= SRC_TAB) {
        for (size_t j = 0; j < sv.index.size(); ++j) {
          tab.comp[tab.n] = slot;
          tab.idx[tab.n] = sv.index[j];
          tab.lo = std::min(tab.lo, sv.index[j]);
          tab.hi = std::max(tab.hi, sv.index[j]);
          ++tab.n;
          trec.push_back({sv.src, sv.amp[j], mult, cmp, toff, 0});
        }
        continue;
      }
      longs.push_back({&sv, slot, mult, toff, integrated});
    }
    std::vector<char> taken(longs.size(), 0);
    for (size_t q = 0; q < longs.size(); ++q) {
      if (taken[q]) continue;
      std::vector<SrcVol *> group;
      for (size_t r = q; r < longs.size(); ++r)
        if (!taken[r] && longs[r].sv->src == longs[q].sv->src && longs[r].slot == longs[q].slot) {
          group.push_back(longs[r].sv);
          taken[r] = 1;
        }
      const MergedBox *m = merged_box(ft2, group);
      if (!m || bx.n >= SRC_BOX) {
        for (SrcVol *sv : group) {
          prepare_src(*sv);
          rest.push_back({sv, longs[q].slot, longs[q].mult, toff, integrated});
        }
        continue;
      }
      const int b = bx.n++;
      bx.comp[b] = longs[q].slot;
      for (int k = 0; k < 3; ++k) {
        const int d = gv.ydir(k);
        const bool has = gv.has(d);
        bx.lo[b][k] = has ? m->blo[d] : 0;
        bx.hi[b][k] = has ? m->bhi[d] : 0;
        bx.bst[b][k] = has ? m->bst[d] : 0;
      }
      bx.amp[b] = static_cast<const double2 *>(m->amp.p);
      bx.mult[b] = longs[q].mult;
      brec.push_back({longs[q].sv->src, 1.0, 1.0, 0, toff, 1});
      brec.push_back({longs[q].sv->src, 1.0, 1.0, 1, toff, 1});
    }
    tab.vbase = int(splan.rec.size());
    splan.rec.insert(splan.rec.end(), trec.begin(), trec.end());
    bx.vbase = int(splan.rec.size());
    splan.rec.insert(splan.rec.end(), brec.begin(), brec.end());
  }
  // One dense box holding the summed amplitudes of several source lists (one source time,
  // one component) over their joint index box, for batched steps.  A volume that crosses a
  // periodic boundary arrives as several lists that meet at the cell's edge points; the
  // batch subtracts their sum once where the per-chunk path subtracts list by list, which
  // differs only in rounding at those points.  Null when the box would exceed 8x its points.
  struct MergedBox {
    int blo[3], bhi[3];
    int64_t bst[3];
    DevBuf amp;
    long version = -1;
    bool ok = false;
  };
  std::map<std::tuple<int, int, int>, std::unique_ptr<MergedBox>> merged; // (ft, src, component)
  long src_version = 0; // bumped whenever a source list is added or changed
  const MergedBox *merged_box(int ft2, const std::vector<SrcVol *> &group) {
    auto &slot = merged[std::make_tuple(ft2, group[0]->src, group[0]->c)];
    if (!slot) slot.reset(new MergedBox);
    MergedBox &m = *slot;
    if (m.version == src_version) return m.ok ? &m : nullptr;
    m.version = src_version;
    m.ok = false;
    for (int d = 0; d < 3; ++d) {
      m.blo[d] = INT32_MAX;
      m.bhi[d] = INT32_MIN;
    }
    for (const SrcVol *sv : group)
      for (size_t j = 0; j < sv->index.size(); ++j)
        for (int d = 0; d < 3; ++d) {
          m.blo[d] = std::min(m.blo[d], sv->jpos[3 * j + d]);
          m.bhi[d] = std::max(m.bhi[d], sv->jpos[3 * j + d]);
        }
    size_t vol = 1;
    for (int d = 2; d >= 0; --d) {
      m.bst[d] = int64_t(vol);
      vol *= size_t(m.bhi[d] - m.blo[d] + 1);
    }
    std::vector<double2> dense(vol, make_double2(0.0, 0.0));
    std::vector<char> seen(vol, 0);
    size_t distinct = 0;
    for (const SrcVol *sv : group)
      for (size_t j = 0; j < sv->index.size(); ++j) {
        int64_t k = 0;
        for (int d = 0; d < 3; ++d) k += int64_t(sv->jpos[3 * j + d] - m.blo[d]) * m.bst[d];
        distinct += !seen[k];
        seen[k] = 1;
        dense[k].x += sv->amp[j].real();
        dense[k].y += sv->amp[j].imag();
      }
    if (vol > 8 * distinct) return nullptr;
    m.amp.alloc(sizeof(double2) * vol);
    HIPCHECK(hipMemcpy(m.amp.p, dense.data(), sizeof(double2) * vol, hipMemcpyHostToDevice));
    m.ok = true;
    return &m;
  }
  // device copies of a long source list: indices and amplitudes
  void prepare_src(SrcVol &sv) {
    if (!sv.dirty) return;
    const size_t n = sv.index.size();
    if (!sv.d_idx) {
      alloc_allowed("a source list");
      HIPCHECK(hipMalloc(&sv.d_idx, sizeof(int64_t) * n));
      HIPCHECK(hipMalloc(&sv.d_amp, sizeof(double2) * n));
    }
    std::vector<double2> amp(n);
    for (size_t j = 0; j < n; ++j) amp[j] = make_double2(sv.amp[j].real(), sv.amp[j].imag());
    HIPCHECK(hipMemcpy(sv.d_idx, sv.index.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(sv.d_amp, amp.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
    sv.dirty = false;
  }
  // A long list applied after the curl kernel: the kernel has already written the folded
  // periodic images of these points, so they are written again from the updated values.
  // (Integrated lists restore D afterwards and need no images.)
  void apply_long(const LongSrc &ls, R *fbase, int cmp, bool images = true) {
    const cdouble scal = long_scal(ls, t);
    kbegin(MP_K_SRC);
    launch_src<R>(fbase, ls.sv->d_idx, ls.sv->d_amp, int(ls.sv->index.size()), make_double2(scal.real(), scal.imag()),
                  ls.mult, cmp, stream);
    if (images) {
      prepare_images(*ls.sv);
      launch_src_images<R>(fbase, ls.sv->d_idx, ls.sv->d_img_k, ls.sv->d_img_dst, ls.sv->d_img_sgn, ls.sv->n_img, stream);
    }
    kend();
  }
  long plan_version = 0;
  // bumped when a step swaps the P / P_prev pointers (pswap) and when a DFT monitor is added
  // or removed: a batch plan records device pointers and DFT chunk lists, so either makes it stale
  long ptr_version = 0, dft_version = 0;
  void prepare_images(SrcVol &sv) {
    if (sv.img_version == plan_version) return;
    std::vector<int> ks;
    std::vector<int64_t> dsts;
    std::vector<int8_t> sgs;
    const int c = sv.c;
    for (size_t k = 0; k < sv.index.size(); ++k) {
      int dd[3], sg[3];
      for (int d = 0; d < 3; ++d) {
        dd[d] = -1;
        sg[d] = 1;
        if (!gv.has(d) || !wrap_dir[d]) continue;
        const int j = sv.jpos[3 * k + d], y = gv.ysh(c, d);
        const int dst = y == 0 ? (j == gv.num[d] ? 0 : -1) : (j == 0 ? gv.num[d] : -1);
        if (dst >= 0) dd[d] = dst, sg[d] = wrap_dir[d];
      }
      for (int sub = 1; sub < 8; ++sub) {
        int64_t off = 0;
        int sign = 1;
        bool ok = true;
        for (int d = 0; d < 3; ++d)
          if ((sub >> d) & 1) {
            ok = ok && dd[d] >= 0;
            if (dd[d] >= 0) {
              off += int64_t(dd[d] - sv.jpos[3 * k + d]) * gv.st[d];
              sign *= sg[d];
            }
          }
        if (!ok) continue;
        const int64_t dst = sv.index[k] + off;
        if (dst < 0 || size_t(dst) >= gv.ntot) fail("internal: source image outside the array");
        ks.push_back(int(k));
        dsts.push_back(dst);
        sgs.push_back(int8_t(sign));
      }
    }
    alloc_allowed("the periodic images of a source list");
    if (sv.d_img_k) (void)hipFree(sv.d_img_k);
    if (sv.d_img_dst) (void)hipFree(sv.d_img_dst);
    if (sv.d_img_sgn) (void)hipFree(sv.d_img_sgn);
    sv.d_img_k = nullptr;
    sv.d_img_dst = nullptr;
    sv.d_img_sgn = nullptr;
    sv.n_img = int(ks.size());
    if (sv.n_img) {
      HIPCHECK(hipMalloc(&sv.d_img_k, sizeof(int) * ks.size()));
      HIPCHECK(hipMalloc(&sv.d_img_dst, sizeof(int64_t) * dsts.size()));
      HIPCHECK(hipMalloc(&sv.d_img_sgn, sgs.size()));
      HIPCHECK(hipMemcpy(sv.d_img_k, ks.data(), sizeof(int) * ks.size(), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(sv.d_img_dst, dsts.data(), sizeof(int64_t) * dsts.size(), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(sv.d_img_sgn, sgs.data(), sgs.size(), hipMemcpyHostToDevice));
    }
    sv.img_version = plan_version;
  }

  // One fields::step() (src/step.cpp:35-139); the finiteness probe is launched when probe.
  void one_step(bool probe) {
    bool int_B = false, int_D = false;
    for (auto &sv : sources[B_ft])
      if (srcs[sv.src].is_integrated) int_B = true;
    for (auto &sv : sources[D_ft])
      if (srcs[sv.src].is_integrated) int_D = true;
    if (int_B) fail("integrated magnetic sources are not supported by the GPU engine");
    bool have_fmp_E = int_D;
    for (auto &p : pols)
      for (int c = 0; c < 5; ++c)
        for (int cmp = 0; cmp < ncmp(); ++cmp)
          if (needs_P(p, c, cmp)) have_fmp_E = true;

    // the polarisations exist from the first step's B phase on (update_pols(H_stuff) allocates
    // them there, src/step.cpp:85; the E ones are zero until the D phase either way)
    alloc_pols();
    if (!plan_valid) rebuild_plan();
    StepArgs<R> A = base_args();
    phase_B(A);
    // flux_vol::update_half (src/step.cpp:93): only the last step's values are ever read
    if (probe)
      for (auto &fv : flux_vols) flux_launch(fv->plane, static_cast<double *>(fv->partial.p));
    // --- D, then E = chi1inv (D - P - dipoles) and the polarisations ---
    alloc_pols();
    if (!plan_valid) rebuild_plan();
    // P/P_prev ping-pong where P has no halo images: with folded or listed periodic halos the
    // P array carries images while P_prev's halo points stay zero, as the reference keeps them
    const bool pswap = !pols.empty() && !p_conn && !(wrap_dir[0] | wrap_dir[1] | wrap_dir[2]);
    for (int cmp = 0; cmp < ncmp(); ++cmp) {
      cur_cmp = cmp;
      int which[3];
      A.nc = curl_comps(D_ft, cmp, A.cc, which);
      A.cmp = cmp;
      upd_comps(E_ft, cmp, which, A.nc, A.uc, have_fmp_E);
      if (!plan_valid) rebuild_plan();
      set_sources(A, 1, cmp, t);
      A.npole = int(pols.size());
      for (int p = 0; p < A.npole; ++p) A.pole[p] = pols[p].coef;
      A.pswap = pswap ? 1 : 0;
      const SrcLaunch &L = splan.L[1][cmp];
      if (L.rest_cur.empty() && L.rest_dip.empty()) {
        step_kernel(A, true, true, MP_K_STEP_DE);
        continue;
      }
      // long lists beyond the boxes: the curl (with the in-kernel sources) on its own, the
      // lists, then the update without the currents
      StepArgs<R> C = A;
      step_kernel(C, true, false, MP_K_STEP_DE);
      for (const LongSrc &ls : L.rest_cur) apply_long(ls, f[dir_comp(10, comp_dir(ls.sv->c))][cmp], cmp);
      // integrated lists beyond the boxes: f_minus_p formed in place in D around the update,
      // then those points of D restored.  Lists may share points (a periodic plane and its
      // image), so every list is saved before any is applied.
      size_t nsave = 0;
      for (const LongSrc &ls : L.rest_dip) nsave += ls.sv->index.size();
      if (nsave > d_scratch.bytes / sizeof(R)) d_scratch.alloc(sizeof(R) * nsave);
      size_t off = 0;
      for (const LongSrc &ls : L.rest_dip) {
        R *D = f[dir_comp(10, comp_dir(ls.sv->c))][cmp];
        launch_gather<R>(static_cast<R *>(d_scratch.p) + off, D, ls.sv->d_idx, int(ls.sv->index.size()), stream);
        off += ls.sv->index.size();
      }
      for (const LongSrc &ls : L.rest_dip) apply_long(ls, f[dir_comp(10, comp_dir(ls.sv->c))][cmp], cmp, false);
      step_kernel(A, false, true, MP_K_STEP_DE);
      off = 0;
      for (const LongSrc &ls : L.rest_dip) {
        R *D = f[dir_comp(10, comp_dir(ls.sv->c))][cmp];
        launch_scatter<R>(D, ls.sv->d_idx, static_cast<const R *>(d_scratch.p) + off, int(ls.sv->index.size()), stream);
        off += ls.sv->index.size();
      }
    }
    // the E pole updates wrote the new P into the P_prev arrays (the B phase does not swap)
    if (pswap) {
      for (auto &p : pols)
        for (int c = 0; c < 5; ++c)
          for (int cmp = 0; cmp < 2; ++cmp) std::swap(p.P[c][cmp], p.Pp[c][cmp]);
      ++ptr_version;
    }
    if (!plan_valid) rebuild_plan();
    boundaries(1); // D, P_E, E (step_boundaries D_stuff .. E_stuff)
    if (probe) // flux_vol::update (src/step.cpp:124)
      for (auto &fv : flux_vols) {
        flux_launch(fv->plane, static_cast<double *>(fv->partial.p) + FLUX_BLOCKS);
        fv->stepped = true;
      }
    t += 1;
    if (adj && adj->recording) record_design();
    if (!dft_off) update_dfts();
    if (probe && n_chk)
      launch_check<R>(static_cast<const R *const *>(d_chk_ptrs.p), static_cast<const int64_t *>(d_chk_idx.p),
                      static_cast<const double *>(d_chk_w.p), n_chk, static_cast<int *>(d_flag.p), stream);
  }

  StepArgs<R> base_args() const {
    StepArgs<R> A;
    memset(&A, 0, sizeof A);
    A.g = grid();
    for (int d = 0; d < 3; ++d) {
      A.pml.sig[d] = d_sig[d];
      A.pml.kap[d] = d_kap[d];
      A.pml.siginv[d] = d_siginv[d];
    }
    return A;
  }
  // B = B - curl E (+ B sources), then H = mu^-1 B where H is its own array, then the B and
  // H halos: step_db(B_stuff), step_source(B_stuff), step_boundaries(B_stuff),
  // update_eh(H_stuff), step_boundaries(H_stuff) (src/step.cpp:64-80)
  // With magnetic susceptibilities the update is H = mu^-1 (B - P) and update_P from the new H
  // (update_eh(H_stuff), update_pols(H_stuff), src/step.cpp:76-88), with P_prev = P; P = new P
  // (no pointer swap in this phase).
  void phase_B(StepArgs<R> &A) {
    const bool fmp_H = have_fmp_H();
    for (int cmp = 0; cmp < ncmp(); ++cmp) {
      cur_cmp = cmp;
      int which[3];
      A.nc = curl_comps(B_ft, cmp, A.cc, which);
      A.cmp = cmp;
      A.npole = fmp_H ? int(pols.size()) : 0;
      for (int p = 0; p < A.npole; ++p) A.pole[p] = pols[p].coef;
      A.pswap = 0;
      set_sources(A, 0, cmp, t);
      const SrcLaunch &L = splan.L[0][cmp];
      if (L.rest_cur.empty() && !upd_pending(cmp, which, A.nc, fmp_H)) {
        const int nu = upd_comps(H_ft, cmp, which, A.nc, A.uc, fmp_H);
        step_kernel(A, true, nu > 0, MP_K_STEP_BH);
        continue;
      }
      step_kernel(A, true, false, MP_K_STEP_BH);
      for (const LongSrc &ls : L.rest_cur) apply_long(ls, f[dir_comp(15, comp_dir(ls.sv->c))][cmp], cmp);
      StepArgs<R> U = A;
      U.src = nullptr;
      if (upd_comps(H_ft, cmp, which, U.nc, U.uc, fmp_H) > 0) step_kernel(U, false, true, MP_K_STEP_BH);
    }
    if (!plan_valid) rebuild_plan();
    boundaries(0); // B, H (step_boundaries B_stuff .. H_stuff)
  }

  // ---- field energy (src/energy_and_flux.cpp:45-178, src/integrate.cpp:46-205) ----
  // 0.5 * fields::integrate of Re(conj(F) G) (src/energy_and_flux.cpp:67-95) for F = E (H)
  // and G = D (B) of each component over [mn, mx], on the component's own grid
  long double energy_sum(int first, const V3 &mn, const V3 &mx) {
    long double sum = 0;
    for (int c = first; c < first + 5; ++c) {
      if (!has_field(c) || !f[c][0]) continue;
      const int pair[1][2] = {{c, c + 10}};
      sum += dot_integral(mn, mx, c, pair, 1) * 0.5;
    }
    return sum;
  }
  long double energy_of(int first) {
    V3 wmin, wmax;
    for (int d = 0; d < 3; ++d) wmax.t[d] = gv.has(d) ? gv.num[d] * gv.inva : 0.0;
    return energy_sum(first, wmin, wmax);
  }
  // fields::field_energy with synchronize_magnetic_fields / restore_magnetic_fields
  // (src/energy_and_flux.cpp:46-62,146-178): one more B phase, B and H averaged with their
  // values half a step earlier, the magnetic energy taken, every magnetic array restored.
  // synchronize_magnetic_fields / restore_magnetic_fields (src/energy_and_flux.cpp:146-178)
  // around fn: B and H advanced half a step on the device and averaged with their values at
  // t - dt/2, fn evaluated, every magnetic array and its auxiliaries restored
  template <class FN> double with_synchronized_magnetic(FN fn) {
    if (!plan_valid) rebuild_plan();
    HIPCHECK(hipStreamSynchronize(stream));
    struct Saved {
      void *p;
      size_t bytes;
      bool average; // a realnum field array, averaged with its saved values
      std::unique_ptr<DevBuf> b;
    };
    std::vector<Saved> saved;
    auto save_bytes = [&](void *p, size_t esz, bool average) {
      if (!p) return;
      Saved sv{p, esz * gv.ntot, average, std::unique_ptr<DevBuf>(new DevBuf)};
      sv.b->alloc(sv.bytes);
      HIPCHECK(hipMemcpyAsync(sv.b->p, p, sv.bytes, hipMemcpyDeviceToDevice, stream));
      saved.push_back(std::move(sv));
    };
    auto save = [&](auto *p, bool average) { save_bytes(p, sizeof(*p), average); };
    for (int cmp = 0; cmp < 2; ++cmp) {
      for (int c = 15; c < 20; ++c) {
        save(f[c][cmp], true);
        save(f_u[c][cmp], false);
        save(f_w[c][cmp], false);
        save(f_cond[c][cmp], false);
      }
      for (int c = 5; c < 10; ++c)
        if (f[c][cmp] && f[c][cmp] != f[c + 10][cmp]) {
          save(f[c][cmp], true);
          save(f_u[c][cmp], false);
          save(f_w[c][cmp], false);
        }
    }
    StepArgs<R> A = base_args();
    rows_hint = 1;
    phase_B(A);
    for (const Saved &sv : saved)
      if (sv.average) launch_average<R>(static_cast<R *>(sv.p), static_cast<const R *>(sv.b->p), gv.ntot, stream);
    HIPCHECK(hipStreamSynchronize(stream));
    const double value = fn();
    for (const Saved &sv : saved) HIPCHECK(hipMemcpyAsync(sv.p, sv.b->p, sv.bytes, hipMemcpyDeviceToDevice, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    return value;
  }
  double field_energy() override {
    const double magnetic = with_synchronized_magnetic([&] { return double(energy_of(5)); });
    return double(energy_of(0)) + magnetic;
  }

  // ---- time-domain flux (src/energy_and_flux.cpp:180-241, flux_vol src/meep.hpp:2322-2352) ----
  struct FluxPlane {
    int d = 0;
    int grid = 20; // the integration grid: a component's, or 20 (centred)
    int n = 0;
    DevBuf idx, w; // the points of the volume on that grid and their weights
  };
  struct FluxVol {
    FluxPlane plane;
    DevBuf partial; // FLUX_BLOCKS partial sums after the last step's H update, then after its E update
    bool stepped = false;
  };
  std::vector<std::unique_ptr<FluxVol>> flux_vols;
  // the pairs of flux_in_box_wrongH (src/energy_and_flux.cpp:190-204): E_0 x H_0 - E_1 x H_1
  static void flux_pairs(int d, int cE[2], int cH[2]) {
    if (d == 0)
      cE[0] = 1, cE[1] = 4, cH[0] = 9, cH[1] = 6;
    else if (d == 1)
      cE[0] = 4, cE[1] = 0, cH[0] = 5, cH[1] = 9;
    else
      cE[0] = 0, cE[1] = 1, cH[0] = 6, cH[1] = 5;
  }
  // add_flux_plane's direction (volume::normal_direction, src/vec.cpp:227-257) when d < 0
  int flux_direction(int d, const V3 &mn, const V3 &mx) const {
    if (d >= 0) {
      if (d > 2) fail("invalid direction in add_flux_vol");
      return d;
    }
    if (gv.dim == 1) return 2;
    int nd = -1, nzero = 0, npos = 0;
    for (int k = 0; k < 3; ++k)
      if (gv.has(k)) {
        const double ext = mx.t[k] - mn.t[k];
        if (ext == 0) nd = k, ++nzero;
        else if (ext > 0) ++npos;
      }
    if (nzero != 1 || npos != gv.dim - 1) fail("invalid direction in add_flux_vol");
    return nd;
  }
  // the integration points of fields::integrate (loop_in_chunks over the volume on `grid`,
  // src/integrate.cpp:122-165) with their weights; |shift phase|^2 folded in, as each field of
  // a product carries the phase of its periodic image
  void plan_points(FluxPlane &fp, int grid, const V3 &mn, const V3 &mx) {
    fp.grid = grid;
    std::vector<int64_t> idx;
    std::vector<double> w;
    loop_in_chunks(mn, mx, grid, [&](const IV &is, const IV &ie, const V3 &s0, const V3 &s1, const V3 &e0, const V3 &e1, double dV0, const IV &,
                                     cdouble sph) {
      const double ph2 = std::norm(sph);
      visit_box(is, ie, s0, s1, e0, e1, dV0, [&](int64_t i, int64_t, const IV &, double ww, const int *) {
        idx.push_back(i);
        w.push_back(ww * ph2);
      });
    });
    // every field read sits within the arrays: the centring offsets are at most one stride up
    int64_t reach = 0;
    for (int d = 0; d < 3; ++d)
      if (gv.has(d) && grid == 20) reach += gv.st[d];
    for (int64_t i : idx)
      if (i < 0 || size_t(i + reach) >= gv.ntot) fail("internal: integration point outside the array");
    fp.n = int(idx.size());
    if (idx.size() != size_t(fp.n)) fail("integration volume too large");
    if (fp.n) {
      fp.idx.alloc(sizeof(int64_t) * idx.size());
      fp.w.alloc(sizeof(double) * w.size());
      HIPCHECK(hipMemcpy(fp.idx.p, idx.data(), sizeof(int64_t) * idx.size(), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(fp.w.p, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice));
    }
  }
  void plan_flux(FluxPlane &fp, int d, const V3 &mn, const V3 &mx) {
    fp.d = d;
    plan_points(fp, 20, mn, mx);
  }
  // sum over the points of sign_p Re(conj(F_p) G_p) w for the pairs (F_p, G_p), signs +1, -1;
  // on the centred grid each field is the mean of its four Yee points, on a component's own
  // grid its value (the reference's four-point mean with zero offsets)
  void dot_launch(const FluxPlane &fp, const int (*pairs)[2], int np, double *partial) {
    if (!fp.n) return;
    FluxPlaneArgs<R> a;
    memset(&a, 0, sizeof a);
    a.idx = static_cast<const int64_t *>(fp.idx.p);
    a.w = static_cast<const double *>(fp.w.p);
    a.n = fp.n;
    for (int p = 0; p < np; ++p) {
      const int cF = pairs[p][0], cG = pairs[p][1];
      // a component this grid lacks reads as zero (src/integrate.cpp:103-111)
      if (!f[cF][0] || !f[cG][0]) continue;
      for (int cmp = 0; cmp < 2; ++cmp) {
        a.e[a.npair][cmp] = f[cF][cmp];
        a.h[a.npair][cmp] = f[cG][cmp];
      }
      if (fp.grid == 20) {
        gv.yee2cent(cF, a.eo[a.npair][0], a.eo[a.npair][1]);
        gv.yee2cent(cG, a.ho[a.npair][0], a.ho[a.npair][1]);
      }
      a.sign[a.npair++] = 1 - 2 * p;
    }
    if (a.npair) launch_flux_plane<R>(a, partial, stream);
  }
  void flux_launch(const FluxPlane &fp, double *partial) {
    int cE[2], cH[2];
    flux_pairs(fp.d, cE, cH);
    const int pairs[2][2] = {{cE[0], cH[0]}, {cE[1], cH[1]}};
    dot_launch(fp, pairs, 2, partial);
  }
  // fields::integrate of the pairs over [mn, mx] on `grid` (a temporary plan and partial sums)
  double dot_integral(const V3 &mn, const V3 &mx, int grid, const int (*pairs)[2], int np) {
    FluxPlane fp;
    plan_points(fp, grid, mn, mx);
    DevBuf part;
    part.alloc(FLUX_BLOCKS * sizeof(double));
    HIPCHECK(hipMemsetAsync(part.p, 0, FLUX_BLOCKS * sizeof(double), stream));
    dot_launch(fp, pairs, np, static_cast<double *>(part.p));
    return flux_sum(part.p, 0);
  }
  double flux_sum(const void *dpartial, int slot) {
    double h[FLUX_BLOCKS];
    HIPCHECK(hipStreamSynchronize(stream));
    HIPCHECK(hipMemcpy(h, static_cast<const double *>(dpartial) + slot * FLUX_BLOCKS, sizeof h, hipMemcpyDeviceToHost));
    long double sum = 0;
    for (double v : h) sum += v;
    return double(sum);
  }
  void *add_flux_vol(int d, V3 mn, V3 mx) override {
    std::unique_ptr<FluxVol> fv(new FluxVol);
    plan_flux(fv->plane, flux_direction(d, mn, mx), mn, mx);
    fv->partial.alloc(2 * FLUX_BLOCKS * sizeof(double));
    HIPCHECK(hipMemset(fv->partial.p, 0, 2 * FLUX_BLOCKS * sizeof(double)));
    flux_vols.push_back(std::move(fv));
    return flux_vols.back().get();
  }
  // flux_vol::flux (src/meep.hpp:2339): (the flux after E + the flux after H) / 2 of the last step
  double flux_vol_value(void *p) override {
    FluxVol *fv = static_cast<FluxVol *>(p);
    if (!fv->stepped) return 0.0;
    const double half = flux_sum(fv->partial.p, 0), full = flux_sum(fv->partial.p, 1);
    return (full + half) * 0.5;
  }
  double flux_in_box(int d, V3 mn, V3 mx, bool sync) override {
    int cE[2], cH[2];
    flux_pairs(flux_direction(d, mn, mx), cE, cH);
    const int pairs[2][2] = {{cE[0], cH[0]}, {cE[1], cH[1]}};
    auto value = [&] { return dot_integral(mn, mx, 20, pairs, 2); };
    return sync ? with_synchronized_magnetic(value) : value();
  }
  // fields::field_energy_in_box (src/energy_and_flux.cpp:54-59): the magnetic part with H
  // synchronised; electric_energy_in_box / magnetic_energy_in_box (:84-95) as they stand
  double energy_in_box(int which, V3 mn, V3 mx) override {
    if (which == 1) return double(energy_sum(0, mn, mx));
    if (which == 2) return double(energy_sum(5, mn, mx));
    const double magnetic = with_synchronized_magnetic([&] { return double(energy_sum(5, mn, mx)); });
    return double(energy_sum(0, mn, mx)) + magnetic;
  }

  // ---- reverse mode ----------------------------------------------------------------------
  // Gradient of J = sum |a|^2 over the accumulators a of a dft_fields monitor with respect to
  // chi1inv of one electric component on a design box, by the transposed step run backwards
  // (adjoint.hip).  The forward steps record D at the design points: every step (interval 0)
  // or, with checkpoints of the whole state every `interval` steps, recomputed per segment
  // during the backward sweep.
  struct Adj {
    Flux<R> *mon = nullptr;
    int c = -1;
    int lo[3], hi[3]; // design box, loop-axis array indices of component c
    size_t npts = 0;
    int interval = 0;
    long t0 = 0, nrec = 0;
    bool recording = false;
    std::vector<std::unique_ptr<DevBuf>> hist; // history mode: blocks of HBLK steps
    struct Checkpoint {
      long t;
      std::vector<std::unique_ptr<DevBuf>> data; // one per state array, in state_arrays() order
    };
    std::vector<Checkpoint> cps;
    DevBuf seg; // checkpoint mode: one segment of history
  };
  static constexpr long HBLK = 256;
  std::unique_ptr<Adj> adj;
  bool dft_off = false; // recomputation steps leave the DFT accumulators alone

  // every distinct state array (real parts; the reverse mode needs real fields) with its
  // element size: realnum fields, FP64 auxiliaries
  struct StateArr {
    void *p;
    size_t esz;
  };
  std::vector<StateArr> state_arrays() const {
    std::vector<StateArr> v;
    auto add = [&](void *p, size_t esz) {
      if (!p) return;
      for (const StateArr &a : v)
        if (a.p == p) return;
      v.push_back(StateArr{p, esz});
    };
    for (int c = 0; c < NC; ++c) {
      add(f[c][0], sizeof(R));
      add(f_u[c][0], sizeof(double));
      add(f_w[c][0], sizeof(double));
      add(f_cond[c][0], sizeof(double));
    }
    return v;
  }
  void adj_begin(void *monitor, int c, V3 mn, V3 mxv, int interval, int *box_lo, int *box_n) override {
    if (!is_real) fail("the reverse mode needs real fields (use_real_fields)");
    if (any_cond) fail("the reverse mode does not support conductivity");
    if (!pols.empty()) fail("the reverse mode does not support dispersive materials");
    for (auto &lst : sources)
      for (auto &sv : lst)
        if (srcs[sv.src].is_integrated) fail("the reverse mode does not support integrated sources");
    if (!is_electric(c) || !has_field(c)) fail("design component must be an electric field component");
    if (!d_chi1inv[c]) fail("design component has no chi1inv array (give the structure a non-trivial eps)");
    if (interval > 0 && t != 0) fail("checkpointed recording must start at t = 0");
    Flux<R> *mon = nullptr;
    for (auto &fl : fluxes)
      if (fl.get() == monitor) mon = fl.get();
    if (!mon || !mon->H.empty()) fail("the objective must be a dft_fields monitor of this fields object");
    std::unique_ptr<Adj> A(new Adj);
    A->mon = mon;
    A->c = c;
    // component c's points whose positions lie in [mn, mxv]
    for (int k = 0; k < 3; ++k) {
      const int d = gv.ydir(k);
      if (!gv.has(d)) {
        A->lo[k] = A->hi[k] = 0;
        continue;
      }
      const int y = gv.ysh(c, d);
      // position of index j: (2 j + y) / (2 a)
      A->lo[k] = std::max(0, int(std::ceil((mn.t[d] * 2 * a - y) / 2 - 1e-9)));
      A->hi[k] = std::min(gv.num[d], int(std::floor((mxv.t[d] * 2 * a - y) / 2 + 1e-9)));
      if (A->hi[k] < A->lo[k]) fail("empty design region");
      box_lo[d] = A->lo[k];
      box_n[d] = A->hi[k] - A->lo[k] + 1;
    }
    A->npts = 1;
    for (int k = 0; k < 3; ++k) A->npts *= size_t(A->hi[k] - A->lo[k] + 1);
    A->interval = interval;
    A->t0 = t;
    A->recording = true;
    adj = std::move(A);
    if (interval > 0) {
      // t = 0: every array is zero, including those still to be allocated lazily
      adj->cps.emplace_back();
      adj->cps.back().t = 0;
    }
  }
  void adj_end() override { adj.reset(); }
  void save_checkpoint() {
    HIPCHECK(hipStreamSynchronize(stream));
    typename Adj::Checkpoint cp;
    cp.t = t;
    for (const StateArr &a : state_arrays()) {
      cp.data.emplace_back(new DevBuf);
      cp.data.back()->alloc(a.esz * gv.ntot);
      HIPCHECK(hipMemcpyAsync(cp.data.back()->p, a.p, a.esz * gv.ntot, hipMemcpyDeviceToDevice, stream));
    }
    adj->cps.push_back(std::move(cp));
  }
  // the D at the design points after step (t - t0 - 1) -> history
  void record_design() {
    Adj &A = *adj;
    const long n = A.nrec;
    R *dst;
    if (A.interval > 0) {
      if (!seg_active) {
        A.nrec += 1;
        if ((t - A.t0) % A.interval == 0) save_checkpoint();
        return;
      }
      dst = static_cast<R *>(A.seg.p) + size_t(n - seg_first) * A.npts;
    }
    else {
      if (size_t(n / HBLK) >= A.hist.size()) {
        A.hist.emplace_back(new DevBuf);
        A.hist.back()->alloc(sizeof(R) * A.npts * HBLK);
      }
      dst = static_cast<R *>(A.hist[n / HBLK]->p) + size_t(n % HBLK) * A.npts;
    }
    const int dc = A.c + 10;
    // the whole grid in its loop-axis view (the kernel indexes by loop axis)
    Grid G = grid();
    finalize_grid(G);
    for (int k = 0; k < 3; ++k) {
      const int n = gv.has(gv.ydir(k)) ? gv.num[gv.ydir(k)] : 0;
      if (A.lo[k] < 0 || A.hi[k] > n || G.an[k] != n + 1 || G.ast[k] != (gv.has(gv.ydir(k)) ? gv.st[gv.ydir(k)] : 0))
        fail("internal: design box outside the grid");
    }
    launch_record<R>(dst, f[dc][0], G, A.lo, A.hi, stream);
    A.nrec += 1;
  }
  bool seg_active = false;
  long seg_first = 0;
  double adj_objective() override {
    if (!adj) fail("no reverse-mode recording");
    double J = 0;
    for (DftChunk<R> *ch : adj->mon->E) {
      const std::vector<cdouble> v = read_dft(*ch);
      for (const cdouble &z : v) J += std::norm(z);
    }
    return J;
  }
  void adj_gradient(double *out) override {
    if (!adj) fail("no reverse-mode recording");
    Adj &A = *adj;
    HIPCHECK(hipStreamSynchronize(stream));
    if (!plan_valid) rebuild_plan();
    const long N = A.nrec;
    // cotangent arrays, aliased as the forward arrays are
    std::vector<StateArr> fwd = state_arrays();
    std::vector<std::unique_ptr<DevBuf>> barbuf;
    std::map<const void *, R *> bar; // cotangents in realnum, of the FP64 auxiliaries too
    for (const StateArr &a : fwd) {
      barbuf.emplace_back(new DevBuf);
      barbuf.back()->alloc(sizeof(R) * gv.ntot);
      HIPCHECK(hipMemsetAsync(barbuf.back()->p, 0, sizeof(R) * gv.ntot, stream));
      bar[a.p] = static_cast<R *>(barbuf.back()->p);
    }
    auto B = [&](const void *p) -> R * { return p ? bar.at(p) : nullptr; };
    DevBuf cbar[3];
    for (int q = 0; q < 3; ++q) {
      cbar[q].alloc(sizeof(R) * gv.ntot);
      HIPCHECK(hipMemsetAsync(cbar[q].p, 0, sizeof(R) * gv.ntot, stream));
    }
    // argument blocks of both phases over the whole grid
    int whichB[3], whichD[3];
    StepArgs<R> AB = base_args(), AD = base_args();
    AB.nc = curl_comps(B_ft, 0, AB.cc, whichB);
    upd_comps(H_ft, 0, whichB, AB.nc, AB.uc, false);
    AD.nc = curl_comps(D_ft, 0, AD.cc, whichD);
    upd_comps(E_ft, 0, whichD, AD.nc, AD.uc, false);
    if (!plan_valid) fail("internal: the reverse sweep found arrays still to allocate");
    int full_lo[3], full_hi[3];
    for (int d = 0; d < 3; ++d) {
      full_lo[d] = 0;
      full_hi[d] = gv.has(d) ? gv.num[d] : 0;
    }
    for (StepArgs<R> *S : {&AB, &AD}) {
      for (int d = 0; d < 3; ++d) {
        S->g.blo[d] = full_lo[d];
        S->g.bhi[d] = full_hi[d];
      }
      finalize_args(*S);
    }
    check_args(AB, true, true, false);
    check_args(AD, true, true, false);
    auto slots = [&](const StepArgs<R> &S) {
      AdjSlots<R> b;
      memset(&b, 0, sizeof b);
      for (int q = 0; q < 3; ++q) {
        b.fb[q] = S.cc[q].f ? B(S.cc[q].f) : nullptr;
        b.fub[q] = B(S.cc[q].fu);
        b.cbar[q] = static_cast<R *>(cbar[q].p);
      }
      return b;
    };
    auto uslots = [&](const StepArgs<R> &S) {
      AdjSlots<R> b;
      memset(&b, 0, sizeof b);
      for (int q = 0; q < 3; ++q)
        if (S.uc[q].f) {
          b.fb[q] = B(S.uc[q].f);
          b.gb[q] = B(S.uc[q].g);
          b.fwb[q] = B(S.uc[q].fw);
        }
      return b;
    };
    const AdjSlots<R> cB = slots(AB), cD = slots(AD), uH = uslots(AB), uE = uslots(AD);
    // partner gathers: the curl of slot q reads g1 at i + s1 and i, g2 at i and i + s2
    auto gathers = [&](const StepArgs<R> &S) {
      std::vector<AdjGather<R>> out;
      auto entry = [&](const R *g, int q, int64_t s, int plus) {
        AdjGather<R> *ga = nullptr;
        for (auto &x : out)
          if (x.gb == B(g)) ga = &x;
        if (!ga) {
          out.emplace_back();
          memset(&out.back(), 0, sizeof(AdjGather<R>));
          out.back().gb = B(g);
          ga = &out.back();
        }
        if (ga->n >= 4) fail("internal: too many gather entries");
        int ax = -1, step = 0;
        for (int k = 0; k < 3; ++k) {
          const int d = gv.ydir(k);
          if (gv.has(d) && gv.st[d] == s) ax = k, step = 1;
          if (gv.has(d) && gv.st[d] == -s) ax = k, step = -1;
        }
        if (ax < 0) fail("internal: curl stride matches no axis");
        const int e = ga->n++;
        ga->cbar[e] = static_cast<const R *>(cbar[q].p);
        ga->s[e] = s;
        ga->ax[e] = ax;
        ga->step[e] = step;
        ga->plus[e] = plus;
      };
      for (int q = 0; q < 3; ++q) {
        const CurlComp<R> &k = S.cc[q];
        if (!k.f) continue;
        if (k.g1) entry(k.g1, q, k.s1, 1);
        if (k.g2) entry(k.g2, q, k.s2, 0);
      }
      return out;
    };
    const std::vector<AdjGather<R>> gD = gathers(AD), gB = gathers(AB);
    // halo transposes (every halo point, folded planes included), by field-type group
    AdjHalo hal[2];
    build_adj_halo(hal, bar);
    // objective accumulators and per-chunk scratch
    struct Obj {
      DftChunk<R> *ch;
      DevBuf acc, G;
    };
    std::vector<std::unique_ptr<Obj>> obj;
    for (DftChunk<R> *ch : A.mon->E) {
      std::unique_ptr<Obj> o(new Obj);
      o->ch = ch;
      const std::vector<cdouble> v = read_dft(*ch);
      o->acc.alloc(sizeof(double2) * v.size());
      HIPCHECK(hipMemcpy(o->acc.p, v.data(), sizeof(double2) * v.size(), hipMemcpyHostToDevice));
      o->G.alloc(sizeof(double) * std::max(ch->npts, 1));
      obj.push_back(std::move(o));
    }
    DevBuf grad;
    grad.alloc(sizeof(double) * A.npts);
    HIPCHECK(hipMemsetAsync(grad.p, 0, sizeof(double) * A.npts, stream));
    Grid gfull = AD.g;
    int dslot = -1;
    for (int q = 0; q < 3; ++q)
      if (whichD[q] == A.c + 10) dslot = q;
    if (dslot < 0 || !AD.uc[dslot].f) fail("internal: design component has no E update");
    const int64_t lo_dev = 0;
    (void)lo_dev;
    // one reverse step for forward step n (t before it: t0 + n); hist points at D after it
    auto reverse_step = [&](long n, const R *hist) {
      const long tafter = A.t0 + n + 1;
      for (auto &o : obj) {
        DftChunk<R> &ch = *o->ch;
        if (tafter % ch.decimation) continue;
        const double time = (is_magnetic(ch.c) || is_B(ch.c)) ? tafter * dt - 0.5 * dt : tafter * dt;
        launch_adj_dft<R>(static_cast<const int64_t *>(ch.idx.p), static_cast<const double *>(ch.w.p),
                          static_cast<const double2 *>(o->acc.p), static_cast<const double *>(ch.omega.p), ch.npts, ch.nfreq, time,
                          ch.scale.real(), ch.scale.imag(), static_cast<double *>(o->G.p), stream);
        R *eb = B(f[ch.c][0]);
        const int64_t offs[4] = {0, ch.avg1, ch.avg2, ch.avg1 + ch.avg2};
        const int nof = ch.avg2 ? 4 : ch.avg1 ? 2 : 1;
        const double coef = nof == 4 ? 0.25 : nof == 2 ? 0.5 : 1.0;
        for (int q = 0; q < nof; ++q)
          launch_adj_dft_scatter<R>(eb, static_cast<const int64_t *>(ch.idx.p), static_cast<const double *>(o->G.p), ch.npts, offs[q],
                                    coef, stream);
      }
      run_adj_halo(hal[1]);
      AdjDesign<R> dz;
      memset(&dz, 0, sizeof dz);
      dz.slot = dslot;
      for (int k = 0; k < 3; ++k) dz.lo[k] = A.lo[k], dz.hi[k] = A.hi[k];
      dz.hist = hist;
      dz.grad = static_cast<double *>(grad.p);
      launch_adj_upd<R>(AD, uE, dz, stream);
      launch_adj_curl<R>(AD, cD, stream);
      for (const AdjGather<R> &ga : gD) launch_adj_gather<R>(gfull, ga, stream);
      run_adj_halo(hal[0]);
      bool any_h = false;
      for (int q = 0; q < 3; ++q) any_h = any_h || AB.uc[q].f;
      if (any_h) {
        AdjDesign<R> none;
        memset(&none, 0, sizeof none);
        none.slot = -1;
        launch_adj_upd<R>(AB, uH, none, stream);
      }
      launch_adj_curl<R>(AB, cB, stream);
      for (const AdjGather<R> &ga : gB) launch_adj_gather<R>(gfull, ga, stream);
    };
    if (A.interval <= 0) {
      for (long n = N - 1; n >= 0; --n)
        reverse_step(n, static_cast<const R *>(A.hist[n / HBLK]->p) + size_t(n % HBLK) * A.npts);
    }
    else {
      // keep the final state, recompute each segment from its checkpoint, sweep it backwards
      HIPCHECK(hipStreamSynchronize(stream));
      const long tend = t;
      typename Adj::Checkpoint fin;
      fin.t = t;
      for (const StateArr &a : fwd) {
        fin.data.emplace_back(new DevBuf);
        fin.data.back()->alloc(a.esz * gv.ntot);
        HIPCHECK(hipMemcpyAsync(fin.data.back()->p, a.p, a.esz * gv.ntot, hipMemcpyDeviceToDevice, stream));
      }
      A.seg.alloc(sizeof(R) * A.npts * size_t(A.interval));
      const long nrec_total = N;
      for (long sidx = long(A.cps.size()) - 1; sidx >= 0; --sidx) {
        const long first = A.cps[sidx].t - A.t0;
        if (first >= nrec_total) continue;
        const long last = std::min(nrec_total, first + A.interval); // steps first .. last-1
        restore_state(A.cps[sidx], fwd);
        t = A.cps[sidx].t;
        seg_active = true;
        seg_first = first;
        A.nrec = first;
        dft_off = true;
        A.recording = true;
        for (long q = first; q < last; ++q) {
          rows_hint = last - q;
          one_step(false);
        }
        dft_off = false;
        seg_active = false;
        for (long n = last - 1; n >= first; --n)
          reverse_step(n, static_cast<const R *>(A.seg.p) + size_t(n - first) * A.npts);
      }
      restore_state(fin, fwd);
      t = tend;
      A.nrec = nrec_total;
    }
    HIPCHECK(hipStreamSynchronize(stream));
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpy(out, grad.p, sizeof(double) * A.npts, hipMemcpyDeviceToHost));
  }
  void restore_state(const typename Adj::Checkpoint &cp, const std::vector<StateArr> &arrays) {
    if (cp.data.empty()) {
      for (const StateArr &a : arrays) HIPCHECK(hipMemsetAsync(a.p, 0, a.esz * gv.ntot, stream));
      return;
    }
    if (cp.data.size() != arrays.size()) fail("internal: the state arrays changed since the checkpoint");
    for (size_t q = 0; q < arrays.size(); ++q)
      HIPCHECK(hipMemcpyAsync(arrays[q].p, cp.data[q]->p, arrays[q].esz * gv.ntot, hipMemcpyDeviceToDevice, stream));
  }
  // Halo transposes: every halo point of every allocated array (the kernels' folded planes
  // included), grouped by owned source index so one thread owns each source.
  struct AdjHalo {
    int nsrc = 0;
    DevBuf first, src, dst, sgn, slot, bars;
  };
  void build_adj_halo(AdjHalo hal[2], const std::map<const void *, R *> &bar) {
    struct E {
      uint32_t src, dst;
      int8_t sgn;
      uint16_t slot;
    };
    std::vector<E> lists[2];
    std::vector<R *> bars;
    std::map<R *, uint16_t> slot_of;
    bool B_redundant[5];
    for (int q = 0; q < 5; ++q) B_redundant[q] = f[5 + q][0] == f[15 + q][0];
    for (int corig = 0; corig < NC; ++corig) {
      if (!f[corig][0]) continue;
      if (is_B(corig) && B_redundant[corig - 15]) continue;
      R *bp = bar.at(f[corig][0]);
      if (!slot_of.count(bp)) {
        slot_of[bp] = uint16_t(bars.size());
        bars.push_back(bp);
      }
      int lo[3], hi[3];
      for (int d = 0; d < 3; ++d) {
        lo[d] = 0;
        hi[d] = gv.has(d) ? gv.num[d] : 0;
      }
      for (int d = 0; d < 3; ++d) {
        if (!gv.has(d)) continue;
        const int plane = gv.ysh(corig, d) ? gv.num[d] : 0;
        int blo[3], bhi[3];
        for (int q = 0; q < 3; ++q) blo[q] = lo[q], bhi[q] = hi[q];
        blo[d] = bhi[d] = plane;
        for (int j0 = blo[0]; j0 <= bhi[0]; ++j0)
          for (int j1 = blo[1]; j1 <= bhi[1]; ++j1)
            for (int j2 = blo[2]; j2 <= bhi[2]; ++j2) {
              int j[3] = {j0, j1, j2};
              IV here = gv.iloc(corig, j);
              const int64_t n = gv.index(corig, here);
              cdouble ph = 1.0;
              if (!locate(&here, &ph) || on_metal_boundary(here)) continue;
              if (std::fabs(ph.imag()) > 1e-13 || std::fabs(std::fabs(ph.real()) - 1.0) > 1e-13)
                fail("the reverse mode needs real Bloch phases");
              const int ft = field_type(corig);
              const int grp = (ft == 1 || ft == 3) ? 0 : 1;
              lists[grp].push_back({uint32_t(gv.index(corig, here)), uint32_t(n), int8_t(ph.real() > 0 ? 1 : -1), slot_of[bp]});
            }
        if (plane == 0)
          lo[d] = 1;
        else
          hi[d] = plane - 1;
      }
    }
    for (int g = 0; g < 2; ++g) {
      std::vector<E> &L = lists[g];
      std::stable_sort(L.begin(), L.end(), [](const E &x, const E &y) { return x.src < y.src; });
      std::vector<int> first;
      std::vector<uint32_t> src, dst;
      std::vector<int8_t> sg;
      std::vector<uint16_t> sl;
      for (size_t q = 0; q < L.size(); ++q) {
        if (q == 0 || L[q].src != L[q - 1].src) {
          first.push_back(int(q));
          src.push_back(L[q].src);
        }
        dst.push_back(L[q].dst);
        sg.push_back(L[q].sgn);
        sl.push_back(L[q].slot);
        if (size_t(L[q].dst) >= gv.ntot || size_t(L[q].src) >= gv.ntot) fail("internal: halo transpose outside the array");
      }
      first.push_back(int(L.size()));
      AdjHalo &h = hal[g];
      h.nsrc = int(src.size());
      auto up = [&](DevBuf &b, const void *d, size_t n) {
        b.alloc(n ? n : 1);
        if (n) HIPCHECK(hipMemcpy(b.p, d, n, hipMemcpyHostToDevice));
      };
      up(h.first, first.data(), sizeof(int) * first.size());
      up(h.src, src.data(), sizeof(uint32_t) * src.size());
      up(h.dst, dst.data(), sizeof(uint32_t) * dst.size());
      up(h.sgn, sg.data(), sg.size());
      up(h.slot, sl.data(), sizeof(uint16_t) * sl.size());
      up(h.bars, bars.data(), sizeof(R *) * bars.size());
    }
  }
  void run_adj_halo(AdjHalo &h) {
    launch_adj_halo<R>(static_cast<const int *>(h.first.p), static_cast<const uint32_t *>(h.src.p), static_cast<const uint32_t *>(h.dst.p),
                       static_cast<const int8_t *>(h.sgn.p), static_cast<const uint16_t *>(h.slot.p), h.nsrc,
                       static_cast<R *const *>(h.bars.p), stream);
  }

  void step_n(long n, bool) override {
    if (batch_split) { // stepped on its own again: the single-chunk layout
      batch_split = false;
      boxes_valid = false;
    }
    if (!plan_valid) rebuild_plan();
    auto t0 = std::chrono::steady_clock::now();
    HIPCHECK(hipStreamSynchronize(stream));
    keep.clear(); // nothing queued reads them any more
    rows_free = true;
    // the reference checks after every step (src/step.cpp:137-138); a batch probes its last step
    dft_defer_steps = n > 1;
    for (long q = 0; q < n; ++q) {
      rows_hint = n - q;
      one_step(q == n - 1);
    }
    flush_dfts(); // every deferred DFT step reaches the accumulators before the call returns
    dft_defer_steps = false;
    HIPCHECK(hipStreamSynchronize(stream));
    HIPCHECK(hipGetLastError());
    timers[6] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (ktiming) {
      resolve_events();
      timers[0] = k_ms[MP_K_STEP_BH] * 1e-3;
      timers[2] = k_ms[MP_K_STEP_DE] * 1e-3;
      timers[4] = k_ms[MP_K_HALO] * 1e-3;
      timers[5] = k_ms[MP_K_DFT] * 1e-3;
    }
    int flag = 0;
    HIPCHECK(hipMemcpy(&flag, d_flag.p, sizeof(int), hipMemcpyDeviceToHost));
    if (flag) fail("simulation fields are NaN or Inf");
  }
  // n steps as step_n does them, with no device allocation or release: the per-step source
  // rows go into `rows` (rows_bytes of device memory the caller owns) in pieces of as many
  // steps as fit, each piece after the stream has drained the previous one; the DFT monitors
  // accumulate every step (the deferred form keeps a ring of point values of its own).
  // Everything else the steps use (state arrays, halo tables, source sets, the finiteness
  // flag) must exist already: the fields must have been stepped by step_n since the last
  // change of sources, monitors or boundaries, or an allocation would be needed and the call
  // fails before its first step.
  void step_n_noalloc(long n, double *rows, size_t rows_bytes) override {
    if (n < 0) fail("negative step count");
    if (t == 0 || !plan_valid || srcplan_stale() || batch_split || !boxes_valid)
      fail("mp_fields_step_n_noalloc: step these fields with mp_fields_step_n first (their launch plan is "
           "built, and their arrays allocated, by a step that may allocate)");
    if (ktiming) fail("mp_fields_step_n_noalloc: kernel timing creates events; switch it off");
    if (adj && adj->recording) fail("mp_fields_step_n_noalloc: the reverse mode records history; use mp_fields_step_n");
    for (auto &cp : dft_chunks)
      if (cp->nsnap) fail("internal: deferred DFT steps pending");
    const size_t rl = splan.rec.size();
    const long per = rl ? long(rows_bytes / (sizeof(double) * rl)) : n;
    if (rl && per < 1)
      fail("mp_fields_step_n_noalloc: one step's source row (%zu values) exceeds the %zu-byte scratch array", rl,
           rows_bytes);
    auto t0 = std::chrono::steady_clock::now();
    struct Guard {
      Fields *F;
      Guard(Fields *f) : F(f) { g_no_alloc = true; }
      ~Guard() {
        g_no_alloc = false;
        F->ext_rows = nullptr;
        F->ext_n = 0;
      }
    } guard_(this);
    HIPCHECK(hipStreamSynchronize(stream));
    dft_defer_steps = false;
    std::vector<double> host;
    for (long q0 = 0; q0 < n;) {
      const long m = std::min(per, n - q0);
      if (rl) {
        host.resize(rl * size_t(m));
        for (long st = 0; st < m; ++st)
          for (size_t q = 0; q < rl; ++q) host[size_t(st) * rl + q] = recipe_value(splan.rec[q], t + st);
        // the previous piece's launches have finished reading the staging array (and the
        // host vector is reused), so the stream is drained first
        HIPCHECK(hipStreamSynchronize(stream));
        HIPCHECK(hipMemcpyAsync(rows, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice, stream));
        ext_rows = rows;
        ext_t0 = t;
        ext_n = m;
      }
      for (long q = 0; q < m; ++q) one_step(q0 + q == n - 1);
      q0 += m;
    }
    HIPCHECK(hipStreamSynchronize(stream));
    HIPCHECK(hipGetLastError());
    timers[6] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    int flag = 0;
    HIPCHECK(hipMemcpy(&flag, d_flag.p, sizeof(int), hipMemcpyDeviceToHost));
    if (flag) fail("simulation fields are NaN or Inf");
  }
  int device_id() const override { return device; }
  long t_now() const override { return t; }
  double dt_now() const override { return dt; }
  double last_source_time() const override {
    double after = 0;
    for (auto &st : srcs) after = std::max(after, st.last_time());
    return after;
  }

  // ---- DFT ----
  // the accumulation arguments of chunk ch at the current time
  DftArgs<R> dft_args(const DftChunk<R> &ch) const {
    DftArgs<R> da;
    da.f0 = f[ch.c][0];
    da.f1 = f[ch.c][1];
    da.idx = static_cast<const int64_t *>(ch.idx.p);
    da.w = static_cast<const double *>(ch.w.p);
    da.omega = static_cast<const double *>(ch.omega.p);
    da.time = (is_magnetic(ch.c) || is_B(ch.c)) ? t * dt - 0.5 * dt : t * dt;
    da.scale_re = ch.scale.real();
    da.scale_im = ch.scale.imag();
    da.avg1 = ch.avg1;
    da.avg2 = ch.avg2;
    da.npts = ch.npts;
    da.nfreq = ch.nfreq;
    da.acc = ch.acc.p;
    da.comp = ch.comp.p;
    da.acc_kind = ch.acc_kind;
    return da;
  }
  bool dft_defer_steps = false; // set by step_n for calls of several steps
  // DFT accumulation of this step (dft_chunk::update_dft for every chunk, src/dft.cpp:250-308):
  // chunks whose phases for DFT_DEFER steps fit the flush kernel's table keep this step's
  // point values and accumulate every DFT_DEFER steps (and at the end of every step_n call)
  void update_dfts() {
    for (auto &cp : dft_chunks) {
      DftChunk<R> &ch = *cp;
      if (t % ch.decimation) continue;
      if (!f[ch.c][0]) continue;
      const DftArgs<R> da = dft_args(ch);
      kbegin(MP_K_DFT);
      if (ch.defer > 1 && dft_defer_steps) {
        const size_t plane = size_t(ch.npts);
        const size_t need = sizeof(R) * plane * DFT_DEFER * (da.f1 ? 2 : 1);
        if (ch.snap.bytes < need) {
          if (ch.nsnap) fail("internal: deferred DFT storage changed with steps pending");
          ch.snap.alloc(need);
        }
        R *base = static_cast<R *>(ch.snap.p);
        R *s0 = base + size_t(ch.nsnap) * plane;
        R *s1 = da.f1 ? base + size_t(DFT_DEFER + ch.nsnap) * plane : nullptr;
        launch_dft_snap<R>(da, s0, s1, stream);
        ch.snap_time[ch.nsnap++] = da.time;
        if (ch.nsnap == ch.defer) flush_dft(ch);
      }
      else
        launch_dft<R>(da, stream);
      kend();
    }
  }
  void flush_dft(DftChunk<R> &ch) {
    if (!ch.nsnap) return;
    const DftArgs<R> da = dft_args(ch);
    DftDefer<R> d;
    memset(&d, 0, sizeof d);
    const R *base = static_cast<const R *>(ch.snap.p);
    d.snap0 = base;
    d.snap1 = da.f1 ? base + size_t(DFT_DEFER) * size_t(ch.npts) : nullptr;
    d.nsteps = ch.nsnap;
    for (int j = 0; j < ch.nsnap; ++j) d.time[j] = ch.snap_time[j];
    launch_dft_flush<R>(da, d, stream);
    ch.nsnap = 0;
  }
  void flush_dfts() {
    for (auto &cp : dft_chunks) flush_dft(*cp);
  }

  // loop_in_chunks for one chunk and the identity symmetry (src/loop_in_chunks.cpp:232-540)
  template <class CB> void loop_in_chunks(const V3 &wmin, const V3 &wmax, int cgrid, CB chunkloop) {
    IV iyee_c;
    for (int d = 0; d < 3; ++d) iyee_c.t[d] = gv.ysh(20, d) - gv.ysh(cgrid, d);
    V3 yc = gv.loc(iyee_c);
    IV is, ie;
    for (int d = 0; d < 3; ++d)
      if (gv.has(d)) {
        is.t[d] = 1 + 2 * int(std::floor((wmin.t[d] + yc.t[d]) * a - .5)) - iyee_c.t[d];
        ie.t[d] = 1 + 2 * int(std::ceil((wmax.t[d] + yc.t[d]) * a - .5)) - iyee_c.t[d];
      }
    V3 s0, s1, e0, e1;
    for (int d = 0; d < 3; ++d) {
      if (!gv.has(d)) continue;
      double w0 = 1. - wmin.t[d] * a + 0.5 * is.t[d];
      double w1 = 1. + wmax.t[d] * a - 0.5 * ie.t[d];
      if (ie.t[d] >= is.t[d] + 3 * 2) {
        s0.t[d] = w0 * w0 / 2;
        s1.t[d] = 1 - (1 - w0) * (1 - w0) / 2;
        e0.t[d] = w1 * w1 / 2;
        e1.t[d] = 1 - (1 - w1) * (1 - w1) / 2;
      }
      else if (ie.t[d] == is.t[d] + 2 * 2) {
        s0.t[d] = w0 * w0 / 2;
        s1.t[d] = 1 - (1 - w0) * (1 - w0) / 2 - (1 - w1) * (1 - w1) / 2;
        e0.t[d] = w1 * w1 / 2;
        e1.t[d] = s1.t[d];
      }
      else if (wmin.t[d] == wmax.t[d]) {
        s0.t[d] = w0;
        s1.t[d] = w1;
        e0.t[d] = w1;
        e1.t[d] = w0;
      }
      else if (ie.t[d] == is.t[d] + 1 * 2) {
        s0.t[d] = w0 * w0 / 2 - (1 - w1) * (1 - w1) / 2;
        e0.t[d] = w1 * w1 / 2 - (1 - w0) * (1 - w0) / 2;
        s1.t[d] = e0.t[d];
        e1.t[d] = s0.t[d];
      }
      else
        fail("bug: impossible looping boundaries");
    }
    IV mins, maxs;
    for (int d = 0; d < 3; ++d)
      if (gv.has(d) && bc[High][d] == Periodic) {
        const double L = gv.num[d] * gv.inva;
        mins.t[d] = int(std::floor((wmin.t[d] - gv.num[d] * gv.inva) / L));
        maxs.t[d] = int(std::ceil((wmax.t[d] - 0.0) / L));
      }
    IV sh = mins;
    IV isco, ieco;
    for (int d = 0; d < 3; ++d)
      if (gv.has(d)) {
        isco.t[d] = 2 - gv.ysh(cgrid, d);
        ieco.t[d] = 2 * gv.num[d] - gv.ysh(cgrid, d);
      }
    while (true) {
      cdouble ph = 1.0;
      IV shifti;
      for (int d = 0; d < 3; ++d)
        if (gv.has(d)) {
          shifti.t[d] = 2 * gv.num[d] * sh.t[d];
          ph *= std::pow(eikna[d], sh.t[d]);
        }
      IV iscS, iecS;
      bool nonempty = true;
      for (int d = 0; d < 3; ++d)
        if (gv.has(d)) {
          iscS.t[d] = std::max(is.t[d] - shifti.t[d], isco.t[d]);
          iecS.t[d] = std::min(ie.t[d] - shifti.t[d], ieco.t[d]);
          if (iscS.t[d] > iecS.t[d]) nonempty = false;
        }
      if (nonempty) {
        V3 s0c, s1c, e0c, e1c;
        for (int d = 0; d < 3; ++d) s0c.t[d] = s1c.t[d] = e0c.t[d] = e1c.t[d] = 1.0;
        for (int d = 0; d < 3; ++d) {
          if (!gv.has(d)) continue;
          const int a0 = iscS.t[d] + shifti.t[d], b0 = iecS.t[d] + shifti.t[d];
          if (a0 == is.t[d]) {
            s0c.t[d] = s0.t[d];
            s1c.t[d] = s1.t[d];
          }
          else if (a0 == is.t[d] + 2)
            s0c.t[d] = s1.t[d];
          if (b0 == ie.t[d]) {
            e0c.t[d] = e0.t[d];
            e1c.t[d] = e1.t[d];
          }
          else if (b0 == ie.t[d] - 2)
            e0c.t[d] = e1.t[d];
          if (b0 == a0) {
            double w = std::min(s0c.t[d], e0c.t[d]);
            s0c.t[d] = e0c.t[d] = s1c.t[d] = e1c.t[d] = w;
          }
          else if (b0 == a0 + 2) {
            double w = std::min(s0c.t[d], e1c.t[d]);
            s0c.t[d] = e1c.t[d] = w;
            w = std::min(s1c.t[d], e0c.t[d]);
            s1c.t[d] = e0c.t[d] = w;
          }
          else if (b0 == a0 + 4) {
            double w = std::min(s1c.t[d], e1c.t[d]);
            s1c.t[d] = e1c.t[d] = w;
          }
        }
        double dV0 = 1.0;
        for (int d = 0; d < 3; ++d)
          if (gv.has(d) && wmax.t[d] - wmin.t[d] > 0.0) dV0 *= gv.inva;
        chunkloop(iscS, iecS, s0c, s1c, e0c, e1c, dV0, shifti, ph);
      }
      bool wrapped = true;
      for (int d = 0; d < 3; ++d) {
        if (!gv.has(d)) continue;
        if (sh.t[d] + 1 <= maxs.t[d]) {
          sh.t[d] += 1;
          wrapped = false;
          break;
        }
        sh.t[d] = mins.t[d];
      }
      if (wrapped) break;
    }
  }
  // Visit the points of box [is, ie] (half-pixel coordinates, step 2) in loop order, giving
  // the array index, the counter and the integration weight (IVEC_LOOP_WEIGHT,
  // src/meep/vec.hpp:338-350).
  template <class CB>
  void visit_box(const IV &is, const IV &ie, const V3 &s0, const V3 &s1, const V3 &e0, const V3 &e1, double dV, CB cb) const {
    int n[3], d_[3];
    for (int kx = 0; kx < 3; ++kx) {
      int d = gv.ydir(kx);
      d_[kx] = gv.has(d) ? d : -1;
      n[kx] = gv.has(d) ? (ie.t[d] - is.t[d]) / 2 + 1 : 1;
    }
    auto w1 = [&](int i, int nn, int d) {
      if (d < 0) return 1.0;
      return (i > 1 && i < nn - 2) ? 1.0 : (i == 0 ? s0.t[d] : (i == 1 ? s1.t[d] : i == nn - 1 ? e0.t[d] : (i == nn - 2 ? e1.t[d] : 1.0)));
    };
    int64_t counter = 0;
    for (int i1 = 0; i1 < n[0]; ++i1)
      for (int i2 = 0; i2 < n[1]; ++i2)
        for (int i3 = 0; i3 < n[2]; ++i3, ++counter) {
          IV p;
          int ii[3] = {i1, i2, i3};
          for (int kx = 0; kx < 3; ++kx)
            if (d_[kx] >= 0) p.t[d_[kx]] = is.t[d_[kx]] + 2 * ii[kx];
          int64_t idx = 0;
          for (int d = 0; d < 3; ++d)
            if (gv.has(d)) idx += p.t[d] / 2 * gv.st[d];
          double w = w1(i3, n[2], d_[2]) * (w1(i2, n[1], d_[1]) * (dV * w1(i1, n[0], d_[0])));
          cb(idx, counter, p, w, ii);
        }
  }

  int add_src_time(const SrcTime &st) {
    for (size_t q = 0; q < srcs.size(); ++q)
      if (srcs[q].equal(st)) return int(q);
    srcs.push_back(st);
    return int(srcs.size()) - 1;
  }
  // fields::add_volume_source + src_vol_chunkloop (src/sources.cpp:243-312,462-490)
  void add_volume_source(int c, const SrcTime &st, V3 wmin, V3 wmax, cdouble amp, mp_complex_func A, void *Au) override {
    if (c < 0 || c >= NC || comp_dir(c) == NODIR || !has_field(c)) fail("invalid source component %d", c);
    for (int d = 0; d < 3; ++d) {
      if (!gv.has(d)) continue;
      const double w = gv.num[d] * gv.inva, wd = wmax.t[d] - wmin.t[d];
      if (wd > w + gv.inva)
        fail("Source width > cell width");
      else if (wd > w) {
        wmin.t[d] -= (wd - w) * 0.5;
        wmax.t[d] = wmin.t[d] + w;
      }
    }
    cdouble damp = amp;
    for (int d = 0; d < 3; ++d)
      if (gv.has(d) && wmax.t[d] - wmin.t[d] == 0.0) damp *= a;
    const int isrc = add_src_time(st);
    V3 cen;
    for (int d = 0; d < 3; ++d) cen.t[d] = (wmin.t[d] + wmax.t[d]) * 0.5;
    loop_in_chunks(wmin, wmax, c, [&](const IV &is, const IV &ie, const V3 &s0, const V3 &s1, const V3 &e0, const V3 &e1, double, const IV &shift, cdouble sph) {
      std::vector<int64_t> index;
      std::vector<cdouble> amps;
      std::vector<int> jpos;
      const cdouble am = damp * std::conj(sph);
      visit_box(is, ie, s0, s1, e0, e1, 1.0, [&](int64_t idx, int64_t, const IV &p, double w, const int *) {
        if (!gv.owns(p)) return;
        double rel[3];
        for (int d = 0; d < 3; ++d) rel[d] = gv.has(d) ? (p.t[d] * 0.5 * gv.inva + shift.t[d] * (0.5 * gv.inva)) - cen.t[d] : -cen.t[d];
        cdouble Av = 1.0;
        if (A) {
          double re = 1, im = 0;
          A(rel, Au, &re, &im);
          Av = cdouble(re, im);
        }
        amps.push_back(w * am * Av);
        index.push_back(idx);
        for (int d = 0; d < 3; ++d) jpos.push_back(gv.has(d) ? (p.t[d] - gv.ysh(c, d)) / 2 : 0);
      });
      for (int64_t i : index)
        if (i < 0 || size_t(i) >= gv.ntot) fail("internal: source point outside the array");
      const int ft = (is_magnetic(c) || is_B(c)) ? B_ft : D_ft;
      const int cc = (is_D(c) || is_B(c)) ? c - 10 : c;
      // a current in a conducting medium is scaled by condinv at its point
      // (fields_chunk::step_source, src/step.cpp:299-309); integrated sources are not
      const int cD = dir_comp(ft == B_ft ? 15 : 10, comp_dir(c));
      if (!srcs[isrc].is_integrated && !h_condinv[cD].empty())
        for (size_t q = 0; q < amps.size(); ++q) {
          const int j[3] = {jpos[3 * q], jpos[3 * q + 1], jpos[3 * q + 2]};
          amps[q] *= h_condinv[cD][gv.ref_of(j)];
        }
      ++src_version;
      for (SrcVol &o : sources[ft])
        if (o.c == cc && o.src == isrc && o.index == index) {
          for (size_t q = 0; q < o.amp.size(); ++q) o.amp[q] += amps[q];
          o.dirty = true;
          return;
        }
      SrcVol sv;
      sv.c = cc;
      sv.src = isrc;
      sv.index = std::move(index);
      sv.amp = std::move(amps);
      sv.jpos = std::move(jpos);
      sources[ft].push_back(std::move(sv));
    });
    require_component(c);
  }

  // a source list of component c at the points with array coordinates j3 (3 per point), as
  // add_volume_source stores one: device indices, a current in a conducting medium scaled by
  // condinv at its point
  void add_points(int c, int isrc, const std::vector<int> &j3, std::vector<cdouble> amps) {
    const int ft = is_magnetic(c) ? B_ft : D_ft;
    const int cD = dir_comp(ft == B_ft ? 15 : 10, comp_dir(c));
    SrcVol sv;
    sv.c = c;
    sv.src = isrc;
    for (size_t q = 0; q < amps.size(); ++q) {
      const int j[3] = {j3[3 * q], j3[3 * q + 1], j3[3 * q + 2]};
      sv.index.push_back(gv.dev_of(j));
      for (int d = 0; d < 3; ++d) sv.jpos.push_back(j[d]);
      if (!srcs[isrc].is_integrated && !h_condinv[cD].empty()) amps[q] *= h_condinv[cD][gv.ref_of(j)];
    }
    sv.amp = std::move(amps);
    ++src_version;
    sources[ft].push_back(std::move(sv));
  }
  // fields::add_srcdata (src/sources.cpp:314-334): a source on points of component c's array
  // (reference-layout indices) with the amplitudes as given.  A point the chunk does not own
  // waits for require_source_components when the list is marked for the boundary fix, and is
  // dropped otherwise: the halo exchange after the source would overwrite it.
  struct PendingFix {
    int c, src;
    std::vector<int> j3;
    std::vector<cdouble> amp;
  };
  std::vector<PendingFix> pending_fix;
  void array_coords(int64_t ref, int j[3]) const {
    for (int d = 0; d < 3; ++d) j[d] = 0;
    int order[3] = {0, 1, 2};
    std::sort(order, order + 3, [&](int u, int v) { return gv.rst[u] > gv.rst[v]; });
    for (int d : order)
      if (gv.has(d)) {
        j[d] = int(ref / gv.rst[d]);
        ref %= gv.rst[d];
      }
  }
  void add_srcdata(int c, const SrcTime &st, const long *ref, long n, const double *amp, bool fix) override {
    if (c < 0 || c >= NC || !(is_electric(c) || is_magnetic(c)) || comp_dir(c) == NODIR || !has_field(c))
      fail("add_srcdata: invalid component %d", c);
    if (n < 0) fail("add_srcdata: negative point count");
    const int isrc = add_src_time(st);
    std::vector<int> jown;
    std::vector<cdouble> aown;
    PendingFix pf{c, isrc, {}, {}};
    for (long q = 0; q < n; ++q) {
      if (ref[q] < 0 || size_t(ref[q]) >= gv.rntot) fail("add_srcdata: point outside the array");
      int j[3];
      array_coords(ref[q], j);
      const cdouble A(amp[2 * q], amp[2 * q + 1]);
      if (gv.owns(gv.iloc(c, j))) {
        jown.insert(jown.end(), j, j + 3);
        aown.push_back(A);
      }
      else if (fix) {
        pf.j3.insert(pf.j3.end(), j, j + 3);
        pf.amp.push_back(A);
      }
    }
    if (!aown.empty()) add_points(c, isrc, jown, std::move(aown));
    if (!pf.amp.empty()) pending_fix.push_back(std::move(pf));
  }
  // fields::require_source_components (src/fields.cpp:508-530) with fix_boundary_sources
  // (src/fix_boundary_sources.cpp:51-197): each waiting point with a nonzero amplitude moves to
  // the point that owns it (its periodic image, the amplitude times the conjugate phase)
  // unless that point is on a metal boundary; the moved points form one list per source time
  // and component; every source component is allocated
  int dimension() const override { return gv.dim; }
  void require_source_components() override {
    struct Moved {
      int src, c;
      int j[3];
      cdouble A;
    };
    std::vector<Moved> moved;
    for (const PendingFix &pf : pending_fix)
      for (size_t q = 0; q < pf.amp.size(); ++q) {
        if (pf.amp[q] == 0.0) continue;
        IV here = gv.iloc(pf.c, &pf.j3[3 * q]);
        cdouble ph = 1.0;
        if (!locate(&here, &ph) || on_metal_boundary(here)) continue;
        Moved m;
        m.src = pf.src;
        m.c = pf.c;
        for (int d = 0; d < 3; ++d) m.j[d] = gv.has(d) ? (here.t[d] - gv.ysh(pf.c, d)) / 2 : 0;
        m.A = pf.amp[q] * std::conj(ph);
        moved.push_back(m);
      }
    pending_fix.clear();
    std::stable_sort(moved.begin(), moved.end(), [](const Moved &a, const Moved &b) { return a.src != b.src ? a.src < b.src : a.c < b.c; });
    for (size_t q0 = 0; q0 < moved.size();) {
      std::vector<int> j3;
      std::vector<cdouble> amps;
      size_t q1 = q0;
      for (; q1 < moved.size() && moved[q1].src == moved[q0].src && moved[q1].c == moved[q0].c; ++q1) {
        j3.insert(j3.end(), moved[q1].j, moved[q1].j + 3);
        amps.push_back(moved[q1].A);
      }
      add_points(moved[q0].c, moved[q0].src, j3, std::move(amps));
      q0 = q1;
    }
    for (int ft : {D_ft, B_ft})
      for (const SrcVol &sv : sources[ft]) require_component(sv.c);
  }

  int auto_decimation(const std::vector<double> &freq) const {
    double src_freq_max = 0;
    for (const SrcTime &st : srcs)
      if (st.get_fwidth() != 0) src_freq_max = std::max(src_freq_max, std::abs(st.freq) + 0.5 * st.get_fwidth());
    double freq_max = 0;
    for (double fr : freq) freq_max = std::max(freq_max, std::abs(fr));
    if (freq_max > 0 && src_freq_max > 0) return std::max(1, int(std::floor(1 / (dt * (freq_max + src_freq_max)))));
    return 1;
  }
  // fields::add_dft for one component (src/dft.cpp:175-235): on the centred grid with the
  // Yee-to-centre average, or on the component's own grid (centered false, no average)
  void add_dft(std::vector<DftChunk<R> *> &out, int c, const V3 &mn, const V3 &mxv, const std::vector<double> &freq, bool include_w,
               cdouble stored_weight, int decimation, bool centered, int vc = -1, bool sqrt_w = false) {
    loop_in_chunks(mn, mxv, centered ? 20 : c, [&](const IV &is, const IV &ie, const V3 &s0, const V3 &s1, const V3 &e0, const V3 &e1, double dV0, const IV &shift,
                                   cdouble sph) {
      if (!f[c][0]) return;
      std::unique_ptr<DftChunk<R>> ch(new DftChunk<R>);
      ch->c = c;
      ch->vc = vc;
      ch->is = is, ch->ie = ie, ch->shift = shift;
      ch->s0 = s0, ch->s1 = s1, ch->e0 = e0, ch->e1 = e1;
      ch->dV0 = dV0;
      ch->weighted = include_w;
      ch->sqrt_w = include_w && sqrt_w;
      ch->stored_weight = stored_weight;
      for (int d = 0; d < 3; ++d) ch->empty[d] = gv.has(d) && mn.t[d] == mxv.t[d];
      if (centered)
        gv.yee2cent(c, ch->avg1, ch->avg2);
      else
        ch->avg1 = ch->avg2 = 0;
      ch->scale = stored_weight * sph * (dt / std::sqrt(2.0 * pi) * decimation);
      ch->decimation = decimation;
      std::vector<int64_t> idx;
      std::vector<double> w;
      visit_box(is, ie, s0, s1, e0, e1, dV0, [&](int64_t i, int64_t, const IV &, double ww, const int *) {
        idx.push_back(i);
        w.push_back(include_w ? (sqrt_w ? std::sqrt(ww) : ww) : 1.0); // src/dft.cpp:277-280
      });
      for (int64_t i : idx)
        if (i < 0 || size_t(i + ch->avg1 + ch->avg2) >= gv.ntot) fail("internal: DFT point outside the array");
      ch->npts = int(idx.size());
      ch->nfreq = int(freq.size());
      std::vector<double> om;
      for (double fr : freq) om.push_back(2 * pi * fr);
      ch->idx.alloc(sizeof(int64_t) * idx.size());
      ch->w.alloc(sizeof(double) * w.size());
      ch->omega.alloc(sizeof(double) * om.size());
      HIPCHECK(hipMemcpy(ch->idx.p, idx.data(), sizeof(int64_t) * idx.size(), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(ch->w.p, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(ch->omega.p, om.data(), sizeof(double) * om.size(), hipMemcpyHostToDevice));
      ch->acc_kind = dft_acc;
      ch->defer = ch->nfreq * DFT_DEFER <= DFT_FLUSH_TAB ? DFT_DEFER : 1;
      const size_t nacc = size_t(ch->npts) * ch->nfreq;
      const size_t esz = dft_acc == MP_DFT_F64 ? 16 : dft_acc == MP_DFT_KAHAN32 ? 8 : 2 * sizeof(R);
      ch->acc.alloc(esz * nacc);
      HIPCHECK(hipMemset(ch->acc.p, 0, esz * nacc));
      if (dft_acc == MP_DFT_KAHAN32) {
        ch->comp.alloc(8 * nacc);
        HIPCHECK(hipMemset(ch->comp.p, 0, 8 * nacc));
      }
      out.push_back(ch.get());
      dft_chunks.push_back(std::move(ch));
      ++dft_version;
    });
  }
  void *add_dft_flux(int d, V3 mn, V3 mxv, const std::vector<double> &freq, bool centered, int dec, int *dec_out) override {
    if (!components_allocated) fail("allocate field components (by adding sources) before adding dft objects");
    const int decim = dec > 0 ? dec : auto_decimation(freq);
    *dec_out = decim;
    int cE[2], cH[2];
    if (d == 0) {
      cE[0] = 1, cE[1] = 4, cH[0] = 9, cH[1] = 6;
    }
    else if (d == 1) {
      cE[0] = 4, cE[1] = 0, cH[0] = 5, cH[1] = 9;
    }
    else {
      cE[0] = 0, cE[1] = 1, cH[0] = 6, cH[1] = 5;
    }
    std::unique_ptr<Flux<R>> fl(new Flux<R>);
    fl->nfreq = int(freq.size());
    fl->wmin = mn, fl->wmax = mxv;
    std::vector<DftChunk<R> *> El[2], Hl[2];
    for (int q = 0; q < 2; ++q) {
      add_dft(El[q], cE[q], mn, mxv, freq, true, cdouble(double(1 - 2 * q), 0), decim, centered);
      add_dft(Hl[q], cH[q], mn, mxv, freq, false, cdouble(1.0, 0), decim, centered);
    }
    // the reference prepends each chunk to its list (src/dft.cpp:120-129,623-641)
    for (int q = 1; q >= 0; --q) {
      for (int j = int(El[q].size()) - 1; j >= 0; --j) fl->E.push_back(El[q][j]);
      for (int j = int(Hl[q].size()) - 1; j >= 0; --j) fl->H.push_back(Hl[q][j]);
    }
    fluxes.push_back(std::move(fl));
    return fluxes.back().get();
  }
  // fields::add_dft_fields (src/dft.cpp:913-927): weights off, stored weight 1, centred grid;
  // each component's chunks go to the front of the list as the reference prepends them
  void *add_dft_fields(const int *comps, int ncomp, V3 mn, V3 mxv, const std::vector<double> &freq, bool centered, int dec,
                       int *dec_out) override {
    if (!components_allocated) fail("allocate field components (by adding sources) before adding dft objects");
    const int decim = dec > 0 ? dec : auto_decimation(freq);
    *dec_out = decim;
    std::unique_ptr<Flux<R>> fl(new Flux<R>);
    fl->nfreq = int(freq.size());
    fl->wmin = mn, fl->wmax = mxv;
    for (int q = 0; q < ncomp; ++q) {
      if (comps[q] < 0 || comps[q] >= NC) fail("invalid dft component %d", comps[q]);
      std::vector<DftChunk<R> *> lst;
      add_dft(lst, comps[q], mn, mxv, freq, false, cdouble(1.0, 0), decim, centered);
      std::vector<DftChunk<R> *> front(lst.rbegin(), lst.rend());
      fl->E.insert(fl->E.begin(), front.begin(), front.end());
    }
    fluxes.push_back(std::move(fl));
    return fluxes.back().get();
  }
  // fields::add_dft_energy (src/dft.cpp:725-750): for every volume and field direction, E and
  // H with integration weights and D and B without, centred grid, each call's chunks in front
  void *add_dft_energy(int nvol, const double *vmins, const double *vmaxs, const std::vector<double> &freq, int dec,
                       int *dec_out) override {
    if (!components_allocated) fail("allocate field components (by adding sources) before adding dft objects");
    if (nvol < 1) fail("add_dft_energy: no volumes");
    const int decim = dec > 0 ? dec : auto_decimation(freq);
    *dec_out = decim;
    std::unique_ptr<Energy<R>> en(new Energy<R>);
    en->nfreq = int(freq.size());
    for (int d = 0; d < 3; ++d) en->wmin.t[d] = vmins[d], en->wmax.t[d] = vmaxs[d];
    std::vector<DftChunk<R> *> *lists[4] = {&en->E, &en->D, &en->H, &en->B};
    const int base[4] = {0, 10, 5, 15};
    for (int w = 0; w < nvol; ++w) {
      V3 mn, mx;
      for (int d = 0; d < 3; ++d) mn.t[d] = vmins[3 * w + d], mx.t[d] = vmaxs[3 * w + d];
      for (int d = 0; d < 3; ++d)
        for (int q = 0; q < 4; ++q) {
          std::vector<DftChunk<R> *> lst;
          add_dft(lst, dir_comp(base[q], d), mn, mx, freq, q == 0 || q == 2, cdouble(1.0, 0), decim, true);
          lists[q]->insert(lists[q]->begin(), lst.rbegin(), lst.rend());
        }
    }
    fluxes.push_back(std::move(en));
    return fluxes.back().get();
  }
  // fields::add_dft_force (src/stress.cpp:153-198): off-diagonal terms (fd != the normal) pair
  // E and H along fd (weighted, times the volume's weight) with E and H along the normal, on
  // the centred grid; diagonal terms take E and H of every direction with square-root weights
  // on their own grids and the readout factor +-0.5 times the weight (+ along fd)
  void *add_dft_force(int nvol, const int *fdirs, const double *vmins, const double *vmaxs, const double *weights,
                      const std::vector<double> &freq, int dec, int *dec_out) override {
    if (!components_allocated) fail("allocate field components (by adding sources) before adding dft objects");
    if (nvol < 1) fail("add_dft_force: no volumes");
    const int decim = dec > 0 ? dec : auto_decimation(freq);
    *dec_out = decim;
    std::unique_ptr<Force<R>> fo(new Force<R>);
    fo->nfreq = int(freq.size());
    for (int d = 0; d < 3; ++d) fo->wmin.t[d] = vmins[d], fo->wmax.t[d] = vmaxs[d];
    auto front = [](std::vector<DftChunk<R> *> &to, std::vector<DftChunk<R> *> &lst) {
      to.insert(to.begin(), lst.rbegin(), lst.rend());
    };
    for (int w = 0; w < nvol; ++w) {
      V3 mn, mx;
      for (int d = 0; d < 3; ++d) mn.t[d] = vmins[3 * w + d], mx.t[d] = vmaxs[3 * w + d];
      const int nd = flux_direction(-1, mn, mx), fd = fdirs[w];
      if (fd < 0 || fd > 2) fail("NO_DIRECTION dft_force is invalid");
      const cdouble wt(weights[2 * w], weights[2 * w + 1]);
      if (fd != nd) {
        for (int base : {0, 5}) {
          std::vector<DftChunk<R> *> l1, l2;
          add_dft(l1, dir_comp(base, fd), mn, mx, freq, true, wt, decim, true);
          add_dft(l2, dir_comp(base, nd), mn, mx, freq, false, cdouble(1.0, 0), decim, true);
          front(fo->E, l1);
          front(fo->H, l2);
        }
      }
      else
        for (int d = 0; d < 3; ++d)
          for (int base : {0, 5}) {
            std::vector<DftChunk<R> *> l;
            add_dft(l, dir_comp(base, d), mn, mx, freq, true, cdouble(1.0, 0), decim, false, -1, true);
            for (DftChunk<R> *ch : l) ch->extra_weight = wt * (d == fd ? 0.5 : -0.5);
            front(fo->D, l);
          }
    }
    fluxes.push_back(std::move(fo));
    return fluxes.back().get();
  }
  // dft_force::force (src/stress.cpp:90-114): Re(extra_weight F1 conj(F2)) over the paired
  // chunks of (offdiag1, offdiag2) and of (diag, diag), the accumulators widened to double
  void force_spectrum(void *fp, double *out) override {
    Force<R> *fo = static_cast<Force<R> *>(fp);
    const int Nf = fo->nfreq;
    for (int i = 0; i < Nf; ++i) out[i] = 0;
    auto sum = [&](const std::vector<DftChunk<R> *> &A, const std::vector<DftChunk<R> *> &B) {
      for (size_t q = 0; q < A.size() && q < B.size(); ++q) {
        if (A[q]->npts != B[q]->npts) fail("internal: dft_force chunks of different sizes");
        const std::vector<cdouble> a = read_dft(*A[q]), b = A[q] == B[q] ? a : read_dft(*B[q]);
        for (size_t k = 0; k < size_t(A[q]->npts); ++k)
          for (int i = 0; i < Nf; ++i) out[i] += std::real(A[q]->extra_weight * a[k * Nf + i] * std::conj(b[k * Nf + i]));
      }
    };
    sum(fo->E, fo->H);
    sum(fo->D, fo->D);
  }
  // dft_energy::electric / magnetic / total (src/dft.cpp:681-723): 0.5 Re(conj(F) G) over the
  // paired chunks, the product in the accumulator's type
  void energy_spectrum(void *ep, int which, double *out) override {
    Energy<R> *en = static_cast<Energy<R> *>(ep);
    const int Nf = en->nfreq;
    std::vector<double> e(Nf, 0.0), m(Nf, 0.0);
    auto part = [&](const std::vector<DftChunk<R> *> &A, const std::vector<DftChunk<R> *> &B, std::vector<double> &acc) {
      for (size_t q = 0; q < A.size() && q < B.size(); ++q) {
        if (A[q]->npts != B[q]->npts) fail("internal: dft_energy chunks of different sizes");
        const std::vector<cdouble> a = read_dft(*A[q]), b = read_dft(*B[q]);
        const bool narrow = A[q]->acc_kind == MP_DFT_REALNUM;
        for (size_t k = 0; k < size_t(A[q]->npts); ++k)
          for (int i = 0; i < Nf; ++i) {
            const size_t o = k * Nf + i;
            if (narrow) {
              const std::complex<R> aa(R(a[o].real()), R(a[o].imag())), bb(R(b[o].real()), R(b[o].imag()));
              acc[i] += 0.5 * double(std::real(std::conj(aa) * bb));
            }
            else
              acc[i] += 0.5 * std::real(std::conj(a[o]) * b[o]);
          }
      }
    };
    part(en->E, en->D, e);
    part(en->H, en->B, m);
    for (int i = 0; i < Nf; ++i) out[i] = which == 1 ? e[i] : which == 2 ? m[i] : e[i] + m[i];
  }
  int flux_chunk_component(void *flp, int which, int i) override {
    Flux<R> *fl = static_cast<Flux<R> *>(flp);
    auto &lst = which ? fl->H : fl->E;
    if (i < 0 || i >= int(lst.size())) fail("no such dft chunk");
    return lst[i]->c;
  }
  std::vector<cdouble> read_dft(const DftChunk<R> &ch) {
    HIPCHECK(hipStreamSynchronize(stream));
    const size_t n = size_t(ch.npts) * ch.nfreq;
    std::vector<cdouble> out(n);
    if (ch.acc_kind == MP_DFT_F64) {
      HIPCHECK(hipMemcpy(out.data(), ch.acc.p, 16 * n, hipMemcpyDeviceToHost));
    }
    else if (ch.acc_kind == MP_DFT_KAHAN32) {
      // compensated sum: the running sum less the carried compensation term
      std::vector<float> v(2 * n), c(2 * n);
      HIPCHECK(hipMemcpy(v.data(), ch.acc.p, 8 * n, hipMemcpyDeviceToHost));
      HIPCHECK(hipMemcpy(c.data(), ch.comp.p, 8 * n, hipMemcpyDeviceToHost));
      for (size_t q = 0; q < n; ++q)
        out[q] = cdouble(double(v[2 * q]) - double(c[2 * q]), double(v[2 * q + 1]) - double(c[2 * q + 1]));
    }
    else {
      std::vector<R> v(2 * n);
      HIPCHECK(hipMemcpy(v.data(), ch.acc.p, 2 * sizeof(R) * n, hipMemcpyDeviceToHost));
      for (size_t q = 0; q < n; ++q) out[q] = cdouble(v[2 * q], v[2 * q + 1]);
    }
    return out;
  }
  // dft_flux::flux (src/dft.cpp:542-556): the product is formed in the accumulator's type
  void flux(void *flp, double *out) override {
    Flux<R> *fl = static_cast<Flux<R> *>(flp);
    const int Nf = fl->nfreq;
    for (int q = 0; q < Nf; ++q) out[q] = 0;
    for (size_t q = 0; q < fl->E.size() && q < fl->H.size(); ++q) {
      if (fl->E[q]->npts != fl->H[q]->npts) fail("flux of E and H chunks with different points (a monitor off the centred grid)");
      std::vector<cdouble> e = read_dft(*fl->E[q]), h = read_dft(*fl->H[q]);
      const bool narrow = fl->E[q]->acc_kind == MP_DFT_REALNUM;
      for (int kk = 0; kk < fl->E[q]->npts; ++kk)
        for (int i = 0; i < Nf; ++i) {
          const size_t o = size_t(kk) * Nf + i;
          if (narrow) {
            std::complex<R> ee(R(e[o].real()), R(e[o].imag())), hh(R(h[o].real()), R(h[o].imag()));
            out[i] += double(std::real(ee * std::conj(hh)));
          }
          else
            out[i] += std::real(e[o] * std::conj(h[o]));
        }
    }
  }
  int flux_chunks(void *flp, int which) override {
    Flux<R> *fl = static_cast<Flux<R> *>(flp);
    return int(which ? fl->H.size() : fl->E.size());
  }
  long flux_points(void *flp, int which, int i) override {
    Flux<R> *fl = static_cast<Flux<R> *>(flp);
    auto &l = which ? fl->H : fl->E;
    if (i < 0 || i >= int(l.size())) fail("no such dft chunk");
    return l[i]->npts;
  }
  void flux_dft(void *flp, int which, int i, double *out) override {
    Flux<R> *fl = static_cast<Flux<R> *>(flp);
    auto &l = which ? fl->H : fl->E;
    if (i < 0 || i >= int(l.size())) fail("no such dft chunk");
    std::vector<cdouble> v = read_dft(*l[i]);
    for (size_t q = 0; q < v.size(); ++q) {
      out[2 * q] = v[q].real();
      out[2 * q + 1] = v[q].imag();
    }
  }
  // the accumulators of one chunk from host values (a load of data saved from the chunk's
  // get_dft), stored as the chunk keeps them; the compensation of compensated ones restarts
  unsigned acc_version = 0; // accumulators written from the host (loads, scale_dfts)
  void write_dft(DftChunk<R> &ch, const std::vector<cdouble> &v) {
    HIPCHECK(hipStreamSynchronize(stream));
    ++acc_version;
    if (ch.nsnap) fail("internal: deferred DFT steps pending");
    const size_t n = size_t(ch.npts) * ch.nfreq;
    if (v.size() != n) fail("internal: DFT data of the wrong size");
    if (ch.acc_kind == MP_DFT_F64) {
      HIPCHECK(hipMemcpy(ch.acc.p, v.data(), 16 * n, hipMemcpyHostToDevice));
    }
    else if (ch.acc_kind == MP_DFT_KAHAN32) {
      std::vector<float> a(2 * n), c(2 * n, 0.0f);
      for (size_t q = 0; q < n; ++q) a[2 * q] = float(v[q].real()), a[2 * q + 1] = float(v[q].imag());
      HIPCHECK(hipMemcpy(ch.acc.p, a.data(), 8 * n, hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(ch.comp.p, c.data(), 8 * n, hipMemcpyHostToDevice));
    }
    else {
      std::vector<R> a(2 * n);
      for (size_t q = 0; q < n; ++q) a[2 * q] = R(v[q].real()), a[2 * q + 1] = R(v[q].imag());
      HIPCHECK(hipMemcpy(ch.acc.p, a.data(), 2 * sizeof(R) * n, hipMemcpyHostToDevice));
    }
  }
  void flux_set_dft(void *flp, int which, int i, const double *in) override {
    Flux<R> *fl = static_cast<Flux<R> *>(flp);
    auto &l = which ? fl->H : fl->E;
    if (i < 0 || i >= int(l.size())) fail("no such dft chunk");
    const size_t n = size_t(l[i]->npts) * l[i]->nfreq;
    std::vector<cdouble> v(n);
    for (size_t q = 0; q < n; ++q) v[q] = cdouble(in[2 * q], in[2 * q + 1]);
    write_dft(*l[i], v);
  }
  // dft_chunk::scale_dft (src/dft.cpp:402-406): dft[i] *= scale, in the accumulator's type
  void flux_scale(void *flp, cdouble sc) override {
    Flux<R> *fl = static_cast<Flux<R> *>(flp);
    std::vector<std::vector<DftChunk<R> *> *> all{&fl->E, &fl->H};
    if (Energy<R> *en = dynamic_cast<Energy<R> *>(fl)) all.push_back(&en->D), all.push_back(&en->B);
    if (Force<R> *fo = dynamic_cast<Force<R> *>(fl)) all.push_back(&fo->D);
    for (auto *lst : all)
      for (DftChunk<R> *ch : *lst) {
        std::vector<cdouble> v = read_dft(*ch);
        for (cdouble &x : v)
          if (ch->acc_kind == MP_DFT_REALNUM)
            x = cdouble(std::complex<R>(R(x.real()), R(x.imag())) * std::complex<R>(R(sc.real()), R(sc.imag())));
          else
            x *= sc;
        write_dft(*ch, v);
      }
  }
  // fields::get_dft_array (src/dft.cpp:1259-1290).  The grid spanned by the chunks of c
  // (fields::get_dft_component_dims, :1061-1110: corners seeded from the cell's corners one
  // step out) has a dimension per direction with more than one point.  Every point holds its
  // accumulator at num_freq over the chunk's stored weight and, for a weighted chunk, over its
  // nonzero integration weight, times its interpolation weight along the monitor's empty
  // directions (dft_chunk::process_dft_component, :931-1057, retain_interp_weights).  The
  // empty directions are then summed out in realnum (collapse_array,
  // src/array_slice.cpp:554-596).
  long dft_array(void *flp, bool both, int c, int nf, int *rank_out, size_t *dims_out, double *out) override {
    Flux<R> *fl = static_cast<Flux<R> *>(flp);
    if (nf < 0 || nf >= fl->nfreq)
      fail("process_dft_component: frequency index %d is outside the range of the frequency array of size %d", nf, fl->nfreq);
    if (c < 0 || c >= NC) fail("get_dft_array: component %d is not a field component", c);
    std::vector<const DftChunk<R> *> chs;
    for (const DftChunk<R> *ch : fl->E)
      if (ch->c == c) chs.push_back(ch);
    if (both)
      for (const DftChunk<R> *ch : fl->H)
        if (ch->c == c) chs.push_back(ch);
    // dft_force: offdiag1, offdiag2 and diag (src/dft.cpp:1270-1280)
    if (const Force<R> *fo = dynamic_cast<const Force<R> *>(fl))
      for (const DftChunk<R> *ch : fo->D)
        if (ch->c == c) chs.push_back(ch);
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) lo[d] = 2 * gv.num[d] + 1, hi[d] = -1;
    for (const DftChunk<R> *ch : chs)
      for (int d = 0; d < 3; ++d)
        if (gv.has(d)) {
          lo[d] = std::min(lo[d], std::min(ch->is.t[d], ch->ie.t[d]) + ch->shift.t[d]);
          hi[d] = std::max(hi[d], std::max(ch->is.t[d], ch->ie.t[d]) + ch->shift.t[d]);
        }
    int rank = 0, dir[3] = {0, 0, 0};
    size_t n[3] = {1, 1, 1}, total = 1;
    for (int d = 0; d < 3; ++d)
      if (gv.has(d) && (hi[d] - lo[d]) / 2 + 1 > 1) {
        dir[rank] = d;
        n[rank++] = size_t((hi[d] - lo[d]) / 2 + 1);
        total *= n[rank - 1];
      }
    *rank_out = 0;
    dims_out[0] = dims_out[1] = dims_out[2] = 1;
    if (rank == 0) return 0;
    // the kept (non-empty) dimensions and their strides in the collapsed array
    int rrank = 0;
    bool keep[3];
    size_t rstride[3] = {0, 0, 0}, rsize = 1;
    for (int r = rank - 1; r >= 0; --r)
      if ((keep[r] = !(fl->wmax.t[dir[r]] - fl->wmin.t[dir[r]] == 0.0))) {
        rstride[r] = rsize;
        rsize *= n[r];
      }
    for (int r = 0; r < rank; ++r)
      if (keep[r]) dims_out[rrank++] = n[r];
    *rank_out = rrank;
    if (!out) return long(rsize);
    std::vector<std::complex<R>> full(total, std::complex<R>(0, 0));
    for (const DftChunk<R> *ch : chs) {
      const std::vector<cdouble> acc = read_dft(*ch);
      // interpolation weights: the chunk's edge weights along empty directions only
      V3 s0i = ch->s0, s1i = ch->s1, e0i = ch->e0, e1i = ch->e1;
      for (int d = 0; d < 3; ++d)
        if (!ch->empty[d]) s0i.t[d] = s1i.t[d] = e0i.t[d] = e1i.t[d] = 1.0;
      std::vector<double> wi;
      wi.reserve(ch->npts);
      visit_box(ch->is, ch->ie, s0i, s1i, e0i, e1i, 1.0, [&](int64_t, int64_t, const IV &, double w, const int *) { wi.push_back(w); });
      visit_box(ch->is, ch->ie, ch->s0, ch->s1, ch->e0, ch->e1, ch->dV0, [&](int64_t, int64_t k, const IV &p, double w, const int *) {
        cdouble v = acc[size_t(k) * ch->nfreq + nf] / ch->stored_weight;
        if (ch->weighted && v != 0.0) v /= ch->sqrt_w ? std::sqrt(w) : w; // src/dft.cpp:1013-1014
        size_t at = 0;
        for (int r = 0; r < rank; ++r) {
          const int off = p.t[dir[r]] + ch->shift.t[dir[r]] - lo[dir[r]];
          if (off < 0 || size_t(off / 2) >= n[r]) fail("internal: DFT array point out of range");
          at = at * n[r] + size_t(off / 2);
        }
        full[at] = std::complex<R>(wi[size_t(k)] * v);
      });
    }
    if (rrank == rank) {
      for (size_t q = 0; q < total; ++q) out[2 * q] = double(full[q].real()), out[2 * q + 1] = double(full[q].imag());
      return long(total);
    }
    std::vector<R> red(2 * rsize, R(0));
    for (size_t q = 0; q < total; ++q) {
      size_t rem = q, at = 0;
      for (int r = rank - 1; r >= 0; --r) {
        at += rstride[r] * (rem % n[r]);
        rem /= n[r];
      }
      red[2 * at] += full[q].real();
      red[2 * at + 1] += full[q].imag();
    }
    for (size_t q = 0; q < 2 * rsize; ++q) out[q] = double(red[q]);
    return long(rsize);
  }
  // fields::get_eps / get_mu at frequency 0 (src/monitor.cpp:135-195): the number of electric
  // (magnetic) components over the sum of their chi1inv (in realnum) interpolated at p, each
  // point located in the cell, vacuum outside it; the magnetic chi1inv is 1 here
  double eps_mu_at(const V3 &p, bool magnetic) {
    double tr = 0;
    int nc = 0;
    for (int c = magnetic ? 5 : 0; c < (magnetic ? 10 : 5); ++c) {
      const bool has = gv.dim == 1 ? (c == 0 || c == 6) : comp_dir(c) != NODIR;
      if (!has) continue;
      IV locs[8];
      double w[8];
      gv.interpolate(c, p, locs, w);
      double v = 0;
      for (int q = 0; q < 8 && w[q] != 0; ++q) {
        IV il = locs[q];
        cdouble ph = 1.0;
        locate(&il, &ph);
        double u = 1.0;
        if (gv.owns(il) && !magnetic && !s->chi1inv[c].empty()) u = double(R(s->chi1inv[c][size_t(gv.index(c, il))]));
        v += w[q] * u;
      }
      tr += v;
      ++nc;
    }
    return nc / tr;
  }
  // fields::add_dft_near2far (src/near2far.cpp:569-654): for each surface, the tangential E and
  // H of the two directions transverse to its normal (cyclic order) accumulate with their
  // integration weights, off the centred grid, as the equivalent currents n x H and -n x E
  // (the other kind's component along the other tangential direction; sign +1 for the first
  // direction, -1 for the second, negated for E; times the surface's complex weight, weights
  // holding (re, im) pairs).  Each call's
  // chunks go in front of the list, as the reference prepends them.
  void *add_dft_near2far(int nsurf, const int *dirs, const double *vmins, const double *vmaxs, const double *weights,
                         const std::vector<double> &freq, int dec, int nperiods, int *dec_out) override {
    if (!components_allocated) fail("allocate field components (by adding sources) before adding dft objects");
    if (nsurf < 1) fail("add_dft_near2far: no surfaces");
    const int decim = dec > 0 ? dec : auto_decimation(freq);
    *dec_out = decim;
    std::unique_ptr<N2F<R>> n(new N2F<R>);
    n->nfreq = int(freq.size());
    n->freq = freq;
    // dft_near2far::where is the first surface's volume (the union in the loop takes where->v)
    auto v3 = [](const double *q) {
      V3 v;
      for (int d = 0; d < 3; ++d) v.t[d] = q[d];
      return v;
    };
    n->wmin = v3(vmins), n->wmax = v3(vmaxs);
    std::vector<DftChunk<R> *> front;
    for (int w = 0; w < nsurf; ++w) {
      const V3 mn = v3(vmins + 3 * w), mx = v3(vmaxs + 3 * w);
      int nd = dirs[w];
      if (nd < 0) nd = flux_direction(-1, mn, mx);
      if (nd > 2) fail("unknown dft_near2far normal");
      V3 cen;
      for (int d = 0; d < 3; ++d) cen.t[d] = 0.5 * (mn.t[d] + mx.t[d]);
      const double weps = eps_mu_at(cen, false), wmu = eps_mu_at(cen, true);
      auto approxeq = [](double a, double b) { return std::fabs(a - b) < 0.5e-11 * (std::fabs(a) + std::fabs(b)); };
      if (w > 0 && !(approxeq(n->eps, weps) && approxeq(n->mu, wmu))) fail("dft_near2far requires surfaces in a homogeneous medium");
      n->eps = weps, n->mu = wmu;
      const int fd[2] = {(nd + 1) % 3, (nd + 2) % 3};
      if (nperiods > 1)
        for (int i = 0; i < 2; ++i) {
          const double user_width = gv.num[fd[i]] * gv.inva;
          if (gv.has(fd[i]) && bc[High][fd[i]] == Periodic && bc[Low][fd[i]] == Periodic &&
              float(mx.t[fd[i]] - mn.t[fd[i]]) >= float(user_width)) {
            n->periodic_d[i] = fd[i];
            n->periodic_n[i] = nperiods;
            n->period[i] = user_width;
            n->periodic_k[i] = 2 * pi * k[fd[i]].real() * user_width;
          }
        }
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
          const int dc[3] = {0, 1, 4};
          const int c = (i == 0 ? 0 : 5) + dc[fd[j]], c0 = (i == 0 ? 5 : 0) + dc[fd[1 - j]];
          const double sgn = (j == 0 ? 1.0 : -1.0) * (i == 0 ? -1.0 : 1.0);
          std::vector<DftChunk<R> *> lst;
          add_dft(lst, c, mn, mx, freq, true, sgn * cdouble(weights[2 * w], weights[2 * w + 1]), decim, false, c0);
          front.insert(front.begin(), lst.rbegin(), lst.rend());
        }
    }
    n->E = front;
    fluxes.push_back(std::move(n));
    return fluxes.back().get();
  }
  void n2f_info(void *h, double *eps_mu, int *periodic, long *npoints) override {
    N2F<R> *n = static_cast<N2F<R> *>(h);
    eps_mu[0] = n->eps, eps_mu[1] = n->mu;
    for (int i = 0; i < 2; ++i) periodic[2 * i] = n->periodic_d[i], periodic[2 * i + 1] = n->periodic_n[i];
    *npoints = 0;
    for (const DftChunk<R> *ch : n->E) *npoints += ch->npts;
  }
  // The equivalent currents of dft_near2far::farfield_lowlevel (src/near2far.cpp:360-393): every
  // accumulator point of every chunk at its location plus the chunk's periodic-image shift,
  // and its periodic copies within periodic_n periods each way with the Bloch phase of their
  // offset folded into the amplitude; uploaded frequency-major.
  void n2f_sources(N2F<R> &n) {
    if (n.built_t == t && n.built_version == dft_version && n.built_acc == acc_version && n.nsrc > 0) return;
    const int Nf = n.nfreq;
    std::vector<double> x, y, z;
    std::vector<int> cc;
    std::vector<cdouble> amp; // source-major here, transposed on upload
    for (DftChunk<R> *ch : n.E) {
      const std::vector<cdouble> acc = read_dft(*ch);
      std::vector<V3> locs;
      visit_box(ch->is, ch->ie, ch->s0, ch->s1, ch->e0, ch->e1, 1.0, [&](int64_t, int64_t, const IV &p, double, const int *) {
        IV q = p;
        for (int d = 0; d < 3; ++d) q.t[d] += ch->shift.t[d];
        locs.push_back(gv.loc(q));
      });
      for (size_t kk = 0; kk < locs.size(); ++kk)
        for (int i0 = -n.periodic_n[0]; i0 <= n.periodic_n[0]; ++i0)
          for (int i1 = -n.periodic_n[1]; i1 <= n.periodic_n[1]; ++i1) {
            V3 xs = locs[kk];
            if (n.periodic_d[0] >= 0) xs.t[n.periodic_d[0]] += i0 * n.period[0];
            if (n.periodic_d[1] >= 0) xs.t[n.periodic_d[1]] += i1 * n.period[1];
            const cdouble ph = std::polar(1.0, i0 * n.periodic_k[0] + i1 * n.periodic_k[1]);
            x.push_back(xs.t[0]), y.push_back(xs.t[1]), z.push_back(gv.dim == 2 ? 0.0 : xs.t[2]);
            cc.push_back(ch->vc);
            for (int i = 0; i < Nf; ++i) amp.push_back(acc[kk * Nf + i] * ph);
          }
    }
    const size_t ns = x.size();
    if (ns > size_t(INT32_MAX)) fail("near2far: too many surface points");
    // sorted by class (Ex Ey Ez Hx Hy Hz), keeping the list order within a class
    auto cls = [](int c) { return c < 5 ? (c == 4 ? 2 : c) : (c == 9 ? 5 : c - 2); };
    std::vector<size_t> order(ns);
    for (size_t j = 0; j < ns; ++j) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](size_t u, size_t v) { return cls(cc[u]) < cls(cc[v]); });
    for (int c = 0; c <= 6; ++c) n.cls_lo[c] = 0;
    for (size_t j = 0; j < ns; ++j) n.cls_lo[cls(cc[j]) + 1]++;
    for (int c = 0; c < 6; ++c) n.cls_lo[c + 1] += n.cls_lo[c];
    std::vector<double> xs(ns), ys(ns), zs(ns), fm(2 * ns * Nf);
    for (size_t q = 0; q < ns; ++q) {
      const size_t j = order[q];
      xs[q] = x[j], ys[q] = y[j], zs[q] = z[j];
      for (int i = 0; i < Nf; ++i) {
        fm[2 * (size_t(i) * ns + q)] = amp[j * Nf + i].real();
        fm[2 * (size_t(i) * ns + q) + 1] = amp[j * Nf + i].imag();
      }
    }
    n.nsrc = int(ns);
    if (ns) {
      auto put = [&](DevBuf &b, const void *src, size_t bytes) {
        if (b.bytes != bytes) b.alloc(bytes);
        HIPCHECK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
      };
      put(n.sx, xs.data(), 8 * ns), put(n.sy, ys.data(), 8 * ns), put(n.sz, zs.data(), 8 * ns);
      put(n.f0, fm.data(), 8 * fm.size());
      put(n.d_freq, n.freq.data(), 8 * n.freq.size());
    }
    n.built_t = t, n.built_version = dft_version, n.built_acc = acc_version;
  }
  // dft_near2far::farfield_lowlevel (src/near2far.cpp:350-397) at np points on the device: the
  // Green's function sums of near2far.hip, in batches of at most 2^20 (point, frequency) pairs
  void n2f_eval(void *h, const double *pts, long np, double *out) override {
    N2F<R> &n = *static_cast<N2F<R> *>(h);
    if (gv.dim != 2 && gv.dim != 3) fail("only 2d or 3d or cylindrical far-field computation is supported");
    HIPCHECK(hipStreamSynchronize(stream));
    n2f_sources(n);
    const int Nf = n.nfreq;
    if (np <= 0) return;
    if (n.nsrc == 0 || Nf == 0) {
      std::fill(out, out + size_t(np) * Nf * 12, 0.0);
      return;
    }
    const long batch = std::max(1L, (1L << 20) / Nf);
    const long nb = std::min(np, batch);
    const long pairs_max = nb * Nf;
    // blocks per segment: the points times the frequency groups a block takes together
    const long fb = farfield_group(gv.dim), bps = nb * ((Nf + fb - 1) / fb);
    const int segs = int(std::max(1L, std::min<long>((n.nsrc + 511) / 512, (4096 + bps - 1) / bps)));
    // uniformly spaced frequencies (to a few ulps, as linspace makes them) let the kernel step
    // e^{ikr} from one frequency to the next
    bool uniform = Nf > 1;
    for (int i = 2; i < Nf && uniform; ++i) {
      const double want = n.freq[0] + i * (n.freq[1] - n.freq[0]);
      uniform = std::fabs(n.freq[i] - want) <= 8 * DBL_EPSILON * std::max(std::fabs(n.freq[i]), std::fabs(n.freq[0]));
    }
    DevBuf dp, partial, dout;
    dp.alloc(8 * 3 * size_t(nb));
    partial.alloc(8 * 12 * size_t(pairs_max) * segs);
    dout.alloc(8 * 12 * size_t(pairs_max));
    std::vector<double> soa(3 * size_t(nb));
    for (long b0 = 0; b0 < np; b0 += nb) {
      const long m = std::min(nb, np - b0);
      for (long q = 0; q < m; ++q)
        for (int d = 0; d < 3; ++d) soa[size_t(d) * nb + q] = (gv.dim == 2 && d == 2) ? 0.0 : pts[3 * (b0 + q) + d];
      HIPCHECK(hipMemcpyAsync(dp.p, soa.data(), 8 * 3 * size_t(nb), hipMemcpyHostToDevice, stream));
      FarArgs a;
      a.sx = static_cast<const double *>(n.sx.p), a.sy = static_cast<const double *>(n.sy.p), a.sz = static_cast<const double *>(n.sz.p);
      for (int c = 0; c <= 6; ++c) a.cls_lo[c] = n.cls_lo[c];
      a.f0 = static_cast<const double2 *>(n.f0.p);
      a.nsrc = n.nsrc;
      a.px = static_cast<const double *>(dp.p), a.py = a.px + nb, a.pz = a.px + 2 * nb;
      a.npts = int(m);
      a.freq = static_cast<const double *>(n.d_freq.p);
      a.nfreq = Nf;
      a.eps = n.eps, a.mu = n.mu;
      a.dim = gv.dim;
      a.segs = segs;
      a.uniform = uniform ? 1 : 0;
      kbegin(7);
      launch_farfield(a, static_cast<double *>(partial.p), static_cast<double *>(dout.p), stream);
      HIPCHECK(hipGetLastError());
      kend();
      HIPCHECK(hipMemcpyAsync(out + size_t(b0) * Nf * 12, dout.p, 8 * 12 * size_t(m) * Nf, hipMemcpyDeviceToHost, stream));
      HIPCHECK(hipStreamSynchronize(stream));
    }
  }
  void flux_delete(void *flp) override {
    for (size_t q = 0; q < fluxes.size(); ++q)
      if (fluxes[q].get() == flp) {
        Flux<R> *fl = fluxes[q].get();
        std::vector<std::vector<DftChunk<R> *> *> all{&fl->E, &fl->H};
        if (Energy<R> *en = dynamic_cast<Energy<R> *>(fl)) all.push_back(&en->D), all.push_back(&en->B);
        if (Force<R> *fo = dynamic_cast<Force<R> *>(fl)) all.push_back(&fo->D);
        for (auto *lst : all)
          for (DftChunk<R> *ch : *lst)
            for (size_t z = 0; z < dft_chunks.size(); ++z)
              if (dft_chunks[z].get() == ch) {
                dft_chunks.erase(dft_chunks.begin() + z);
                break;
              }
        fluxes.erase(fluxes.begin() + q);
        ++dft_version;
        return;
      }
  }

  // ---- readout ----
  R read1(const R *p, int64_t i) {
    R v;
    HIPCHECK(hipMemcpy(&v, p + i, sizeof(R), hipMemcpyDeviceToHost));
    return v;
  }
  cdouble get_field(int c, const V3 &p) override {
    if (c < 0 || c >= NC || comp_dir(c) == NODIR) fail("invalid component %d", c);
    HIPCHECK(hipStreamSynchronize(stream));
    IV locs[8];
    double w[8];
    gv.interpolate(c, p, locs, w);
    cdouble res = 0.0;
    for (int q = 0; q < 8 && w[q]; ++q) {
      IV il = locs[q];
      cdouble kph = 1.0;
      locate(&il, &kph);
      cdouble v = 0.0;
      if (gv.owns(il) && f[c][0]) {
        const int64_t i = gv.index(c, il);
        v = f[c][1] ? cdouble(read1(f[c][0], i), read1(f[c][1], i)) : cdouble(read1(f[c][0], i), 0);
        v *= kph;
      }
      res += w[q] * v;
    }
    return res;
  }
  R *array_ptr(int c, int cmp, int kind, int ipol) {
    if (c < 0 || c >= NC || cmp < 0 || cmp > 1) fail("invalid component");
    if (kind == MP_ARR_F) return f[c][cmp];
    for (auto &p : pols)
      if (p.isus == ipol && p.allocated) return kind == MP_ARR_P ? p.P[c][cmp] : p.Pp[c][cmp];
    return nullptr;
  }
  double *aux_ptr(int c, int cmp, int kind) {
    if (c < 0 || c >= NC || cmp < 0 || cmp > 1) fail("invalid component");
    if (kind == MP_ARR_F_U) return f_u[c][cmp];
    if (kind == MP_ARR_F_W) return f_w[c][cmp];
    return f_cond[c][cmp];
  }
  static bool is_aux(int kind) { return kind == MP_ARR_F_U || kind == MP_ARR_F_W || kind == MP_ARR_F_COND; }
  template <class T> void read_layout(const T *p, double *out) {
    std::vector<T> v(gv.ntot);
    HIPCHECK(hipMemcpy(v.data(), p, sizeof(T) * gv.ntot, hipMemcpyDeviceToHost));
    gv.for_each_point([&](int64_t r, int64_t d) { out[r] = double(v[d]); });
  }
  template <class T> void write_layout(T *p, const double *in) {
    std::vector<T> v(gv.ntot);
    HIPCHECK(hipMemcpy(v.data(), p, sizeof(T) * gv.ntot, hipMemcpyDeviceToHost));
    gv.for_each_point([&](int64_t r, int64_t d) { v[d] = T(in[r]); });
    HIPCHECK(hipMemcpy(p, v.data(), sizeof(T) * gv.ntot, hipMemcpyHostToDevice));
  }
  int get_array(int c, int cmp, int kind, int ipol, double *out) override {
    HIPCHECK(hipStreamSynchronize(stream));
    if (is_aux(kind)) {
      double *p = aux_ptr(c, cmp, kind);
      if (p) {
        read_layout(p, out);
        return 0;
      }
    }
    else if (R *p = array_ptr(c, cmp, kind, ipol)) {
      read_layout(p, out);
      return 0;
    }
    for (size_t i = 0; i < gv.rntot; ++i) out[i] = 0;
    return 1;
  }
  void set_array(int c, int cmp, int kind, int ipol, const double *in) override {
    HIPCHECK(hipStreamSynchronize(stream));
    if (is_aux(kind)) {
      double *p = aux_ptr(c, cmp, kind);
      if (!p) fail("array not allocated");
      write_layout(p, in);
      return;
    }
    R *p = array_ptr(c, cmp, kind, ipol);
    if (!p) fail("array not allocated");
    write_layout(p, in);
  }
  bool have_component(int c) const override { return c >= 0 && c < NC && f[c][0]; }
  void times(double *t7) const override {
    for (int q = 0; q < 7; ++q) t7[q] = timers[q];
  }
  void set_dft_acc(int kind) override {
    if (kind < 0 || kind > 2) fail("unknown DFT accumulator kind");
    if (kind == MP_DFT_KAHAN32 && sizeof(R) != sizeof(float)) fail("compensated FP32 accumulators need FP32 fields");
    dft_acc = kind;
  }
};

// ---- many independent chunks stepped together --------------------------------------------
// Every launch group of a step (B phase, B halos, D phase, D halos, DFT) covers all chunks in
// one kernel: the chunks' launch boxes become items in device memory, source values come from
// per-step rows computed on the host before the loop, and the whole loop can be captured in a
// hipGraph.  The block size is chosen by timing the first steps with each candidate (the
// results do not depend on it).
struct BatchBase {
  virtual ~BatchBase() {}
  virtual void step_n(long n) = 0;
  virtual void options(int graph, int tune, int block) = 0;
  virtual void stats(int *block, double *ms_per_step) const = 0;
};

template <class R> struct Batch : BatchBase {
  std::vector<Fields<R> *> sims;
  hipStream_t stream = nullptr;
  bool use_graph = true, tune = true, tuned = false;
  int block = 256;
  double last_ms = 0;
  // plan
  bool planned = false;
  std::vector<long> plan_versions, src_versions, ptr_versions, dft_versions;
  struct Group {
    int phase;
    int kind; // 0 generic, 1 vectorised, 2 vectorised with per-point uPML
    bool curl, upd;
    DevBuf items;
    std::vector<BatchItem<R>> host;
    DevBuf blk_item[4], blk_first[4];
    unsigned nblk[4];
  };
  std::vector<std::unique_ptr<Group>> groups;
  DevBuf conn_items[2], conn_item_tab[2], conn_first_tab[2];
  unsigned conn_nblk[2] = {0, 0};
  std::vector<BatchDft<R>> dft_host;
  DevBuf dft_items, dft_item_tab, dft_first_tab;
  unsigned dft_nblk = 0;
  bool dft_defer = false; // deferred accumulation (DftDefer) over the chunks' snap rings
  DevBuf dft_snap_item, dft_snap_first, dft_flush_item, dft_flush_first;
  unsigned dft_snap_nblk = 0, dft_flush_nblk = 0;
  long call_steps = 0; // steps of the current step_n call (the last one flushes)
  std::vector<size_t> row_off; // each chunk's offset in the step's source row
  size_t row_len = 0;
  DevBuf rows;

  explicit Batch(std::vector<Fields<R> *> s) : sims(std::move(s)) {
    if (sims.empty()) fail("a batch needs at least one fields object");
    HIPCHECK(hipStreamCreate(&stream));
  }
  ~Batch() override {
    if (stream) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
  }
  // the launch-box layout of the member chunks: split end slabs while batched (a chunk's own
  // step_n restores its default)
  void set_split(bool on) {
    for (Fields<R> *f : sims)
      if (f->batch_split != on) {
        f->batch_split = on;
        f->boxes_valid = false;
      }
  }
  void options(int graph, int t, int b) override {
    use_graph = graph != 0;
    tune = t != 0;
    if (b) {
      bool ok = false;
      for (int c : BATCH_BLOCKS) ok = ok || c == b;
      if (!ok) fail("batch block size must be 64, 128, 256 or 512");
      block = b;
      tuned = true;
    }
  }
  void stats(int *b, double *ms) const override {
    *b = block;
    *ms = last_ms;
  }
  static int bs_index(int b) { return b == 64 ? 0 : b == 128 ? 1 : b == 256 ? 2 : 3; }
  // block tables of a list of item thread counts for block size bs, padded to 8 blocks
  static void tables(const std::vector<uint32_t> &counts, unsigned bs, std::vector<uint32_t> &item, std::vector<uint32_t> &first) {
    item.clear();
    first.clear();
    for (size_t k = 0; k < counts.size(); ++k)
      for (uint32_t f = 0; f < counts[k]; f += bs) {
        item.push_back(uint32_t(k));
        first.push_back(f);
      }
    while (item.size() % 8) {
      item.push_back(0);
      first.push_back(0x7fffffffu); // beyond every item's count: the block exits at once
    }
  }
  template <class T> static void upload(DevBuf &b, const std::vector<T> &v) {
    b.alloc(sizeof(T) * std::max<size_t>(v.size(), 1));
    if (!v.empty()) HIPCHECK(hipMemcpy(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  }
  bool stale() const {
    if (!planned) return true;
    for (size_t q = 0; q < sims.size(); ++q)
      if (sims[q]->plan_version != plan_versions[q] || sims[q]->src_version != src_versions[q] ||
          sims[q]->ptr_version != ptr_versions[q] || sims[q]->dft_version != dft_versions[q] || !sims[q]->plan_valid)
        return true;
    return false;
  }
  // plan from the chunks' next step; returns the steps taken on the way (0 or 1)
  long plan() {
    long taken = 0;
    for (Fields<R> *f : sims)
      if (!f->flux_vols.empty()) fail("batch: time-domain flux planes (add_flux_vol) are not stepped in batches");
    for (int attempt = 0; attempt < 2; ++attempt) {
      // each chunk's source row sits at row_off[q] of the batch's row
      row_off.assign(sims.size(), 0);
      row_len = 0;
      for (size_t q = 0; q < sims.size(); ++q) {
        sims[q]->ensure_srcplan();
        row_off[q] = row_len;
        row_len += sims[q]->splan.rec.size();
      }
      std::vector<typename Fields<R>::Collect> cs(sims.size());
      bool ok = true;
      for (size_t q = 0; q < sims.size(); ++q) {
        cs[q].vofs = int64_t(row_off[q]);
        const long pv = sims[q]->splan.pv;
        ok = sims[q]->collect_step(cs[q]) && ok;
        if (sims[q]->srcplan_stale() || sims[q]->splan.pv != pv) ok = false;
      }
      if (!ok) {
        if (attempt) fail("batch: a chunk keeps allocating arrays or needs the unfused source path");
        // the first step of a chunk allocates lazily; every chunk takes it on its own
        for (Fields<R> *f : sims) f->step_n(1, false);
        set_split(true);
        taken = 1;
        continue;
      }
      build(cs);
      return taken;
    }
    return taken;
  }
  void build(std::vector<typename Fields<R>::Collect> &cs) {
    groups.clear();
    for (int phase = 0; phase < 2; ++phase)
      for (int key = 0; key < 20; ++key) {
        std::unique_ptr<Group> gp(new Group);
        Group &g = *gp;
        g.phase = phase;
        g.kind = key / 4;
        g.curl = key & 2;
        g.upd = key & 1;
        for (auto &c : cs) {
          auto it = c.items[phase].find(key);
          if (it != c.items[phase].end()) g.host.insert(g.host.end(), it->second.begin(), it->second.end());
        }
        if (g.host.empty()) continue;
        upload(g.items, g.host);
        std::vector<uint32_t> counts;
        for (auto &it : g.host) counts.push_back(it.count);
        for (int b = 0; b < 4; ++b) {
          std::vector<uint32_t> item, first;
          tables(counts, unsigned(BATCH_BLOCKS[b]), item, first);
          upload(g.blk_item[b], item);
          upload(g.blk_first[b], first);
          g.nblk[b] = unsigned(item.size());
        }
        groups.push_back(std::move(gp));
      }
    for (int grp = 0; grp < 2; ++grp) {
      std::vector<BatchConn> ci;
      for (auto &c : cs) ci.insert(ci.end(), c.conn[grp].begin(), c.conn[grp].end());
      std::vector<uint32_t> counts;
      for (auto &c : ci) counts.push_back(uint32_t(c.n));
      std::vector<uint32_t> item, first;
      tables(counts, 256, item, first);
      upload(conn_items[grp], ci);
      upload(conn_item_tab[grp], item);
      upload(conn_first_tab[grp], first);
      conn_nblk[grp] = ci.empty() ? 0 : unsigned(item.size());
    }
    dft_host.clear();
    for (auto &c : cs) dft_host.insert(dft_host.end(), c.dft.begin(), c.dft.end());
    {
      std::vector<uint32_t> counts;
      for (auto &d : dft_host) counts.push_back(uint32_t(size_t(d.a.npts) * d.a.nfreq));
      std::vector<uint32_t> item, first;
      tables(counts, 256, item, first);
      upload(dft_item_tab, item);
      upload(dft_first_tab, first);
      dft_nblk = dft_host.empty() ? 0 : unsigned(item.size());
    }
    // deferred accumulation when every chunk accumulates every step and its phases for
    // DFT_DEFER steps fit the flush kernel's table
    dft_defer = !dft_host.empty();
    for (auto &d : dft_host)
      if (d.decimation != 1 || d.a.nfreq * DFT_DEFER > DFT_FLUSH_TAB || !d.snap0) dft_defer = false;
    if (dft_defer) {
      std::vector<uint32_t> counts, item, first;
      for (auto &d : dft_host) counts.push_back(uint32_t(d.a.npts));
      tables(counts, 256, item, first);
      upload(dft_snap_item, item);
      upload(dft_snap_first, first);
      dft_snap_nblk = unsigned(item.size());
      counts.clear();
      for (auto &d : dft_host) counts.push_back(uint32_t(size_t(d.a.npts) * size_t((d.a.nfreq + 3) / 4)));
      tables(counts, 256, item, first);
      upload(dft_flush_item, item);
      upload(dft_flush_first, first);
      dft_flush_nblk = unsigned(item.size());
    }
    plan_versions.clear();
    src_versions.clear();
    ptr_versions.clear();
    dft_versions.clear();
    for (Fields<R> *f : sims) {
      plan_versions.push_back(f->plan_version);
      src_versions.push_back(f->src_version);
      ptr_versions.push_back(f->ptr_version);
      dft_versions.push_back(f->dft_version);
    }
    planned = true;
  }
  void launch_step(long s, int bidx, const double *vrow) {
    for (int phase = 0; phase < 2; ++phase) {
      for (auto &gp : groups)
        if (gp->phase == phase)
          launch_step_batch<R>(static_cast<const BatchItem<R> *>(gp->items.p), static_cast<const uint32_t *>(gp->blk_item[bidx].p),
                               static_cast<const uint32_t *>(gp->blk_first[bidx].p), gp->nblk[bidx], gp->kind, gp->curl, gp->upd,
                               BATCH_BLOCKS[bidx], vrow, stream);
      launch_conn_batch<R>(static_cast<const BatchConn *>(conn_items[phase].p), static_cast<const uint32_t *>(conn_item_tab[phase].p),
                           static_cast<const uint32_t *>(conn_first_tab[phase].p), conn_nblk[phase], stream);
    }
    const BatchDft<R> *di = static_cast<const BatchDft<R> *>(dft_items.p);
    if (!dft_defer) {
      launch_dft_batch<R>(di, static_cast<const uint32_t *>(dft_item_tab.p), static_cast<const uint32_t *>(dft_first_tab.p),
                          dft_nblk, s, stream);
      return;
    }
    const int slot = int(s % DFT_DEFER);
    launch_dft_snap_batch<R>(di, static_cast<const uint32_t *>(dft_snap_item.p), static_cast<const uint32_t *>(dft_snap_first.p),
                             dft_snap_nblk, slot, stream);
    if (slot == DFT_DEFER - 1 || s == call_steps - 1)
      launch_dft_flush_batch<R>(di, static_cast<const uint32_t *>(dft_flush_item.p), static_cast<const uint32_t *>(dft_flush_first.p),
                                dft_flush_nblk, s - slot, slot + 1, stream);
  }
  void step_n(long n) override {
    if (n <= 0) return;
    for (Fields<R> *f : sims) HIPCHECK(hipStreamSynchronize(f->stream));
    long done = 0;
    set_split(true);
    if (stale()) done = plan();
    const long m = n - done;
    if (m <= 0) return;
    for (Fields<R> *f : sims)
      if (!f->plan_valid) fail("internal: a chunk's plan changed during batch planning");
    // source rows for the m steps, and the DFT chunks' starting times
    const size_t rl = std::max<size_t>(row_len, 1);
    std::vector<double> hrows(rl * size_t(m), 0.0);
    for (size_t q = 0; q < sims.size(); ++q) {
      Fields<R> *F = sims[q];
      const std::vector<SrcRecipe> &rec = F->splan.rec;
      for (long st = 0; st < m; ++st)
        for (size_t e = 0; e < rec.size(); ++e) hrows[size_t(st) * rl + row_off[q] + e] = F->recipe_value(rec[e], F->t + st);
    }
    upload(rows, hrows);
    {
      size_t k = 0;
      for (size_t q = 0; q < sims.size(); ++q)
        for (auto &cp : sims[q]->dft_chunks)
          if (sims[q]->f[cp->c][0]) dft_host[k++].t_start = sims[q]->t;
      if (k != dft_host.size()) fail("internal: batch DFT chunks changed");
      upload(dft_items, dft_host);
    }
    const double *drows = static_cast<const double *>(rows.p);
    call_steps = m;
    auto t0 = std::chrono::steady_clock::now();
    long st = 0;
    if (tune && !tuned) {
      // two steps with each candidate block size, timed on the stream
      double best = 1e300;
      for (int b = 0; b < 4 && st + 2 <= m; ++b) {
        hipEvent_t e0, e1;
        HIPCHECK(hipEventCreate(&e0));
        HIPCHECK(hipEventCreate(&e1));
        HIPCHECK(hipEventRecord(e0, stream));
        for (int r = 0; r < 2; ++r, ++st) launch_step(st, b, drows + size_t(st) * rl);
        HIPCHECK(hipEventRecord(e1, stream));
        HIPCHECK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (ms < best) best = ms, block = BATCH_BLOCKS[b];
      }
      tuned = st > 0;
    }
    const int bidx = bs_index(block);
    if (st < m) {
      if (use_graph) {
        hipGraph_t graph;
        hipGraphExec_t exec;
        HIPCHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        for (long q = st; q < m; ++q) launch_step(q, bidx, drows + size_t(q) * rl);
        HIPCHECK(hipStreamEndCapture(stream, &graph));
        HIPCHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        HIPCHECK(hipGraphLaunch(exec, stream));
        HIPCHECK(hipStreamSynchronize(stream));
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
      }
      else
        for (long q = st; q < m; ++q) launch_step(q, bidx, drows + size_t(q) * rl);
    }
    for (Fields<R> *f : sims) f->probe_on(stream);
    HIPCHECK(hipStreamSynchronize(stream));
    HIPCHECK(hipGetLastError());
    last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / double(m);
    for (Fields<R> *f : sims) {
      f->t += m;
      f->check_flag();
    }
  }
};

template <class F> int guard(F fn) {
  try {
    fn();
    return 0;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return -1;
  }
}

static V3 to_v3(const double *p) {
  V3 v;
  for (int d = 0; d < 3; ++d) v.t[d] = p ? p[d] : 0;
  return v;
}

struct FluxHandle {
  FieldsBase *owner;
  void *fl;
  int nfreq, decimation;
};

// The output grid of dft_near2far::get_farfields_array (src/near2far.cpp:411-427):
// floor(extent * resolution) points along each direction of the cell's dimensionality, spaced
// extent / (n - 1) from the volume's low corner (one point at the low corner when n <= 1).
struct FarGrid {
  int rank = 0, dirs[3] = {0, 1, 2};
  size_t dims[3] = {1, 1, 1}, N = 1;
  double dx[3] = {0, 0, 0};
};
static FarGrid far_grid(int dim, const V3 &mn, const V3 &mx, double resolution) {
  FarGrid g;
  for (int d = 0; d < 3; ++d) {
    const bool has = dim == 1 ? d == 2 : (dim == 2 ? d != 2 : true);
    if (!has) continue;
    const double ext = mx.t[d] - mn.t[d];
    size_t n = size_t(std::max(0, int(std::floor(ext * resolution))));
    if (n <= 1)
      n = 1;
    else
      g.dx[g.rank] = ext / double(n - 1);
    g.dims[g.rank] = n;
    g.N *= n;
    g.dirs[g.rank++] = d;
  }
  return g;
}
// dft_near2far::get_farfields_array (src/near2far.cpp:411-483): every grid point's fields in one
// far-field evaluation, rearranged to the 6 x 2 x N x Nfreq row-major array
static std::vector<double> farfields_array(FluxHandle *h, const FarGrid &g, const V3 &mn) {
  const size_t N = g.N, Nf = size_t(h->nfreq);
  std::vector<double> pts(3 * N, 0.0);
  for (size_t i0 = 0; i0 < g.dims[0]; ++i0)
    for (size_t i1 = 0; i1 < g.dims[1]; ++i1)
      for (size_t i2 = 0; i2 < g.dims[2]; ++i2) {
        const size_t idx = (i0 * g.dims[1] + i1) * g.dims[2] + i2, ii[3] = {i0, i1, i2};
        for (int r = 0; r < g.rank; ++r) pts[3 * idx + g.dirs[r]] = mn.t[g.dirs[r]] + double(ii[r]) * g.dx[r];
      }
  std::vector<double> raw(N * Nf * 12);
  h->owner->n2f_eval(h->fl, pts.data(), long(N), raw.data());
  std::vector<double> out(12 * N * Nf);
  for (size_t idx = 0; idx < N; ++idx)
    for (size_t i = 0; i < Nf; ++i)
      for (int k = 0; k < 6; ++k)
        for (int ri = 0; ri < 2; ++ri) out[((k * 2 + ri) * N + idx) * Nf + i] = raw[(idx * Nf + i) * 12 + 2 * k + ri];
  return out;
}

} // namespace mx

using namespace mx;

extern "C" {

const char *mp_last_error(void) { return g_last_error.c_str(); }
const char *mp_engine_name(void) { return "hip-gfx950"; }

mp_structure *mp_structure_new(int dim, double xs, double ys, double zs, double a, mp_scalar_func eps, void *eu, int n_pml,
                               const int *pd, const int *ps, const double *pt, const double *pR, double courant, int single) {
  return mp_structure_new_pml(dim, xs, ys, zs, a, eps, eu, n_pml, pd, ps, pt, pR, nullptr, nullptr, nullptr, nullptr, nullptr,
                              courant, single);
}
mp_structure *mp_structure_new_pml(int dim, double xs, double ys, double zs, double a, mp_scalar_func eps, void *eu, int n_pml,
                                   const int *pd, const int *ps, const double *pt, const double *pR, const double *pstretch,
                                   const mp_scalar_func *pprof, void *const *puser, const double *pint, const double *pint_u,
                                   double courant, int single) {
  HostStructure *s = new HostStructure;
  int rc = guard([&] {
    auto np = [&](double sz) { return sz == 0 ? 1 : int(sz * a + 0.5); };
    if (dim == 1)
      s->gv.init(1, a, 0, 0, int(zs * a + 0.5));
    else if (dim == 2)
      s->gv.init(2, a, np(xs), np(ys), 0);
    else if (dim == 3)
      s->gv.init(3, a, np(xs), np(ys), np(zs));
    else
      fail("dim must be 1, 2 or 3");
    s->single = single != 0;
    s->courant = courant;
    s->dt = courant / a;
    const Geom &gv = s->gv;
    for (int i = 0; i < n_pml; ++i) {
      const int d = pd[i];
      if (d < 0 || d > 2 || !gv.has(d) || gv.num[d] <= 1) continue;
      PmlShape sh;
      if (pstretch) sh.mean_stretch = pstretch[i];
      if (pprof) sh.profile = pprof[i];
      if (puser) sh.user = puser[i];
      if (pint && pint[i] > 0) sh.integral = pint[i];
      if (pint_u && pint_u[i] > 0) sh.integral_u = pint_u[i];
      host_use_pml(*s, d, pt[i], ps[i], pR && pR[i] > 0 ? pR[i] : 1e-15, sh);
    }
    if (eps) // set_chi1inv without subpixel averaging (src/anisotropic_averaging.cpp:90-98)
      for (int c = 0; c < 5; ++c) {
        if (comp_dir(c) == NODIR || (gv.dim == 1 && c != 0)) continue;
        std::vector<double> v(gv.ntot, 1.0);
        bool trivial = true;
        int lo[3], hi[3];
        for (int d = 0; d < 3; ++d) {
          lo[d] = 0;
          hi[d] = gv.has(d) ? gv.num[d] : 0;
        }
        for (int j0 = lo[0]; j0 <= hi[0]; ++j0)
          for (int j1 = lo[1]; j1 <= hi[1]; ++j1)
            for (int j2 = lo[2]; j2 <= hi[2]; ++j2) {
              int j[3] = {j0, j1, j2};
              IV here = gv.iloc(c, j);
              V3 h = gv.loc(here);
              double r[3];
              for (int d = 0; d < 3; ++d) r[d] = gv.has(d) ? ((h.t[d] - 0.5 * gv.inva) + (h.t[d] + 0.5 * gv.inva)) * 0.5 : 0.0;
              const int64_t idx = gv.index(c, here);
              const double e = eps(r, eu);
              const double val = single ? double(float(1 / e)) : 1 / e;
              v[idx] = val;
              if (val != 1.0) trivial = false;
            }
        if (!trivial) s->chi1inv[c] = std::move(v);
      }
  });
  if (rc) {
    delete s;
    return nullptr;
  }
  return reinterpret_cast<mp_structure *>(s);
}
void mp_structure_delete(mp_structure *s) { delete reinterpret_cast<HostStructure *>(s); }
size_t mp_structure_ntot(const mp_structure *s) { return reinterpret_cast<const HostStructure *>(s)->gv.ntot; }
int mp_structure_num(const mp_structure *s, int *n) {
  for (int d = 0; d < 3; ++d) n[d] = reinterpret_cast<const HostStructure *>(s)->gv.num[d];
  return 0;
}
double mp_structure_dt(const mp_structure *s) { return reinterpret_cast<const HostStructure *>(s)->dt; }
int mp_structure_set_chi1inv(mp_structure *sp, int c, const double *arr) {
  return guard([&] {
    HostStructure *s = reinterpret_cast<HostStructure *>(sp);
    if (c < 0 || c >= 5 || comp_dir(c) == NODIR) fail("chi1inv: electric component expected");
    s->chi1inv[c].assign(arr, arr + s->gv.ntot);
  });
}
int mp_structure_get_chi1inv(const mp_structure *sp, int c, double *arr) {
  return guard([&] {
    const HostStructure *s = reinterpret_cast<const HostStructure *>(sp);
    if (c < 0 || c >= 5) fail("chi1inv: electric component expected");
    for (size_t i = 0; i < s->gv.ntot; ++i) arr[i] = s->chi1inv[c].empty() ? 1.0 : s->chi1inv[c][i];
  });
}
int mp_structure_add_susceptibility(mp_structure *sp, double omega_0, double gamma, int drude, mp_scalar_func sigma, void *su) {
  return mp_structure_add_susceptibility_ft(sp, 0, omega_0, gamma, drude, sigma, su);
}
int mp_structure_add_susceptibility_ft(mp_structure *sp, int ft, double omega_0, double gamma, int drude, mp_scalar_func sigma,
                                       void *su) {
  int idx = -1;
  int rc = guard([&] {
    if (ft != 0 && ft != 1) fail("add_susceptibility: field type must be E_stuff or H_stuff");
    HostStructure *s = reinterpret_cast<HostStructure *>(sp);
    const Geom &gv = s->gv;
    HostSus sus;
    sus.omega_0 = s->single ? double(float(omega_0)) : omega_0;
    sus.gamma = s->single ? double(float(gamma)) : gamma;
    sus.drude = drude != 0;
    sus.ft = ft;
    const int c0 = ft == 0 ? 0 : 5;
    for (int c = c0; c < c0 + 5; ++c) {
      if (comp_dir(c) == NODIR || (gv.dim == 1 && c != (ft == 0 ? 0 : 6))) continue;
      std::vector<double> v(gv.ntot, 0.0);
      bool trivial = true;
      for (int j0 = 0; j0 <= (gv.has(0) ? gv.num[0] : 0); ++j0)
        for (int j1 = 0; j1 <= (gv.has(1) ? gv.num[1] : 0); ++j1)
          for (int j2 = 0; j2 <= (gv.has(2) ? gv.num[2] : 0); ++j2) {
            int j[3] = {j0, j1, j2};
            IV here = gv.iloc(c, j);
            V3 h = gv.loc(here);
            double r[3] = {h.t[0], h.t[1], h.t[2]};
            const double val = sigma ? sigma(r, su) : 0.0;
            v[gv.index(c, here)] = val;
            if (val != 0) trivial = false;
          }
      if (!trivial) sus.sigma[c] = std::move(v);
    }
    s->sus.push_back(std::move(sus));
    idx = int(s->sus.size()) - 1;
  });
  return rc ? rc : idx;
}
// structure_chunk::set_conductivity (src/structure.cpp:868-905): sampled at the D (B)
// component's points over the whole array; an E (H) component's value is multiplied by its
// chi1inv (the magnetic chi1inv is 1 here, mu = 1)
int mp_structure_set_conductivity(mp_structure *sp, int c, mp_scalar_func C, void *user) {
  return guard([&] {
    HostStructure *s = reinterpret_cast<HostStructure *>(sp);
    const Geom &gv = s->gv;
    if (c < 0 || c >= NC || comp_dir(c) == NODIR || !C) fail("invalid component for conductivity");
    const int d = comp_dir(c);
    const int cC = is_electric(c) ? dir_comp(10, d) : is_magnetic(c) ? dir_comp(15, d) : c;
    if (gv.dim == 1 && cC != 10 && cC != 16) return; // the component is absent from the grid
    const std::vector<double> *multby = is_electric(c) && !s->chi1inv[c].empty() ? &s->chi1inv[c] : nullptr;
    auto rn = [&](double x) { return s->single ? double(float(x)) : x; };
    std::vector<double> v(gv.ntot, 0.0);
    bool trivial = true;
    for (int j0 = 0; j0 <= (gv.has(0) ? gv.num[0] : 0); ++j0)
      for (int j1 = 0; j1 <= (gv.has(1) ? gv.num[1] : 0); ++j1)
        for (int j2 = 0; j2 <= (gv.has(2) ? gv.num[2] : 0); ++j2) {
          int j[3] = {j0, j1, j2};
          IV here = gv.iloc(cC, j);
          V3 h = gv.loc(here);
          double r[3] = {h.t[0], h.t[1], h.t[2]};
          const int64_t i = gv.index(cC, here);
          const double val = multby ? rn(C(r, user) * rn((*multby)[i])) : rn(C(r, user));
          v[i] = val;
          if (val != 0.0) trivial = false;
        }
    if (trivial)
      s->cond[cC].clear();
    else
      s->cond[cC] = std::move(v);
  });
}
int mp_structure_set_sigma(mp_structure *sp, int isus, int c, const double *sigma) {
  return guard([&] {
    HostStructure *s = reinterpret_cast<HostStructure *>(sp);
    if (isus < 0 || isus >= int(s->sus.size())) fail("no such susceptibility");
    if (c < 0 || c >= 5) fail("sigma: electric component expected");
    bool trivial = true;
    for (size_t i = 0; i < s->gv.ntot; ++i)
      if (sigma[i] != 0) trivial = false;
    if (trivial)
      s->sus[isus].sigma[c].clear();
    else
      s->sus[isus].sigma[c].assign(sigma, sigma + s->gv.ntot);
  });
}

mp_fields *mp_fields_new(mp_structure *sp) {
  FieldsBase *f = nullptr;
  int rc = guard([&] {
    HostStructure *s = reinterpret_cast<HostStructure *>(sp);
    if (s->single)
      f = new Fields<float>(s);
    else
      f = new Fields<double>(s);
  });
  return rc ? nullptr : reinterpret_cast<mp_fields *>(f);
}
void mp_fields_delete(mp_fields *f) { delete reinterpret_cast<FieldsBase *>(f); }
#define F_ reinterpret_cast<FieldsBase *>(f)
int mp_fields_use_real_fields(mp_fields *f) {
  return guard([&] { F_->use_real_fields(); });
}
int mp_fields_use_bloch(mp_fields *f, const double *k) {
  return guard([&] {
    int j = 0;
    for (int d = 0; d < 3; ++d)
      if (F_->s->gv.has(d)) F_->use_bloch(d, k[j++]);
  });
}
int mp_fields_use_bloch_dir(mp_fields *f, int d, double kr, double ki) {
  return guard([&] { F_->use_bloch(d, cdouble(kr, ki)); });
}
int mp_fields_set_boundary(mp_fields *f, int side, int d, int cond) {
  return guard([&] { F_->set_boundary(side, d, cond); });
}
int mp_fields_add_point_source_legacy(mp_fields *f, int c, double freq, double width, double peaktime, double cutoff, const double *p,
                                      double ar, double ai, int is_c) {
  return guard([&] {
    // src/sources.cpp:193-213
    const double dt = F_->dt_now(), now = F_->t_now() * dt;
    width /= freq;
    SrcTime st;
    if (is_c)
      st = SrcTime::continuous(freq, width, now, 1e300, cutoff);
    else {
      cutoff = F_->s->gv.inva + cutoff * width;
      if (peaktime <= 0.0) peaktime = now + cutoff;
      peaktime += is_magnetic(c) ? -dt * 0.5 : dt;
      st = SrcTime::gaussian(freq, width, peaktime - cutoff, peaktime + cutoff);
    }
    st.is_integrated = !is_magnetic(c);
    V3 pv = to_v3(p);
    F_->add_volume_source(c, st, pv, pv, cdouble(ar, ai), nullptr, nullptr);
  });
}
int mp_fields_add_volume_source(mp_fields *f, int c, const mp_src_time *src, const double *vmin, const double *vmax, double ar, double ai,
                                mp_complex_func A, void *Au) {
  return guard([&] {
    SrcTime st = src->kind == 0 ? SrcTime::gaussian(src->freq, src->width, src->start_time, src->end_time)
                                : SrcTime::continuous(src->freq, src->width, src->start_time, src->end_time, src->slowness);
    st.is_integrated = src->is_integrated != 0;
    F_->add_volume_source(c, st, to_v3(vmin), to_v3(vmax), cdouble(ar, ai), A, Au);
  });
}
static SrcTime src_time_of(const mp_src_time *src) {
  SrcTime st = src->kind == 0 ? SrcTime::gaussian(src->freq, src->width, src->start_time, src->end_time)
                              : SrcTime::continuous(src->freq, src->width, src->start_time, src->end_time, src->slowness);
  st.is_integrated = src->is_integrated != 0;
  return st;
}
// amp_file_func (src/sources.cpp:345-376) over a caller's array: the point's position
// relative to the volume's centre mapped to [0, 1] along each direction the volume spans,
// then linearly interpolated with mirror boundaries (map_coordinates and linear_interpolate,
// src/fields.cpp:754-810)
struct AmpGrid {
  const double *v; // interleaved (re, im), row major
  int n[3];
  double span[3];
};
static int mirror_at(int i, int n) { return i >= n ? 2 * n - 1 - i : (i < 0 ? -1 - i : i); }
static double amp_grid_value(const AmpGrid &G, const double r_in[3], int part) {
  int lo[3], hi[3];
  double w[3];
  for (int d = 0; d < 3; ++d) {
    double r = r_in[d];
    r = r < 0.0 ? -r : (r > 1.0 ? 1.0 - r : r);
    const int n = G.n[d];
    lo[d] = mirror_at(int(r * n), n);
    const double off = r * n - lo[d] - 0.5;
    hi[d] = mirror_at(off >= 0.0 ? lo[d] + 1 : lo[d] - 1, n);
    w[d] = std::fabs(off);
  }
  auto at = [&](int x, int y, int z) { return G.v[2 * ((size_t(x) * G.n[1] + y) * G.n[2] + z) + part]; };
  const double z0 = (at(lo[0], lo[1], lo[2]) * (1.0 - w[0]) + at(hi[0], lo[1], lo[2]) * w[0]) * (1.0 - w[1]) +
                    (at(lo[0], hi[1], lo[2]) * (1.0 - w[0]) + at(hi[0], hi[1], lo[2]) * w[0]) * w[1];
  const double z1 = (at(lo[0], lo[1], hi[2]) * (1.0 - w[0]) + at(hi[0], lo[1], hi[2]) * w[0]) * (1.0 - w[1]) +
                    (at(lo[0], hi[1], hi[2]) * (1.0 - w[0]) + at(hi[0], hi[1], hi[2]) * w[0]) * w[1];
  return z0 * (1.0 - w[2]) + z1 * w[2];
}
static void amp_grid_func(const double *rel, void *user, double *re, double *im) {
  const AmpGrid &G = *static_cast<const AmpGrid *>(user);
  double r[3];
  for (int d = 0; d < 3; ++d) r[d] = G.span[d] == 0 ? 0 : 0.5 + rel[d] / G.span[d];
  *re = amp_grid_value(G, r, 0);
  *im = amp_grid_value(G, r, 1);
}
int mp_fields_add_volume_source_array(mp_fields *f, int c, const mp_src_time *src, const double *vmin, const double *vmax,
                                      const double *arr, size_t dim1, size_t dim2, size_t dim3, double ar, double ai) {
  return guard([&] {
    if (!arr || dim1 == 0 || dim2 == 0 || dim3 == 0) fail("add_volume_source: empty amplitude array");
    if (dim1 > (size_t(1) << 30) || dim2 > (size_t(1) << 30) || dim3 > (size_t(1) << 30)) fail("add_volume_source: amplitude array too large");
    const V3 mn = to_v3(vmin), mx = to_v3(vmax);
    AmpGrid G;
    G.v = arr;
    G.n[0] = int(dim1), G.n[1] = int(dim2), G.n[2] = int(dim3);
    const int dim = F_->dimension();
    for (int d = 0; d < 3; ++d) {
      const bool spans = dim == 1 ? d == 2 : (dim == 2 ? d != 2 : true);
      G.span[d] = spans ? mx.t[d] - mn.t[d] : 0.0;
    }
    F_->add_volume_source(c, src_time_of(src), mn, mx, cdouble(ar, ai), amp_grid_func, &G);
  });
}
int mp_fields_add_srcdata(mp_fields *f, int c, const mp_src_time *src, long n, const long *index, const double *amp,
                          int needs_boundary_fix) {
  return guard([&] { F_->add_srcdata(c, src_time_of(src), index, n, amp, needs_boundary_fix != 0); });
}
int mp_fields_require_source_components(mp_fields *f) {
  return guard([&] { F_->require_source_components(); });
}
int mp_fields_step(mp_fields *f) {
  return guard([&] { F_->step_n(1, true); });
}
int mp_fields_step_n(mp_fields *f, long n) {
  return guard([&] { F_->step_n(n, false); });
}

// Staging array of mp_fields_step_n_noalloc: the per-step source rows of the call.  Each GPU
// the library runs on holds its own copy until the process ends; calls on one GPU share it.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(1) << 20;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_BYTES / sizeof(double)];
// the address of each GPU's copy, looked up once per GPU with that GPU current (the symbol's
// name is a host-side stand-in and is never passed to a kernel or a copy)
constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mutex;
static double *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
static double *scratch_on(int dev) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) fail("GPU %d beyond the %d the staging table holds", dev, NOALLOC_MAX_DEVICES);
  std::lock_guard<std::mutex> lk(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    HIPCHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) fail("no device address for the staging array on GPU %d", dev);
    g_scratch_ptr[dev] = static_cast<double *>(ptr);
  }
  return g_scratch_ptr[dev];
}
int mp_fields_step_n_noalloc(mp_fields *f, long n) {
  return guard([&] {
    FieldsBase *b = F_;
    const int dev = b->device_id();
    int cur = 0;
    HIPCHECK(hipGetDevice(&cur));
    if (cur != dev) HIPCHECK(hipSetDevice(dev));
    struct Restore {
      int cur, dev;
      ~Restore() {
        if (cur != dev) (void)hipSetDevice(cur);
      }
    } restore{cur, dev};
    b->step_n_noalloc(n, scratch_on(dev), NOALLOC_SCRATCH_BYTES);
  });
}
double mp_fields_time(const mp_fields *f) {
  auto *b = reinterpret_cast<const FieldsBase *>(f);
  return b->t_now() * b->dt_now();
}
double mp_fields_round_time(const mp_fields *f) {
  auto *b = reinterpret_cast<const FieldsBase *>(f);
  return double(float(b->t_now() * b->dt_now()));
}
long mp_fields_t(const mp_fields *f) { return reinterpret_cast<const FieldsBase *>(f)->t_now(); }
double mp_fields_dt(const mp_fields *f) { return reinterpret_cast<const FieldsBase *>(f)->dt_now(); }
double mp_fields_last_source_time(mp_fields *f) { return F_->last_source_time(); }
mp_batch *mp_batch_new(mp_fields *const *fields, int n) {
  BatchBase *b = nullptr;
  int rc = guard([&] {
    if (n <= 0 || !fields) fail("a batch needs at least one fields object");
    std::vector<Fields<float> *> fs;
    std::vector<Fields<double> *> fd;
    for (int q = 0; q < n; ++q) {
      FieldsBase *F = reinterpret_cast<FieldsBase *>(fields[q]);
      if (auto *x = dynamic_cast<Fields<float> *>(F)) fs.push_back(x);
      if (auto *x = dynamic_cast<Fields<double> *>(F)) fd.push_back(x);
    }
    if (int(fs.size()) == n)
      b = new Batch<float>(fs);
    else if (int(fd.size()) == n)
      b = new Batch<double>(fd);
    else
      fail("the fields objects of a batch must share one precision");
  });
  return rc ? nullptr : reinterpret_cast<mp_batch *>(b);
}
void mp_batch_delete(mp_batch *b) { delete reinterpret_cast<BatchBase *>(b); }
int mp_batch_step_n(mp_batch *b, long n) {
  return guard([&] { reinterpret_cast<BatchBase *>(b)->step_n(n); });
}
int mp_batch_set_options(mp_batch *b, int use_graph, int autotune, int block) {
  return guard([&] { reinterpret_cast<BatchBase *>(b)->options(use_graph, autotune, block); });
}
int mp_batch_stats(mp_batch *b, int *block, double *ms_per_step) {
  return guard([&] { reinterpret_cast<const BatchBase *>(b)->stats(block, ms_per_step); });
}
struct AdjHandle {
  FieldsBase *owner;
};
mp_adjoint *mp_adjoint_new(mp_fields *f, mp_dft_fields *objective, int c, const double *vmin, const double *vmax, int interval,
                           int *lo_xyz, int *n_xyz) {
  AdjHandle *h = new AdjHandle{reinterpret_cast<FieldsBase *>(f)};
  int lo[3] = {0, 0, 0}, box[3] = {1, 1, 1};
  int rc = guard([&] {
    FluxHandle *mon = reinterpret_cast<FluxHandle *>(objective);
    if (!mon || mon->owner != h->owner) fail("the objective monitor belongs to another fields object");
    h->owner->adj_begin(mon->fl, c, to_v3(vmin), to_v3(vmax), interval, lo, box);
  });
  if (rc) {
    delete h;
    return nullptr;
  }
  for (int d = 0; d < 3; ++d) {
    if (lo_xyz) lo_xyz[d] = lo[d];
    if (n_xyz) n_xyz[d] = box[d];
  }
  return reinterpret_cast<mp_adjoint *>(h);
}
void mp_adjoint_delete(mp_adjoint *a) {
  AdjHandle *h = reinterpret_cast<AdjHandle *>(a);
  if (!h) return;
  guard([&] { h->owner->adj_end(); });
  delete h;
}
int mp_adjoint_objective(mp_adjoint *a, double *J) {
  AdjHandle *h = reinterpret_cast<AdjHandle *>(a);
  return guard([&] { *J = h->owner->adj_objective(); });
}
int mp_adjoint_gradient(mp_adjoint *a, double *grad) {
  AdjHandle *h = reinterpret_cast<AdjHandle *>(a);
  return guard([&] { h->owner->adj_gradient(grad); });
}
int mp_fields_field_energy(mp_fields *f, double *energy) {
  return guard([&] { *energy = F_->field_energy(); });
}
int mp_fields_get_field(mp_fields *f, int c, const double *p, double *re, double *im) {
  return guard([&] {
    cdouble v = F_->get_field(c, to_v3(p));
    *re = v.real();
    *im = v.imag();
  });
}
int mp_fields_get_array(mp_fields *f, int c, int cmp, int kind, int ipol, double *out) {
  int missing = 0;
  int rc = guard([&] { missing = F_->get_array(c, cmp, kind, ipol, out); });
  return rc ? rc : missing;
}
int mp_fields_set_array(mp_fields *f, int c, int cmp, int kind, int ipol, const double *in) {
  return guard([&] { F_->set_array(c, cmp, kind, ipol, in); });
}
int mp_fields_have_component(mp_fields *f, int c) { return F_->have_component(c) ? 1 : 0; }
int mp_fields_get_times(mp_fields *f, double *t7) {
  F_->times(t7);
  return 0;
}
int mp_fields_set_kernel_timing(mp_fields *f, int on) {
  return guard([&] { F_->set_ktiming(on != 0); });
}
int mp_fields_get_kernel_timing(mp_fields *f, int which, double *ms, long *n) {
  return guard([&] { F_->get_ktiming(which, ms, n); });
}
int mp_fields_reset_kernel_timing(mp_fields *f) {
  return guard([&] { F_->reset_ktiming(); });
}
int mp_fields_set_dft_accumulator(mp_fields *f, int kind) {
  return guard([&] { F_->set_dft_acc(kind); });
}
mp_dft_flux *mp_fields_add_dft_flux(mp_fields *f, int d, const double *vmin, const double *vmax, const double *freq, int nfreq,
                                    int use_symmetry, int centered, int dec) {
  (void)use_symmetry; // one chunk with the identity symmetry: nothing to reduce
  FluxHandle *h = new FluxHandle;
  int rc = guard([&] {
    std::vector<double> fr(freq, freq + nfreq);
    h->owner = F_;
    h->nfreq = nfreq;
    h->fl = F_->add_dft_flux(d, to_v3(vmin), to_v3(vmax), fr, centered != 0, dec, &h->decimation);
  });
  if (rc) {
    delete h;
    return nullptr;
  }
  return reinterpret_cast<mp_dft_flux *>(h);
}
void mp_dft_flux_delete(mp_dft_flux *flp) {
  FluxHandle *h = reinterpret_cast<FluxHandle *>(flp);
  if (!h) return;
  h->owner->flux_delete(h->fl);
  delete h;
}
int mp_dft_flux_flux(mp_dft_flux *flp, double *out) {
  FluxHandle *h = reinterpret_cast<FluxHandle *>(flp);
  return guard([&] { h->owner->flux(h->fl, out); });
}
int mp_dft_flux_num_chunks(mp_dft_flux *flp, int which) {
  FluxHandle *h = reinterpret_cast<FluxHandle *>(flp);
  return h->owner->flux_chunks(h->fl, which);
}
long mp_dft_flux_chunk_points(mp_dft_flux *flp, int which, int i) {
  FluxHandle *h = reinterpret_cast<FluxHandle *>(flp);
  long n = -1;
  guard([&] { n = h->owner->flux_points(h->fl, which, i); });
  return n;
}
int mp_dft_flux_decimation(mp_dft_flux *flp) { return reinterpret_cast<FluxHandle *>(flp)->decimation; }
int mp_dft_flux_get_dft(mp_dft_flux *flp, int which, int i, double *out) {
  FluxHandle *h = reinterpret_cast<FluxHandle *>(flp);
  return guard([&] { h->owner->flux_dft(h->fl, which, i, out); });
}
int mp_dft_flux_set_dft(mp_dft_flux *flp, int which, int i, const double *in) {
  FluxHandle *h = reinterpret_cast<FluxHandle *>(flp);
  return guard([&] { h->owner->flux_set_dft(h->fl, which, i, in); });
}
int mp_dft_flux_scale_dfts(mp_dft_flux *flp, double re, double im) {
  FluxHandle *h = reinterpret_cast<FluxHandle *>(flp);
  return guard([&] { h->owner->flux_scale(h->fl, cdouble(re, im)); });
}

mp_dft_fields *mp_fields_add_dft_fields(mp_fields *f, const int *comps, int ncomp, const double *vmin, const double *vmax,
                                        const double *freq, int nfreq, int centered, int dec) {
  FluxHandle *h = new FluxHandle;
  h->owner = reinterpret_cast<FieldsBase *>(f);
  h->nfreq = nfreq;
  std::vector<double> fr(freq, freq + nfreq);
  int rc = guard([&] {
    h->fl = h->owner->add_dft_fields(comps, ncomp, to_v3(vmin), to_v3(vmax), fr, centered != 0, dec, &h->decimation);
  });
  if (rc) {
    delete h;
    return nullptr;
  }
  return reinterpret_cast<mp_dft_fields *>(h);
}
void mp_dft_fields_delete(mp_dft_fields *df) { mp_dft_flux_delete(reinterpret_cast<mp_dft_flux *>(df)); }
int mp_dft_fields_num_chunks(mp_dft_fields *df) { return mp_dft_flux_num_chunks(reinterpret_cast<mp_dft_flux *>(df), 0); }
long mp_dft_fields_chunk_points(mp_dft_fields *df, int i) {
  return mp_dft_flux_chunk_points(reinterpret_cast<mp_dft_flux *>(df), 0, i);
}
int mp_dft_fields_chunk_component(mp_dft_fields *df, int i) {
  FluxHandle *h = reinterpret_cast<FluxHandle *>(df);
  int c = -1;
  if (guard([&] { c = h->owner->flux_chunk_component(h->fl, 0, i); })) return -1;
  return c;
}
int mp_dft_fields_get_dft(mp_dft_fields *df, int i, double *out) {
  return mp_dft_flux_get_dft(reinterpret_cast<mp_dft_flux *>(df), 0, i, out);
}

struct FluxVolHandle {
  FieldsBase *owner;
  void *fv;
};
mp_flux_vol *mp_fields_add_flux_vol(mp_fields *f, int d, const double *vmin, const double *vmax) {
  FieldsBase *F = reinterpret_cast<FieldsBase *>(f);
  FluxVolHandle *h = new FluxVolHandle{F, nullptr};
  if (guard([&] { h->fv = F->add_flux_vol(d, to_v3(vmin), to_v3(vmax)); })) {
    delete h;
    return nullptr;
  }
  return reinterpret_cast<mp_flux_vol *>(h);
}
void mp_flux_vol_delete(mp_flux_vol *h) { delete reinterpret_cast<FluxVolHandle *>(h); }
int mp_flux_vol_flux(mp_flux_vol *hp, double *out) {
  FluxVolHandle *h = reinterpret_cast<FluxVolHandle *>(hp);
  return guard([&] { *out = h->owner->flux_vol_value(h->fv); });
}
int mp_fields_flux_in_box(mp_fields *f, int d, const double *vmin, const double *vmax, double *out) {
  return guard([&] { *out = reinterpret_cast<FieldsBase *>(f)->flux_in_box(d, to_v3(vmin), to_v3(vmax), true); });
}
int mp_fields_flux_in_box_wrongH(mp_fields *f, int d, const double *vmin, const double *vmax, double *out) {
  return guard([&] { *out = reinterpret_cast<FieldsBase *>(f)->flux_in_box(d, to_v3(vmin), to_v3(vmax), false); });
}
int mp_fields_energy_in_box(mp_fields *f, int which, const double *vmin, const double *vmax, double *out) {
  return guard([&] {
    if (which < 0 || which > 2) fail("energy_in_box: which must be 0 (field), 1 (electric) or 2 (magnetic)");
    *out = reinterpret_cast<FieldsBase *>(f)->energy_in_box(which, to_v3(vmin), to_v3(vmax));
  });
}

static long dft_array(void *hp, bool both, int c, int nf, int *rank, size_t *dims, double *out) {
  FluxHandle *h = reinterpret_cast<FluxHandle *>(hp);
  long n = 0;
  const int rc = guard([&] { n = h->owner->dft_array(h->fl, both, c, nf, rank, dims, out); });
  return rc ? -1 : n;
}
long mp_dft_flux_get_dft_array(mp_dft_flux *fl, int c, int num_freq, int *rank, size_t *dims, double *out) {
  return dft_array(fl, true, c, num_freq, rank, dims, out);
}
long mp_dft_fields_get_dft_array(mp_dft_fields *df, int c, int num_freq, int *rank, size_t *dims, double *out) {
  return dft_array(df, false, c, num_freq, rank, dims, out);
}
long mp_dft_force_get_dft_array(mp_dft_force *fo, int c, int num_freq, int *rank, size_t *dims, double *out) {
  return dft_array(fo, true, c, num_freq, rank, dims, out);
}

mp_near2far *mp_fields_add_dft_near2far(mp_fields *f, int nsurf, const int *dirs, const double *vmins, const double *vmaxs,
                                        const double *weights, const double *freq, int nfreq, int dec, int nperiods) {
  FluxHandle *h = new FluxHandle;
  int rc = guard([&] {
    std::vector<double> fr(freq, freq + nfreq);
    h->owner = F_;
    h->nfreq = nfreq;
    h->fl = F_->add_dft_near2far(nsurf, dirs, vmins, vmaxs, weights, fr, dec, nperiods, &h->decimation);
  });
  if (rc) {
    delete h;
    return nullptr;
  }
  return reinterpret_cast<mp_near2far *>(h);
}
void mp_near2far_delete(mp_near2far *n) { mp_dft_flux_delete(reinterpret_cast<mp_dft_flux *>(n)); }
int mp_near2far_info(const mp_near2far *n, double *eps_mu, int *periodic, long *npoints) {
  const FluxHandle *h = reinterpret_cast<const FluxHandle *>(n);
  if (guard([&] { h->owner->n2f_info(h->fl, eps_mu, periodic, npoints); })) return -1;
  return h->nfreq;
}
int mp_near2far_farfield(mp_near2far *n, const double *x, double *EH) {
  FluxHandle *h = reinterpret_cast<FluxHandle *>(n);
  return guard([&] { h->owner->n2f_eval(h->fl, x, 1, EH); });
}
long mp_near2far_get_farfields_array(mp_near2far *n, const double *wmin, const double *wmax, double resolution, int *rank,
                                     size_t *dims, double *out) {
  FluxHandle *h = reinterpret_cast<FluxHandle *>(n);
  long count = -1;
  const int rc = guard([&] {
    const V3 mn = to_v3(wmin), mx = to_v3(wmax);
    const FarGrid g = far_grid(h->owner->dimension(), mn, mx, resolution);
    count = long(g.N);
    int rr = 0;
    dims[0] = dims[1] = dims[2] = 1;
    for (int r = 0; r < g.rank; ++r)
      if (g.dims[r] > 1) dims[rr++] = g.dims[r];
    *rank = rr;
    if (out) {
      const std::vector<double> a = farfields_array(h, g, mn);
      std::copy(a.begin(), a.end(), out);
    }
  });
  return rc ? -1 : count;
}
// dft_near2far::flux (src/near2far.cpp:517-565): Re(E1 conj(H1)) - Re(E2 conj(H2)) of the
// transverse pair in cyclic order, summed over the grid, times the grid's cell volume
int mp_near2far_flux(mp_near2far *n, int d, const double *wmin, const double *wmax, double resolution, double *out) {
  FluxHandle *h = reinterpret_cast<FluxHandle *>(n);
  return guard([&] {
    const int dim = h->owner->dimension();
    const bool has = dim == 1 ? d == 2 : (dim == 2 ? d == 0 || d == 1 : d >= 0 && d <= 2);
    if (!has) fail("cannot get flux for near2far: co-ordinate mismatch");
    const V3 mn = to_v3(wmin), mx = to_v3(wmax);
    const FarGrid g = far_grid(dim, mn, mx, resolution);
    const std::vector<double> EH = farfields_array(h, g, mn);
    const size_t N = g.N, Nf = size_t(h->nfreq);
    const int cE[3][2] = {{1, 2}, {2, 0}, {0, 1}}, cH[3][2] = {{5, 4}, {3, 5}, {4, 3}};
    for (size_t i = 0; i < Nf; ++i) out[i] = 0;
    auto at = [&](int k, size_t idx, size_t i) {
      return cdouble(EH[((k * 2 + 0) * N + idx) * Nf + i], EH[((k * 2 + 1) * N + idx) * Nf + i]);
    };
    for (size_t idx = 0; idx < N; ++idx)
      for (size_t i = 0; i < Nf; ++i)
        for (int j = 0; j < 2; ++j) out[i] += std::real(at(cE[d][j], idx, i) * std::conj(at(cH[d][j], idx, i))) * (1 - 2 * j);
    double dV = 1;
    for (int q = 0; q < 3; ++q) {
      const bool hq = dim == 1 ? q == 2 : (dim == 2 ? q != 2 : true);
      if (!hq) continue;
      const int nn = int(std::floor((mx.t[q] - mn.t[q]) * resolution));
      if (nn > 1) dV *= (mx.t[q] - mn.t[q]) / (nn - 1);
    }
    for (size_t i = 0; i < Nf; ++i) out[i] *= dV;
  });
}
long mp_near2far_get_dft_array(mp_near2far *n, int c, int num_freq, int *rank, size_t *dims, double *out) {
  return dft_array(n, false, c, num_freq, rank, dims, out);
}
mp_dft_energy *mp_fields_add_dft_energy(mp_fields *f, int nvol, const double *vmins, const double *vmaxs, const double *freq,
                                        int nfreq, int dec) {
  FluxHandle *h = new FluxHandle;
  int rc = guard([&] {
    std::vector<double> fr(freq, freq + nfreq);
    h->owner = F_;
    h->nfreq = nfreq;
    h->fl = F_->add_dft_energy(nvol, vmins, vmaxs, fr, dec, &h->decimation);
  });
  if (rc) {
    delete h;
    return nullptr;
  }
  return reinterpret_cast<mp_dft_energy *>(h);
}
void mp_dft_energy_delete(mp_dft_energy *e) { mp_dft_flux_delete(reinterpret_cast<mp_dft_flux *>(e)); }
mp_dft_force *mp_fields_add_dft_force(mp_fields *f, int nvol, const int *comps, const double *vmins, const double *vmaxs,
                                      const double *weights, const double *freq, int nfreq, int dec) {
  FluxHandle *h = new FluxHandle;
  int rc = guard([&] {
    std::vector<double> fr(freq, freq + nfreq);
    std::vector<int> fd(size_t(std::max(nvol, 0)));
    for (int w = 0; w < nvol; ++w) {
      const int c = comps[w];
      fd[w] = c >= 100 ? (c == 100 ? 0 : c == 101 ? 1 : c == 104 ? 2 : -1) : (c >= 0 && c < NC ? comp_dir(c) : -1);
      if (fd[w] == NODIR) fd[w] = -1;
    }
    h->owner = F_;
    h->nfreq = nfreq;
    h->fl = F_->add_dft_force(nvol, fd.data(), vmins, vmaxs, weights, fr, dec, &h->decimation);
  });
  if (rc) {
    delete h;
    return nullptr;
  }
  return reinterpret_cast<mp_dft_force *>(h);
}
void mp_dft_force_delete(mp_dft_force *fo) { mp_dft_flux_delete(reinterpret_cast<mp_dft_flux *>(fo)); }
int mp_dft_force_scale_dfts(mp_dft_force *fo, double re, double im) {
  return mp_dft_flux_scale_dfts(reinterpret_cast<mp_dft_flux *>(fo), re, im);
}
int mp_dft_energy_scale_dfts(mp_dft_energy *e, double re, double im) {
  return mp_dft_flux_scale_dfts(reinterpret_cast<mp_dft_flux *>(e), re, im);
}
int mp_dft_force_force(mp_dft_force *fo, double *out) {
  FluxHandle *h = reinterpret_cast<FluxHandle *>(fo);
  return guard([&] { h->owner->force_spectrum(h->fl, out); });
}
int mp_dft_energy_get(mp_dft_energy *e, int which, double *out) {
  FluxHandle *h = reinterpret_cast<FluxHandle *>(e);
  return guard([&] {
    if (which < 0 || which > 2) fail("dft_energy: which must be 0 (total), 1 (electric) or 2 (magnetic)");
    h->owner->energy_spectrum(h->fl, which, out);
  });
}
int mp_near2far_num_chunks(mp_near2far *n) { return mp_dft_flux_num_chunks(reinterpret_cast<mp_dft_flux *>(n), 0); }
long mp_near2far_chunk_points(mp_near2far *n, int i) {
  return mp_dft_flux_chunk_points(reinterpret_cast<mp_dft_flux *>(n), 0, i);
}
int mp_near2far_get_dft(mp_near2far *n, int i, double *out) {
  return mp_dft_flux_get_dft(reinterpret_cast<mp_dft_flux *>(n), 0, i, out);
}
int mp_near2far_set_dft(mp_near2far *n, int i, const double *in) {
  return mp_dft_flux_set_dft(reinterpret_cast<mp_dft_flux *>(n), 0, i, in);
}
int mp_near2far_scale_dfts(mp_near2far *n, double re, double im) {
  return mp_dft_flux_scale_dfts(reinterpret_cast<mp_dft_flux *>(n), re, im);
}
// green2d / green3d on the device: the far-field kernel with one current and one point
int mp_near2far_green(int dim, const double *x, double freq, double eps, double mu, const double *x0, int c0, const double *f0,
                      double *EH) {
  return guard([&] {
    ensure_gpu();
    if (dim != 2 && dim != 3) fail("only 2d or 3d Green's functions");
    if (c0 < 0 || c0 >= 10 || comp_dir(c0) == NODIR) fail("unrecognized source type");
    // [f0.re f0.im sx sy sz px py pz freq] (the amplitude first, 16-byte aligned), then the
    // component, the partial and the output
    double host[9] = {f0[0], f0[1], x0[0], x0[1], dim == 3 ? x0[2] : 0.0, x[0], x[1], dim == 3 ? x[2] : 0.0, freq};
    DevBuf buf, part, outb;
    buf.alloc(sizeof host), part.alloc(12 * 8), outb.alloc(12 * 8);
    HIPCHECK(hipMemcpy(buf.p, host, sizeof host, hipMemcpyHostToDevice));
    const double *d = static_cast<const double *>(buf.p);
    FarArgs a;
    a.f0 = reinterpret_cast<const double2 *>(d);
    a.sx = d + 2, a.sy = d + 3, a.sz = d + 4, a.px = d + 5, a.py = d + 6, a.pz = d + 7, a.freq = d + 8;
    const int cls = c0 < 5 ? (c0 == 4 ? 2 : c0) : (c0 == 9 ? 5 : c0 - 2);
    for (int c = 0; c <= 6; ++c) a.cls_lo[c] = c > cls ? 1 : 0;
    a.nsrc = 1, a.npts = 1, a.nfreq = 1, a.eps = eps, a.mu = mu, a.dim = dim, a.segs = 1, a.uniform = 0;
    launch_farfield(a, static_cast<double *>(part.p), static_cast<double *>(outb.p), nullptr);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpy(EH, outb.p, 12 * 8, hipMemcpyDeviceToHost));
  });
}

} // extern "C"
