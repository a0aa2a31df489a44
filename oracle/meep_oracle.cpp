/* meep_oracle.cpp — CPU restatement of Meep's FDTD time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the CPU reference the MI355X engine is compared with; it
 * is never linked into, loaded by or called from the product library.  Only tests/, the
 * smoke check and bench.py's CPU-baseline leg use it.
 *
 * It restates, for one chunk spanning the whole cell with the identity symmetry, the
 * reference's algorithm for fields::step() and everything it calls, following the loops of
 * NanoComp/meep (paths relative to the reference tree):
 *   grid / layout      src/vec.cpp:262-296,333-344,395-479,528-605; src/meep/vec.hpp:151-350
 *   structure, uPML    src/structure.cpp:252-273,625-691; src/anisotropic_averaging.cpp:90-98
 *   susceptibility     src/susceptibility.cpp:76-281; src/structure.cpp:471-507
 *   fields, allocation src/fields.cpp:32-86,428-456,480-574,660-679
 *   step order         src/step.cpp:35-139,225-325
 *   curl + uPML        src/step_db.cpp:44-146; src/step_generic.cpp:65-249
 *   E/H update         src/update_eh.cpp:66-215; src/step_generic.cpp:566-785
 *   polarisation       src/update_pols.cpp:40-62
 *   boundaries         src/boundaries.cpp:78-119,191-230,286-345,368-600
 *   sources            src/sources.cpp:67-145,193-312,462-490
 *   loop_in_chunks     src/loop_in_chunks.cpp:232-540
 *   DFT / flux         src/dft.cpp:57-129,175-308,542-556,602-650
 *   point probes       src/monitor.cpp:35-108
 * It implements the C-ABI of include/meep_amd.h so tests drive it and the engine with the
 * same calls.  realnum is a template parameter (double, or float for --enable-single).
 *
 * Pinning (what holds this restatement to the reference, beyond reading it):
 *   - the whole step (curl, uPML, E/H update, Lorentz poles, Bloch/metal walls, integrated
 *     legacy sources) and field_energy: the reference's 13 known answers
 *     (tests/known_results.cpp:137-157), FP64 and FP32, in tests/test_oracle_known_results.py;
 *   - closed forms in tests/closed_forms.py, asserted in tests/test_oracle_closed_forms.py:
 *     the DFT's scale, phase convention and E/H time offset, the current injection's dt and
 *     resolution factors and the flux of a point (a 1D current sheet against the exact
 *     frequency-domain solution of the Yee lattice, 1e-9); the flux integration weights (a
 *     periodic 2D line current, 1e-8); the Lorentz and Drude recurrences (relaxation and free
 *     oscillation against the continuum model, src/susceptibility.cpp:297-307); the magnetic
 *     and metallic walls (an exact quarter-wave lattice eigenmode); energy conservation;
 *   - tests/test_oracle_mutations.py: each planted mistake below fails one of those checks;
 *   - add_dft_fields shares the accumulation code the flux checks pin.
 * The reverse mode and batched stepping entry points are not restated (they fail here); the
 * engine's are checked against finite differences and against one-at-a-time runs.
 *
 * Planted mistakes (tests/test_oracle_mutations.py): building with -DMUT_<name> plants one
 * plausible mistake, which the checks of tests/closed_forms.py must detect:
 *   MUT_DFT_SCALE     the DFT scale loses its dt/sqrt(2 pi) factor
 *   MUT_DFT_SIGN      the DFT phase exp(i omega t) becomes exp(-i omega t)
 *   MUT_DFT_HTIME     H is transformed at t dt instead of t dt - dt/2
 *   MUT_SRC_DT        a current source is injected without its factor dt
 *   MUT_FLUX_WEIGHT   the flux loses the E chunks' integration weights
 *   MUT_DRUDE         a Drude pole keeps the restoring term of a Lorentz pole
 *   MUT_GAMMA         the pole damping uses gamma dt instead of pi gamma dt
 *   MUT_MAGNETIC      a Low magnetic wall zeroes the metallic wall's plane instead
 *   MUT_COND_HALF     the conductive curl update uses dt instead of dt/2
 *   MUT_COND_CHI      an electric conductivity is not multiplied by chi1inv
 *   MUT_COND_SRC      a current source in a conductive medium is not scaled by condinv
 *   MUT_COND_PML      the PML-with-conductivity branch drops its (kap - sig) factor
 *   MUT_PML_CHUNKWIDE a direction's uPML branch is taken at every point of the chunk, as if
 *                     the uPML regions were not chunks of their own
 *   MUT_DFTARR_WEIGHT get_dft_array leaves a weighted chunk's integration weights in
 *   MUT_DFTARR_SIGN   get_dft_array leaves a chunk's stored weight (the flux sign) in
 *   MUT_DFTARR_INTERP get_dft_array sums the planes around an off-grid monitor unweighted
 *   MUT_DFTARR_SHIFT  get_dft_array places a periodic image piece without its shift
 *   MUT_FLUXT_SIGN    the time-domain flux adds its second E x H pair instead of subtracting it
 *   MUT_FLUXT_CENTRE  integrals on the centred grid read each field at its own Yee point
 *   MUT_FLUXT_HALF    a flux plane keeps the flux after the E update only, not the mean with
 *                     the one after the H update
 *   MUT_PML_INTEGRAL  a uPML term's sigma ignores its profile's integral (takes 1/3)
 *   MUT_PML_STRETCH   a uPML term's kappa divides by the profile's integral, not that of u profile(u)
 *   MUT_AMPARR_OFFSET an amplitude array is interpolated without its half-cell offset
 *   MUT_FIXBND_PHASE  a source point moved to its periodic owner keeps the phase, not its conjugate
 *   MUT_N2F_SIGN      near2far's equivalent magnetic current takes +n x E instead of -n x E
 *   MUT_N2F_PHASE     near2far's periodic copies add without their Bloch phase
 *   MUT_HPOLE_SUB     a magnetic susceptibility's P is not subtracted from B (H = mu^-1 B)
 *   MUT_SCALE_DFTS    scale_dfts scales a flux's E accumulators only
 *   MUT_DFTENERGY_WEIGHT  dft_energy weights D and B as well as E and H (volume counted twice)
 *   MUT_FORCE_DIAG    the diagonal stress terms all take +0.5 (no -1/2 delta_ij |F|^2 part)
 *   MUT_DFTARR_SQRT   get_dft_array divides square-root-weighted accumulators by the full weight
 */
#include "../include/meep_amd.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

typedef std::complex<double> cdouble;
const double pi = 3.141592653589793238462643383279502884197;
constexpr int NC = 20;
enum { X = 0, Y = 1, Z = 2, NODIR = 5 };
enum { E_stuff = 0, H_stuff, D_stuff, B_stuff, PE_stuff, PH_stuff, WE_stuff, WH_stuff };
enum { CONNECT_PHASE = 0, CONNECT_NEGATE = 1, CONNECT_COPY = 2 };
enum { Periodic = 0, Metallic, Magnetic, None };
enum { High = 0, Low = 1 };

thread_local std::string last_error;

struct oracle_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};
[[noreturn]] void fail(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw oracle_error(std::string("meep: ") + buf);
}

// ---- components (src/meep/vec.hpp:420-560) ---------------------------------------------
inline bool is_electric(int c) { return c < 5; }
inline bool is_magnetic(int c) { return c >= 5 && c < 10; }
inline bool is_D(int c) { return c >= 10 && c < 15; }
inline bool is_B(int c) { return c >= 15 && c < 20; }
inline int ftype(int c) { return is_electric(c) ? E_stuff : is_magnetic(c) ? H_stuff : is_D(c) ? D_stuff : B_stuff; }
inline int component_direction(int c) {
  switch (c % 5) {
    case 0: return X;
    case 1: return Y;
    case 4: return Z;
    default: return NODIR; // R, P: not supported
  }
}
inline int direction_component(int c, int d) {
  int start = (c / 5) * 5;
  return d == X ? start : d == Y ? start + 1 : start + 4;
}
inline int first_field_component(int ft) {
  switch (ft) {
    case E_stuff: return 0;
    case H_stuff: return 5;
    case D_stuff: return 10;
    case B_stuff: return 15;
  }
  return 0;
}
inline bool valid_cart(int c) { return c >= 0 && c < NC && component_direction(c) != NODIR; }

// ---- ivec / vec -------------------------------------------------------------------------
struct ivec {
  int t[3] = {0, 0, 0};
};
struct vec3 {
  double t[3] = {0, 0, 0};
};

// ---- grid_volume restricted to D1/D2/D3 with origin 0 (src/vec.cpp, src/meep/vec.hpp) --
struct GV {
  int dim = 3; // 1, 2, 3
  double a = 1, inva = 1;
  int num[3] = {0, 0, 0};
  ptrdiff_t st[3] = {0, 0, 0};
  size_t ntot = 1;

  bool has(int d) const { return dim == 1 ? d == Z : (dim == 2 ? d != Z : true); }
  // yucky_dir: D2 loops Z, X, Y; others X, Y, Z (src/vec.cpp:262-270)
  int ydir(int k) const { return dim == 2 ? (k + 2) % 3 : k; }
  int ydirs_present(int k) const { return has(ydir(k)); }
  void init(int d, double aa, int nx, int ny, int nz) {
    dim = d;
    a = aa;
    inva = 1.0 / aa;
    num[0] = nx;
    num[1] = ny;
    num[2] = nz;
    ntot = 1;
    for (int dd = 0; dd < 3; ++dd)
      if (has(dd)) ntot *= size_t(num[dd] + 1);
    st[0] = st[1] = st[2] = 0;
    // src/vec.cpp:482-494
    for (int dd = 0; dd < 3; ++dd)
      if (has(dd)) {
        if (dd == Z) st[dd] = 1;
        if (dd == Y) st[dd] = num[Z] + 1;
        if (dd == X) st[dd] = ptrdiff_t(num[Z] + 1) * (num[Y] + 1);
      }
  }
  ivec iyee_shift(int c) const {
    ivec r;
    for (int d = 0; d < 3; ++d)
      if (has(d)) {
        bool s = c == 20 || ((is_electric(c) || is_D(c)) && d == component_direction(c)) ||
                 ((is_magnetic(c) || is_B(c)) && d != component_direction(c));
        r.t[d] = s ? 1 : 0;
      }
    return r;
  }
  ivec little_corner() const { return ivec(); }
  ivec big_corner() const {
    ivec r;
    for (int d = 0; d < 3; ++d)
      if (has(d)) r.t[d] = 2 * num[d];
    return r;
  }
  ivec little_owned_corner(int c) const {
    ivec r, s = iyee_shift(c);
    for (int d = 0; d < 3; ++d)
      if (has(d)) r.t[d] = 2 - s.t[d];
    return r;
  }
  ivec big_owned_corner(int c) const {
    ivec r = big_corner(), s = iyee_shift(c);
    for (int d = 0; d < 3; ++d)
      if (has(d)) r.t[d] -= s.t[d];
    return r;
  }
  bool owns(const ivec &p) const {
    for (int d = 0; d < 3; ++d)
      if (has(d) && !(p.t[d] > 0 && p.t[d] <= 2 * num[d])) return false;
    return true;
  }
  bool contains(const ivec &p) const {
    for (int d = 0; d < 3; ++d)
      if (has(d) && (p.t[d] < 0 || p.t[d] >= (num[d] + 1) * 2)) return false;
    return true;
  }
  bool contains(const vec3 &p) const {
    for (int d = 0; d < 3; ++d)
      if (has(d) && (p.t[d] < -inva || p.t[d] > num[d] * inva + inva)) return false;
    return true;
  }
  // grid_volume::iloc(c, index) (src/vec.cpp:547-556): the point of c's array at index
  ivec iloc_of(int c, ptrdiff_t index) const {
    ivec r, s = iyee_shift(c);
    int order[3] = {X, Y, Z};
    std::sort(order, order + 3, [&](int u, int v) { return st[u] > st[v]; });
    for (int d : order)
      if (has(d)) {
        r.t[d] = int(index / st[d]) * 2 + s.t[d];
        index %= st[d];
      }
    return r;
  }
  ptrdiff_t index(int c, const ivec &p) const {
    ivec s = iyee_shift(c);
    ptrdiff_t idx = 0;
    for (int d = 0; d < 3; ++d)
      if (has(d)) idx += (p.t[d] - s.t[d]) / 2 * st[d];
    return idx;
  }
  vec3 loc(const ivec &p) const {
    vec3 r;
    for (int d = 0; d < 3; ++d)
      if (has(d)) r.t[d] = p.t[d] * (0.5 * inva);
    return r;
  }
  ivec round_vec(const vec3 &p) const {
    ivec r;
    for (int d = 0; d < 3; ++d)
      if (has(d)) {
        double x = p.t[d] * 2 * a;
        r.t[d] = int(floor(fabs(x) + 0.5) * (x < 0 ? -1 : 1));
      }
    return r;
  }
  vec3 yee_shift(int c) const { return loc(iyee_shift(c)); }
  ivec icenter() const {
    ivec r;
    for (int d = 0; d < 3; ++d)
      if (has(d)) r.t[d] = num[d] - num[d] % 2;
    return r;
  }
  // src/vec.cpp:692-720: the walls sit on the Ez (Ex for Z in 3D and 1D) points of index 0 and
  // ntot-1, which have no Yee shift along d, so they are at 0 and n_d / a.
  double boundary_location(int side, int d) const { return side == High ? num[d] * inva : 0.0; }
  // src/vec.cpp:333-344
  void yee2cent_offsets(int c, ptrdiff_t &o1, ptrdiff_t &o2) const {
    o1 = o2 = 0;
    ivec s = iyee_shift(c);
    for (int d = 0; d < 3; ++d)
      if (has(d) && !s.t[d]) {
        if (o1)
          o2 = st[d];
        else
          o1 = st[d];
      }
  }
  // src/vec.cpp:395-430
  bool get_boundary_icorners(int c, int ib, ivec *cs, ivec *ce) const {
    ivec s = iyee_shift(c), cl, cb, clo = little_owned_corner(c), cbo = big_owned_corner(c);
    for (int d = 0; d < 3; ++d)
      if (has(d)) {
        cl.t[d] = s.t[d];
        cb.t[d] = 2 * num[d] + s.t[d];
      }
    *cs = cl;
    *ce = cb;
    int jb = 0;
    for (int d = 0; d < 3; ++d) {
      if (!has(d)) continue;
      if (cl.t[d] < clo.t[d]) {
        if (jb == ib) {
          ce->t[d] = cs->t[d];
          return true;
        }
        cs->t[d] = clo.t[d];
        jb++;
      }
      if (cb.t[d] > cbo.t[d]) {
        if (jb == ib) {
          cs->t[d] = ce->t[d];
          return true;
        }
        ce->t[d] = cbo.t[d];
        jb++;
      }
    }
    return false;
  }
  // interpolate(component, vec, ivec locs[8], weights[8]) — src/vec.cpp:558-605
  void interpolate(int c, const vec3 &pc, ivec locs[8], double w[8]) const {
    const double SMALL = 1e-13;
    vec3 ys = yee_shift(c), p;
    for (int d = 0; d < 3; ++d)
      if (has(d)) p.t[d] = (pc.t[d] - ys.t[d]) * a;
    ivec middle, sh = iyee_shift(c);
    for (int d = 0; d < 3; ++d)
      if (has(d)) middle.t[d] = int(floor(p.t[d])) * 2 + 1 + sh.t[d];
    vec3 midv = loc(middle), dv;
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
        locs[i].t[d] = middle.t[d] - 1;
        w[i] *= 0.5 * (1.0 - dv.t[d]);
        locs[have + i].t[d] = middle.t[d] + 1;
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
        if (-w[i] >= SMALL * 1e5) fail("large negative interpolation weight[%d] = %e", i, w[i]);
        w[i] = 0.0;
      }
      else if (w[i] < SMALL)
        w[i] = 0.0;
    }
    // stupidsort (src/vec.cpp:511-526)
    {
      ivec *l = locs;
      double *ww = w;
      int n = have;
      while (n) {
        if (fabs(ww[0]) < 2e-15) {
          ww[0] = ww[n - 1];
          l[0] = l[n - 1];
          ww[n - 1] = 0.0;
          l[n - 1] = ivec();
        }
        else {
          ww += 1;
          l += 1;
        }
        n -= 1;
      }
    }
    bool all_same = true;
    for (int i = 0; i < 8 && w[i]; ++i)
      if (w[i] != w[0]) all_same = false;
    if (all_same) {
      int nw = 0;
      for (int i = 0; i < 8 && w[i]; ++i) nw++;
      for (int i = 0; i < 8 && w[i]; ++i) w[i] = 1.0 / nw;
    }
  }
};

// LOOP_OVER_IVECS (src/meep/vec.hpp:151-169) with the loop counters exposed.
struct Loop {
  ptrdiff_t n[3], s[3], is[3], idx0;
  int d[3];
  bool pres[3];
};
inline Loop make_loop(const GV &gv, const ivec &is, const ivec &ie) {
  Loop L;
  L.idx0 = 0;
  for (int k = 0; k < 3; ++k) {
    int d = gv.ydir(k);
    L.d[k] = d;
    L.pres[k] = gv.has(d);
    if (gv.has(d)) {
      L.is[k] = is.t[d];
      L.n[k] = (ie.t[d] - is.t[d]) / 2 + 1;
      L.s[k] = gv.st[d];
      L.idx0 += is.t[d] / 2 * gv.st[d];
    }
    else {
      L.is[k] = 0;
      L.n[k] = 1;
      L.s[k] = 0;
    }
  }
  return L;
}
template <class F> inline void loop_ivecs(const GV &gv, const ivec &is, const ivec &ie, F body) {
  Loop L = make_loop(gv, is, ie);
  for (ptrdiff_t i1 = 0; i1 < L.n[0]; ++i1)
    for (ptrdiff_t i2 = 0; i2 < L.n[1]; ++i2)
      for (ptrdiff_t i3 = 0; i3 < L.n[2]; ++i3) {
        ptrdiff_t idx = L.idx0 + i1 * L.s[0] + i2 * L.s[1] + i3 * L.s[2];
        body(idx, L, i1, i2, i3);
      }
}
// The same loop with the two outer counters shared out over OpenMP threads (the reference's
// collapse(2) loops, src/meep/vec.hpp:201-230), for bodies that touch only their own point.
// Loops of fewer than ORACLE_PAR_MIN points stay on one thread (waking the team costs more
// than they take); the CPU-baseline build lowers it.
#ifndef ORACLE_PAR_MIN
#define ORACLE_PAR_MIN (1 << 16)
#endif
template <class F> inline void loop_ivecs_par(const GV &gv, const ivec &is, const ivec &ie, F body) {
  const Loop L = make_loop(gv, is, ie);
#pragma omp parallel for collapse(2) schedule(static) if (L.n[0] * L.n[1] * L.n[2] >= ORACLE_PAR_MIN)
  for (ptrdiff_t i1 = 0; i1 < L.n[0]; ++i1)
    for (ptrdiff_t i2 = 0; i2 < L.n[1]; ++i2)
      for (ptrdiff_t i3 = 0; i3 < L.n[2]; ++i3) {
        ptrdiff_t idx = L.idx0 + i1 * L.s[0] + i2 * L.s[1] + i3 * L.s[2];
        body(idx, L, i1, i2, i3);
      }
}
inline ivec loop_iloc(const Loop &L, ptrdiff_t i1, ptrdiff_t i2, ptrdiff_t i3) {
  ivec r;
  ptrdiff_t ii[3] = {i1, i2, i3};
  for (int k = 0; k < 3; ++k)
    if (L.pres[k]) r.t[L.d[k]] = int(L.is[k] + 2 * ii[k]);
  return r;
}
// IVEC_LOOP_WEIGHT (src/meep/vec.hpp:338-350)
inline double weight1(const vec3 &s0, const vec3 &s1, const vec3 &e0, const vec3 &e1, ptrdiff_t i,
                      ptrdiff_t n, int dir, bool present) {
  auto g = [&](const vec3 &v) { return present ? v.t[dir] : 1.0; };
  return (i > 1 && i < n - 2) ? 1.0
                              : (i == 0 ? g(s0)
                                        : (i == 1 ? g(s1) : i == n - 1 ? g(e0) : (i == n - 2 ? g(e1) : 1.0)));
}
inline double loop_weight(const GV &gv, const Loop &L, ptrdiff_t i1, ptrdiff_t i2, ptrdiff_t i3,
                          const vec3 &s0, const vec3 &s1, const vec3 &e0, const vec3 &e1, double dV) {
  double w1 = weight1(s0, s1, e0, e1, i1, L.n[0], L.d[0], gv.has(L.d[0]));
  double w2 = weight1(s0, s1, e0, e1, i2, L.n[1], L.d[1], gv.has(L.d[1]));
  double w3 = weight1(s0, s1, e0, e1, i3, L.n[2], L.d[2], gv.has(L.d[2]));
  return w3 * (w2 * (dV * w1));
}

inline int cycle_direction(int d, int shift) { return (d + shift + 99) % 3; }

// ---- source time dependence (src/sources.cpp:67-145) ------------------------------------
struct SrcTime {
  int kind = 0; // 0 gaussian, 1 continuous
  double freq = 0, width = 0, peak_time = 0, cutoff = 0, fwidth = 0;
  double start_time = 0, end_time = 0, slowness = 3.0;
  bool is_integrated = true;
  double current_time = NAN;
  cdouble cur_dipole = 0, cur_current = 0;

  static double gaussian_bandwidth(double width) { return sqrt(-2.0 * log(1e-7)) / (width * pi); }
  static SrcTime gaussian(double f, double w, double st, double et) {
    SrcTime s;
    s.kind = 0;
    s.freq = f;
    s.width = w;
    s.peak_time = 0.5 * (st + et);
    s.cutoff = (et - st) * 0.5;
    s.fwidth = gaussian_bandwidth(w);
    while (exp(-s.cutoff * s.cutoff / (2 * w * w)) < 1e-100) s.cutoff *= 0.9;
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
    if (kind == 0) {
      double tt = time - peak_time;
      if (float(fabs(tt)) > cutoff) return 0.0;
      cdouble amp = 1.0 / cdouble(0, -2 * pi * freq);
      return exp(-tt * tt / (2 * width * width)) * std::polar(1.0, -2 * pi * freq * tt) * amp;
    }
    float rtime = float(time);
    if (rtime < start_time || rtime > end_time) return 0.0;
    cdouble amp = 1.0 / (cdouble(0, -1.0) * (2 * pi) * freq);
    if (width == 0.0) return exp(cdouble(0, -1.0) * (2 * pi) * freq * time) * amp;
    double ts = (time - start_time) / width - slowness;
    double te = (end_time - time) / width - slowness;
    return exp(cdouble(0, -1.0) * (2 * pi) * freq * time) * amp * (1.0 + tanh(ts)) * (1.0 + tanh(te)) * 0.25;
  }
  cdouble current(double time, double dt) const { return (dipole(time + dt) - dipole(time)) / dt; }
  void update(double time, double dt) {
    if (time != current_time) {
      cur_dipole = dipole(time);
      cur_current = current(time, dt);
      current_time = time;
    }
  }
  double last_time() const { return kind == 0 ? float(peak_time + cutoff) : end_time; }
  double get_fwidth() const { return kind == 0 ? fwidth : 0.0; }
  bool equal(const SrcTime &o) const {
    if (kind != o.kind || is_integrated != o.is_integrated) return false;
    if (kind == 0) return freq == o.freq && width == o.width && peak_time == o.peak_time && cutoff == o.cutoff;
    return freq == o.freq && width == o.width && start_time == o.start_time && end_time == o.end_time &&
           slowness == o.slowness;
  }
};

// ---- structure ---------------------------------------------------------------------------
template <class R> struct Sus {
  double omega_0 = 0, gamma = 0;
  bool drude = false;
  int ft = E_stuff; // the field type it polarises (E_stuff or H_stuff)
  std::vector<R> sigma[NC];
  bool trivial[NC];
};

template <class R> struct Structure {
  GV gv;
  double courant = 0.5, dt = 0.5, a = 1;
  std::vector<R> chi1inv[NC]; // diagonal only; empty == NULL (trivial)
  // conductivity[c][dir(c)] by D or B component c (src/meep.hpp:599); empty == NULL; condinv
  // = 1 / (1 + conductivity dt / 2), recomputed when stale (src/structure.cpp:693-707)
  std::vector<R> cond[NC], condinv[NC];
  bool condinv_stale = false;
  std::vector<R> sig[3], kap[3], siginv[3];
  int sigsize[6] = {0, 0, 0, 0, 0, 0};
  std::vector<Sus<R>> sus; // insertion order; the reference's chiP list is newest first
};

// pml_x (src/structure.cpp:625-628)
inline double pml_x(int i, double dx, double bloc, double a) {
  double here = i * 0.5 / a;
  return (0.5 / a * ((int)(dx * (2 * a) + 0.5) - (int)(fabs(bloc - here) * (2 * a) + 0.5)));
}

// structure_chunk::use_pml (src/structure.cpp:630-691).  The reference splits each PML slab
// into its own chunk so no chunk sees both sides of one direction; one chunk here holds
// both sides, so each side's profile is written into the same array.
// one uPML term of a boundary_region (src/meep.hpp:665-672): the profile (null: the quadratic
// pml_quadratic_profile, src/structure.cpp:252-255) and its integrals
struct PmlTerm {
  double mean_stretch = 1.0;
  mp_scalar_func profile = nullptr;
  void *user = nullptr;
  double integral = 1. / 3., integral_u = 1. / 4.;
  double eval(double u) const {
    if (!profile) return u * u;
    const double r[3] = {u, 0, 0};
    return profile(r, user);
  }
};
template <class R> void use_pml(Structure<R> &s, int d, double dx, double bloc, double Rasym, const PmlTerm &term) {
  if (dx <= 0.0) return;
#ifdef MUT_PML_INTEGRAL
  const double integral = 1. / 3., integral_u = term.integral_u, mean_stretch = term.mean_stretch;
#elif defined(MUT_PML_STRETCH)
  const double integral = term.integral, integral_u = term.integral, mean_stretch = term.mean_stretch;
#else
  const double integral = term.integral, integral_u = term.integral_u, mean_stretch = term.mean_stretch;
#endif
  const double prefac = (-log(Rasym)) / (4 * dx * integral);
  const double kappa_prefac = (mean_stretch - 1) / integral_u;
  const GV &gv = s.gv;
  bool found = false;
  for (int i = 0; i <= 2 * gv.num[d] + 1; ++i)
    if (pml_x(i, dx, bloc, s.a) > 0) {
      found = true;
      break;
    }
  if (!found) return;
  for (int dd = 0; dd < 3; ++dd)
    if (s.sig[dd].empty()) {
      int spml = (dd == d) ? (2 * gv.num[d] + 2) : 1;
      s.sigsize[dd] = spml;
      s.sig[dd].assign(spml, R(0));
      s.kap[dd].assign(spml, R(1));
      s.siginv[dd].assign(spml, R(1));
    }
  if (s.sigsize[d] == 1) { // first PML side in this direction: widen the 1-point placeholder
    int spml = 2 * gv.num[d] + 2;
    s.sigsize[d] = spml;
    s.sig[d].assign(spml, R(0));
    s.kap[d].assign(spml, R(1));
    s.siginv[d].assign(spml, R(1));
  }
  for (int i = 0; i <= 2 * gv.num[d] + 1; ++i) {
    double x = pml_x(i, dx, bloc, s.a);
    if (x > 0) {
      double sp = term.eval(x / dx);
      s.sig[d][i] = 0.5 * s.dt * prefac * sp;
      s.kap[d][i] = 1 + kappa_prefac * sp * (x / dx);
      s.siginv[d][i] = 1 / (s.kap[d][i] + s.sig[d][i]);
    }
  }
}

// structure_chunk::update_condinv (src/structure.cpp:693-707)
template <class R> void update_condinv(Structure<R> &s) {
  if (!s.condinv_stale) return;
  for (int c = 0; c < NC; ++c) {
    if (!s.cond[c].empty()) {
      s.condinv[c].assign(s.gv.ntot, R(1));
      for (size_t i = 0; i < s.gv.ntot; ++i) s.condinv[c][i] = 1 / (1 + s.cond[c][i] * s.dt * 0.5);
    }
    else
      s.condinv[c].clear();
  }
  s.condinv_stale = false;
}

// ---- fields --------------------------------------------------------------------------------
struct SrcVol {
  int c;       // E or H component
  int src;     // index into Fields::srcs
  std::vector<ptrdiff_t> index;
  std::vector<cdouble> amp;
};

template <class R> struct DftChunk {
  int c;
  ivec is, ie;
  vec3 s0, s1, e0, e1;
  double dV0 = 1, dV1 = 0;
  ptrdiff_t avg1 = 0, avg2 = 0;
  cdouble scale, stored_weight = 1.0;
  ivec shift;                   // the periodic image this piece of the monitor stands for
  bool empty_dim[3] = {};       // the monitor volume has no extent in this direction
  int vc = -1;                  // near2far: the equivalent source component
  bool sqrt_w = false;          // accumulate with the square root of the integration weight
  cdouble extra_weight = 1.0;   // dft_force: the readout's factor (src/dft.cpp:97,224)
  bool include_w;
  std::vector<double> omega;
  int decimation = 1;
  size_t N = 0;
  int acc = MP_DFT_REALNUM;
  std::vector<std::complex<R>> dft;
  std::vector<cdouble> dft64;
};

template <class R> struct Fields {
  Structure<R> *s;
  GV gv;
  double a, dt, courant;
  int t = 0;
  bool is_real = false;
  R *f[NC][2], *f_u[NC][2], *f_w[NC][2], *fmp[NC][2], *f_cond[NC][2];
  std::vector<std::unique_ptr<R[]>> pool;
  int boundaries[2][3];
  cdouble k[3], eikna[3];
  std::vector<SrcTime> srcs;
  std::vector<SrcVol> sources[4];
  bool components_allocated = false;
  struct Pol {
    int isus;
    bool allocated = false;
    R *P[NC][2], *Pp[NC][2];
  };
  std::vector<Pol> pols;
  int plus_c[NC], minus_c[NC], plus_d[NC], minus_d[NC];
  bool have_p[NC], have_m[NC];
  bool connections_valid = false;
  struct Conn {
    R *dst[2], *src[2];
    int cls;
    std::complex<R> ph;
  };
  std::vector<Conn> conns[8];
  std::vector<R *> zeroes[8];
  std::vector<std::unique_ptr<DftChunk<R>>> dft_chunks;
  int dft_acc = MP_DFT_REALNUM;
  double times[7] = {0, 0, 0, 0, 0, 0, 0};

  R *alloc(const R *copy_from) {
    pool.emplace_back(new R[gv.ntot]);
    R *p = pool.back().get();
    if (copy_from)
      memcpy(p, copy_from, sizeof(R) * gv.ntot);
    else
      for (size_t i = 0; i < gv.ntot; ++i) p[i] = 0;
    return p;
  }
  int ncmp() const { return is_real ? 1 : 2; }

  explicit Fields(Structure<R> *st) : s(st), gv(st->gv) {
    a = gv.a;
    dt = s->dt;
    courant = s->courant;
    for (int c = 0; c < NC; ++c)
      for (int cmp = 0; cmp < 2; ++cmp) f[c][cmp] = f_u[c][cmp] = f_w[c][cmp] = fmp[c][cmp] = f_cond[c][cmp] = nullptr;
    for (int d = 0; d < 3; ++d) {
      k[d] = 0.0;
      eikna[d] = 1.0;
    }
    // default boundaries: Metallic where the grid has a boundary (src/fields.cpp:70-76)
    for (int b = 0; b < 2; ++b)
      for (int d = 0; d < 3; ++d) boundaries[b][d] = gv.has(d) ? Metallic : None;
    // polarisation states, newest susceptibility first (src/fields.cpp:262-277)
    for (int i = int(s->sus.size()) - 1; i >= 0; --i) {
      Pol p;
      p.isus = i;
      for (int c = 0; c < NC; ++c) p.P[c][0] = p.P[c][1] = p.Pp[c][0] = p.Pp[c][1] = nullptr;
      pols.push_back(p);
    }
    for (int c = 0; c < NC; ++c) have_p[c] = have_m[c] = false;
    // unit directions are periodic (src/fields.cpp:80-85)
    for (int d = 0; d < 3; ++d)
      if (gv.has(d) && gv.num[d] == 1) use_bloch(d, 0.0);
  }

  bool has_field(int c) const {
    if (gv.dim == 1) return c == 0 || c == 6 || c == 10 || c == 16;
    return component_direction(c) != NODIR;
  }

  // figure_out_step_plan (src/fields.cpp:428-456)
  static bool cross_negative(int a, int b) { return ((3 + b - a) % 3) == 2; }
  static int cross(int a, int b) { return (3 + 2 * a - b) % 3; }
  void figure_out_step_plan() {
    for (int c = 0; c < NC; ++c) have_p[c] = have_m[c] = false;
    for (int c1 = 0; c1 < NC; ++c1) {
      if (!f[c1][0]) continue;
      int dc1 = component_direction(c1);
      for (int c2 = 0; c2 < NC; ++c2)
        if ((is_electric(c1) && is_magnetic(c2)) || (is_D(c1) && is_magnetic(c2)) ||
            (is_magnetic(c1) && is_electric(c2)) || (is_B(c1) && is_electric(c2))) {
          int dc2 = component_direction(c2);
          if (dc2 == NODIR || dc1 == dc2 || !has_field(c2) || !has_field(c1)) continue;
          int dd = cross(dc1, dc2);
          if (!gv.has(dd)) continue;
          if (cross_negative(dc2, dc1)) {
            minus_c[c1] = c2;
            have_m[c1] = true;
            minus_d[c1] = dd;
          }
          else {
            plus_c[c1] = c2;
            have_p[c1] = true;
            plus_d[c1] = dd;
          }
        }
    }
  }

  static bool is_tm(int c) { return c == 5 || c == 6 || c == 15 || c == 16 || c == 4 || c == 14; }
  bool is_like(int c1, int c2) const { return gv.dim != 2 || !(is_tm(c1) ^ is_tm(c2)); }

  // alloc_f (src/fields.cpp:480-504)
  bool alloc_f(int c) {
    bool changed = false;
    for (int cmp = 0; cmp < ncmp(); ++cmp)
      if (!f[c][cmp]) {
        changed = true;
        if (is_magnetic(c)) {
          int bc = direction_component(15, component_direction(c));
          if (!f[bc][cmp]) f[bc][cmp] = alloc(nullptr);
          f[c][cmp] = f[bc][cmp];
        }
        else
          f[c][cmp] = alloc(nullptr);
      }
    return changed;
  }
  void require_component(int c) {
    if (!has_field(c)) fail("cannot require a component absent from this grid");
    components_allocated = true;
    int need = 0;
    for (int ca = 0; ca < NC; ++ca)
      if (has_field(ca) && is_like(c, ca))
        if (alloc_f(ca)) need++;
    if (need) {
      figure_out_step_plan();
      connections_valid = false;
    }
  }

  // ---- boundaries (src/boundaries.cpp) ----
  void set_boundary(int b, int d, int cond) {
    if (boundaries[b][d] != cond) {
      boundaries[b][d] = cond;
      if (cond == Periodic || boundaries[1 - b][d] == Periodic) boundaries[1 - b][d] = cond;
      connections_valid = false;
    }
  }
  void use_bloch(int d, cdouble kk) {
    k[d] = kk;
    for (int b = 0; b < 2; ++b) set_boundary(b, d, Periodic);
    if (real(kk) * gv.num[d] == 0.5 * a)
      eikna[d] = -exp(-imag(kk) * ((2 * pi / a) * gv.num[d]));
    else
      eikna[d] = exp(cdouble(0, 1) * kk * ((2 * pi / a) * gv.num[d]));
    if (is_real && kk != 0.0) fail("Can't use real fields with bloch boundary conditions!");
    connections_valid = false;
  }
  ivec ilattice_vector(int d) const {
    ivec r;
    r.t[d] = 2 * gv.num[d];
    return r;
  }
  bool on_metal_boundary(const ivec &here) const {
    for (int d = 0; d < 3; ++d) {
      if (!gv.has(d)) continue;
      if (here.t[d] == 2 * gv.num[d] && boundaries[High][d] == Metallic) return true;
#ifdef MUT_MAGNETIC
      if (boundaries[Low][d] == Magnetic && here.t[d] == 0) return true;
#else
      if (boundaries[Low][d] == Magnetic && here.t[d] == 1) return true;
#endif
      if (boundaries[Low][d] == Metallic && here.t[d] == 0) return true;
    }
    return false;
  }
  bool locate_point_in_user_volume(ivec *there, cdouble *phase) const {
    if (!gv.owns(*there)) {
      for (int d = 0; d < 3; ++d) {
        if (!gv.has(d)) continue;
        int L = 2 * gv.num[d];
        if (boundaries[High][d] == Periodic && there->t[d] <= 0) {
          while (there->t[d] <= 0) {
            there->t[d] += L;
            *phase *= conj(eikna[d]);
          }
        }
        else if (boundaries[High][d] == Periodic && there->t[d] - L > 0) {
          while (there->t[d] - L > 0) {
            there->t[d] -= L;
            *phase *= eikna[d];
          }
        }
      }
    }
    return gv.owns(*there);
  }
  static int connect_phase_from_phase(cdouble ph) {
    auto close = [](cdouble p, double r) { return fabs(p.imag()) < 1e-13 && fabs(p.real() - r) < 1e-13; };
    return close(ph, 1.0) ? CONNECT_COPY : close(ph, -1.0) ? CONNECT_NEGATE : CONNECT_PHASE;
  }
  // find_metals + connect_the_chunks for one chunk (src/boundaries.cpp:315-345,368-600)
  void connect_chunks() {
    if (connections_valid) return;
    for (int ft = 0; ft < 8; ++ft) {
      conns[ft].clear();
      zeroes[ft].clear();
    }
    // find_metals: owned points on a metal wall
    for (int ft = 0; ft < 4; ++ft)
      for (int cmp = 0; cmp < ncmp(); ++cmp)
        for (int c = 0; c < NC; ++c) {
          if (ftype(c) != ft || !f[c][cmp]) continue;
          loop_ivecs(gv, gv.little_owned_corner(c), gv.big_corner(), [&](ptrdiff_t n, const Loop &L, ptrdiff_t i1, ptrdiff_t i2, ptrdiff_t i3) {
            bool at_b = (L.s[0] != 0 && (i1 == 0 || i1 == L.n[0] - 1)) || (L.s[1] != 0 && (i2 == 0 || i2 == L.n[1] - 1)) ||
                        (L.s[2] != 0 && (i3 == 0 || i3 == L.n[2] - 1));
            if (at_b) {
              ivec here = loop_iloc(L, i1, i2, i3);
              if (on_metal_boundary(here)) zeroes[ft].push_back(f[c][cmp] + n);
            }
          });
        }
    bool B_redundant[5];
    for (int i = 0; i < 5; ++i) B_redundant[i] = f[5 + i][0] == f[15 + i][0];
    for (int corig = 0; corig < NC; ++corig) {
      if (!f[corig][0]) continue;
      for (int ib = 0;; ++ib) {
        ivec cs, ce;
        if (!gv.get_boundary_icorners(corig, ib, &cs, &ce)) break;
        loop_ivecs(gv, cs, ce, [&](ptrdiff_t n, const Loop &L, ptrdiff_t i1, ptrdiff_t i2, ptrdiff_t i3) {
          ivec here = loop_iloc(L, i1, i2, i3);
          int c = corig;
          cdouble ph = 1.0;
          if (!locate_point_in_user_volume(&here, &ph) || on_metal_boundary(here)) return;
          if (is_B(corig) && B_redundant[corig - 15]) return;
          int ip = connect_phase_from_phase(ph);
          ptrdiff_t m = gv.index(c, here);
          Conn cn;
          cn.cls = ip;
          cn.ph = std::complex<R>(R(ph.real()), R(ph.imag()));
          for (int cmp = 0; cmp < 2; ++cmp) {
            cn.dst[cmp] = cmp < ncmp() && f[corig][cmp] ? f[corig][cmp] + n : nullptr;
            cn.src[cmp] = cmp < ncmp() && f[c][cmp] ? f[c][cmp] + m : nullptr;
          }
          conns[ftype(c)].push_back(cn);
          if (is_electric(corig) || is_magnetic(corig)) {
            int fP = is_electric(corig) ? PE_stuff : PH_stuff;
            for (auto &p : pols)
              if (p.allocated && p.P[corig][0]) {
                Conn cp = cn;
                for (int cmp = 0; cmp < 2; ++cmp) {
                  cp.dst[cmp] = cmp < ncmp() && p.P[corig][cmp] ? p.P[corig][cmp] + n : nullptr;
                  cp.src[cmp] = cmp < ncmp() && p.P[c][cmp] ? p.P[c][cmp] + m : nullptr;
                }
                conns[fP].push_back(cp);
              }
          }
        });
      }
    }
    connections_valid = true;
  }
  void step_boundaries(int ft) {
    connect_chunks();
    for (R *z : zeroes[ft]) *z = 0.0;
    for (const Conn &cn : conns[ft]) {
      if (cn.cls == CONNECT_PHASE) {
        std::complex<R> v = cn.ph * std::complex<R>(*cn.src[0], *cn.src[1]);
        *cn.dst[0] = v.real();
        *cn.dst[1] = v.imag();
      }
      else
        for (int cmp = 0; cmp < ncmp(); ++cmp) *cn.dst[cmp] = cn.cls == CONNECT_NEGATE ? -*cn.src[cmp] : *cn.src[cmp];
    }
  }

  // ---- curl ----
  // step_curl (src/step_generic.cpp:65-249): the eight branches {no PML, PML in f, PML via
  // fu, both} x {conductivity cnd / cndinv (with the fcnd auxiliary when PML acts on f), none}
  void step_curl(R *fp, const R *g1, const R *g2, ptrdiff_t s1, ptrdiff_t s2, const ivec &is, const ivec &ie, R dtdx,
                 int dsig, R *fu, int dsigu, const R *cnd, const R *cndinv, R *fcnd) {
    if (!g1) {
      std::swap(g1, g2);
      std::swap(s1, s2);
      dtdx = -dtdx;
    }
    const R *sig = dsig != NODIR ? s->sig[dsig].data() : nullptr, *kap = dsig != NODIR ? s->kap[dsig].data() : nullptr,
            *siginv = dsig != NODIR ? s->siginv[dsig].data() : nullptr;
    const R *sigu = dsigu != NODIR ? s->sig[dsigu].data() : nullptr, *kapu = dsigu != NODIR ? s->kap[dsigu].data() : nullptr,
            *siginvu = dsigu != NODIR ? s->siginv[dsigu].data() : nullptr;
#ifdef MUT_COND_HALF
    const R dt2 = R(dt);
#else
    const R dt2 = R(dt * 0.5);
#endif
    loop_ivecs_par(gv, is, ie, [&](ptrdiff_t i, const Loop &L, ptrdiff_t i1, ptrdiff_t i2, ptrdiff_t i3) {
      ivec here = loop_iloc(L, i1, i2, i3);
      R c = g2 ? (g1[i + s1] - g1[i] + g2[i] - g2[i + s2]) : (g1[i + s1] - g1[i]);
      // The reference breaks the uPML regions off into chunks of their own
      // (src/structure.cpp:128-137, structure::use_pml skipping a chunk that does not meet a
      // direction's uPML, :645-654), so a point is stepped with dsig (dsigu) only where that
      // direction's profile acts on it.  Where it does not (kap = 1, sig = 0) the branches
      // agree on the curl term, but not on what a source or a conductivity has put into f:
      // through f_u or f_cond, f's own increments would bypass the damping.  Each point
      // therefore takes the branch of its region here.
      const int kk = dsig != NODIR ? here.t[dsig] : 0, ku = dsigu != NODIR ? here.t[dsigu] : 0;
#ifdef MUT_PML_CHUNKWIDE
      const bool pml = dsig != NODIR, pmlu = dsigu != NODIR;
#else
      const bool pml = dsig != NODIR && (sig[kk] != 0 || kap[kk] != 1);
      const bool pmlu = dsigu != NODIR && (sigu[ku] != 0 || kapu[ku] != 1);
#endif
      if (cnd) {
        if (!pml && !pmlu)
          fp[i] = ((1 - dt2 * cnd[i]) * fp[i] - dtdx * c) * cndinv[i];
        else if (!pml) {
          R fprev = fu[i];
          fu[i] = ((1 - dt2 * cnd[i]) * fprev - dtdx * c) * cndinv[i];
          fp[i] = siginvu[ku] * ((kapu[ku] - sigu[ku]) * fp[i] + fu[i] - fprev);
        }
        else {
          R *g = pmlu ? fu : fp;
          const R gprev = pmlu ? fu[i] : R(0);
          R fcnd_prev = fcnd[i];
          fcnd[i] = ((1 - dt2 * cnd[i]) * fcnd[i] - dtdx * c) * cndinv[i];
#ifdef MUT_COND_PML
          g[i] = (g[i] + (fcnd[i] - fcnd_prev)) * siginv[kk];
#else
          g[i] = ((kap[kk] - sig[kk]) * g[i] + (fcnd[i] - fcnd_prev)) * siginv[kk];
#endif
          if (pmlu) fp[i] = siginvu[ku] * ((kapu[ku] - sigu[ku]) * fp[i] + fu[i] - gprev);
        }
        return;
      }
      if (!pml) {
        if (!pmlu)
          fp[i] -= dtdx * c;
        else {
          R fprev = fu[i];
          fu[i] -= dtdx * c;
          fp[i] = siginvu[ku] * ((kapu[ku] - sigu[ku]) * fp[i] + fu[i] - fprev);
        }
      }
      else {
        if (!pmlu)
          fp[i] = ((kap[kk] - sig[kk]) * fp[i] - dtdx * c) * siginv[kk];
        else {
          R fprev = fu[i];
          fu[i] = ((kap[kk] - sig[kk]) * fu[i] - dtdx * c) * siginv[kk];
          fp[i] = siginvu[ku] * ((kapu[ku] - sigu[ku]) * fp[i] + fu[i] - fprev);
        }
      }
    });
  }
  // fields_chunk::step_db (src/step_db.cpp:44-146), Cartesian part
  bool step_db(int ft) {
    bool allocated_u = false;
    for (int cmp = 0; cmp < ncmp(); ++cmp)
      for (int cc = first_field_component(ft); cc < first_field_component(ft) + 5; ++cc) {
        if (!f[cc][cmp]) continue;
        int dc = component_direction(cc);
        int dsig0 = cycle_direction(dc, 1), dsigu0 = cycle_direction(dc, 2);
        int dsig = s->sigsize[dsig0] > 1 ? dsig0 : NODIR;
        int dsigu = s->sigsize[dsigu0] > 1 ? dsigu0 : NODIR;
        ptrdiff_t sp = have_p[cc] ? gv.st[plus_d[cc]] : 0, sm = have_m[cc] ? gv.st[minus_d[cc]] : 0;
        const R *fpp = have_p[cc] ? f[plus_c[cc]][cmp] : nullptr;
        const R *fmm = have_m[cc] ? f[minus_c[cc]][cmp] : nullptr;
        const R *cnd = s->cond[cc].empty() ? nullptr : s->cond[cc].data();
        const R *cndinv = s->condinv[cc].empty() ? nullptr : s->condinv[cc].data();
        if (dsig != NODIR && cnd && !f_cond[cc][cmp]) f_cond[cc][cmp] = alloc(nullptr);
        if (dsigu != NODIR && !f_u[cc][cmp]) {
          f_u[cc][cmp] = alloc(f[cc][cmp]);
          allocated_u = true;
        }
        if (ft == D_stuff) {
          sp = -sp;
          sm = -sm;
        }
        step_curl(f[cc][cmp], fpp, fmm, sp, sm, gv.little_owned_corner(cc), gv.big_corner(), R(courant), dsig, f_u[cc][cmp], dsigu,
                  cnd, cndinv, f_cond[cc][cmp]);
      }
    return allocated_u;
  }
  // fields_chunk::step_source (src/step.cpp:295-318)
  void step_source(int ft) {
    for (const SrcVol &sv : sources[ft]) {
      int c = direction_component(first_field_component(ft), component_direction(sv.c));
      const SrcTime &st = srcs[sv.src];
      const R *cndinv = s->condinv[c].empty() ? nullptr : s->condinv[c].data();
#ifdef MUT_COND_SRC
      cndinv = nullptr;
#endif
      if (!st.is_integrated && f[c][0] && ((ft == D_stuff && is_electric(sv.c)) || (ft == B_stuff && is_magnetic(sv.c))))
        for (size_t j = 0; j < sv.index.size(); ++j) {
          ptrdiff_t i = sv.index[j];
#ifdef MUT_SRC_DT
          cdouble A = sv.amp[j] * st.cur_current;
#else
          cdouble A = sv.amp[j] * st.cur_current * dt;
#endif
          if (cndinv) A = sv.amp[j] * st.cur_current * dt * double(cndinv[i]);
          f[c][0][i] -= real(A);
          if (!is_real) f[c][1][i] -= imag(A);
        }
    }
  }
  void calc_sources(double tim) {
    for (SrcTime &st : srcs) st.update(tim, dt);
  }

  bool needs_P(const Sus<R> &su, int c, int cmp) const {
    if (!is_electric(c) && !is_magnetic(c)) return false;
    int d = component_direction(c);
    return !su.trivial[c] && f[direction_component(c, d)][cmp];
  }
  // fields_chunk::update_eh (src/update_eh.cpp:66-215), diagonal chi1inv, no chi2/chi3
  bool update_eh(int ft) {
    int ft2 = ft == E_stuff ? D_stuff : B_stuff;
    bool allocated_eh = false;
    bool have_int = false;
    for (const SrcVol &sv : sources[ft2])
      if (srcs[sv.src].is_integrated) have_int = true;
    int c0 = first_field_component(ft), c20 = first_field_component(ft2);
    for (int ec = c0; ec < c0 + 5; ++ec) {
      int dc = c20 + (ec - c0);
      for (int cmp = 0; cmp < ncmp(); ++cmp) {
        bool need = false;
        if (f[ec][cmp]) {
          need = have_int;
          for (auto &p : pols) need = need || needs_P(s->sus[p.isus], ec, cmp);
        }
        if (need) {
          if (!fmp[dc][cmp]) fmp[dc][cmp] = alloc(nullptr);
        }
        else
          fmp[dc][cmp] = nullptr;
      }
    }
    bool have_fmp = false;
    for (int dc = c20; dc < c20 + 5; ++dc)
      if (fmp[dc][0]) have_fmp = true;
    const size_t ntot = gv.ntot;
    for (int ec = c0; ec < c0 + 5; ++ec)
      if (f[ec][0]) {
        int dc = c20 + (ec - c0);
        for (int cmp = 0; cmp < ncmp(); ++cmp)
          if (fmp[dc][cmp]) memcpy(fmp[dc][cmp], f[dc][cmp], sizeof(R) * ntot);
      }
#ifdef MUT_HPOLE_SUB
    for (auto &p : pols)
      if (p.allocated && s->sus[p.isus].ft == ft && ft == E_stuff)
#else
    for (auto &p : pols)
      if (p.allocated && s->sus[p.isus].ft == ft) // subtract_P (src/susceptibility.cpp:264-281)
#endif
          for (int ec = c0; ec < c0 + 5; ++ec)
            for (int cmp = 0; cmp < 2; ++cmp)
              if (p.P[ec][cmp]) {
                int dc = c20 + (ec - c0);
                if (fmp[dc][cmp])
                  for (size_t i = 0; i < ntot; ++i) fmp[dc][cmp][i] -= p.P[ec][cmp][i];
              }
    if (have_fmp)
      for (const SrcVol &sv : sources[ft2]) {
        const SrcTime &st = srcs[sv.src];
        if (st.is_integrated && f[sv.c][0] && ft == ftype(sv.c)) {
          int c = direction_component(first_field_component(ft2), component_direction(sv.c));
          for (size_t j = 0; j < sv.index.size(); ++j) {
            cdouble A = sv.amp[j] * st.cur_dipole;
            for (int cmp = 0; cmp < ncmp(); ++cmp) fmp[c][cmp][sv.index[j]] -= cmp ? imag(A) : real(A);
          }
        }
      }
    for (int cmp = 0; cmp < ncmp(); ++cmp)
      for (int ec = c0; ec < c0 + 5; ++ec) {
        if (!f[ec][cmp]) continue;
        int dc = c20 + (ec - c0);
        int d_ec = component_direction(ec);
        int dsigw = s->sigsize[d_ec] > 1 ? d_ec : NODIR;
        const R *u = s->chi1inv[ec].empty() ? nullptr : s->chi1inv[ec].data();
        if (f[ec][cmp] == f[dc][cmp] && (u || have_fmp || dsigw != NODIR)) {
          f[ec][cmp] = alloc(f[dc][cmp]);
          allocated_eh = true;
        }
        if (!f_w[ec][cmp] && dsigw != NODIR) f_w[ec][cmp] = alloc(f[ec][cmp]);
        if (f[ec][cmp] == f[dc][cmp]) continue;
        const R *g = fmp[dc][cmp] ? fmp[dc][cmp] : f[dc][cmp];
        R *fe = f[ec][cmp], *fw = f_w[ec][cmp];
        const R *sigw = dsigw != NODIR ? s->sig[dsigw].data() : nullptr, *kapw = dsigw != NODIR ? s->kap[dsigw].data() : nullptr;
        // step_update_EDHB diagonal branches (src/step_generic.cpp:682-701,774-782)
        loop_ivecs_par(gv, gv.little_owned_corner(ec), gv.big_corner(), [&](ptrdiff_t i, const Loop &L, ptrdiff_t i1, ptrdiff_t i2, ptrdiff_t i3) {
          if (dsigw != NODIR) {
            int kw = loop_iloc(L, i1, i2, i3).t[dsigw];
            R fwprev = fw[i], kapwkw = kapw[kw], sigwkw = sigw[kw];
            fw[i] = u ? (g[i] * u[i]) : g[i];
            fe[i] += (kapwkw + sigwkw) * fw[i] - (kapwkw - sigwkw) * fwprev;
          }
          else
            fe[i] = u ? (g[i] * u[i]) : g[i];
        });
      }
    return allocated_eh;
  }
  // fields_chunk::update_pols (src/update_pols.cpp:40-62) + lorentzian update_P, for the
  // susceptibilities of field type ft (the chiP[ft] list)
  bool update_pols(int ft) {
    bool allocated = false;
    for (auto &p : pols) {
      const Sus<R> &su = s->sus[p.isus];
      if (su.ft != ft) continue;
      if (!p.allocated) {
        for (int c = 0; c < NC; ++c)
          for (int cmp = 0; cmp < 2; ++cmp)
            if (needs_P(su, c, cmp)) {
              p.P[c][cmp] = alloc(nullptr);
              p.Pp[c][cmp] = alloc(nullptr);
            }
        p.allocated = true;
        allocated = true;
      }
      const R omega2pi = 2 * pi * R(su.omega_0), g2pi = R(su.gamma) * 2 * pi;
      const R rdt = R(dt);
      const R omega0dtsqr = omega2pi * omega2pi * rdt * rdt;
#ifdef MUT_GAMMA
      (void)g2pi;
      const R gamma1inv = 1 / (1 + R(su.gamma) * rdt), gamma1 = (1 - R(su.gamma) * rdt);
#else
      const R gamma1inv = 1 / (1 + g2pi * rdt / 2), gamma1 = (1 - g2pi * rdt / 2);
#endif
#ifdef MUT_DRUDE
      const R omega0dtsqr_denom = omega0dtsqr;
#else
      const R omega0dtsqr_denom = su.drude ? 0 : omega0dtsqr;
#endif
      for (int c = 0; c < NC; ++c)
        for (int cmp = 0; cmp < 2; ++cmp) {
          if (!p.P[c][cmp]) continue;
          const R *w = f_w[c][cmp] ? f_w[c][cmp] : f[c][cmp];
          const R *sg = su.sigma[c].empty() ? nullptr : su.sigma[c].data();
          if (!w || !sg) continue;
          R *pp = p.P[c][cmp], *ppr = p.Pp[c][cmp];
          loop_ivecs_par(gv, gv.little_owned_corner(c), gv.big_corner(), [&](ptrdiff_t i, const Loop &, ptrdiff_t, ptrdiff_t, ptrdiff_t) {
            R pcur = pp[i];
            pp[i] = gamma1inv * (pcur * (2 - omega0dtsqr_denom) - gamma1 * ppr[i] + omega0dtsqr * (sg[i] * w[i]));
            ppr[i] = pcur;
          });
        }
    }
    return allocated;
  }

  // ---- DFT ----
  void update_dfts() {
#ifdef MUT_DFT_HTIME
    double timeE = t * dt, timeH = t * dt;
#else
    double timeE = t * dt, timeH = t * dt - 0.5 * dt;
#endif
    for (auto &cp : dft_chunks) {
      DftChunk<R> &ch = *cp;
      if (t % ch.decimation) continue;
      if (!f[ch.c][0]) continue;
      double time = (is_magnetic(ch.c) || is_B(ch.c)) ? timeH : timeE;
      const int Nw = int(ch.omega.size());
      std::vector<std::complex<R>> ph(Nw);
      std::vector<cdouble> ph64(Nw);
      for (int i = 0; i < Nw; ++i) {
#ifdef MUT_DFT_SIGN
        ph64[i] = std::polar(1.0, -ch.omega[i] * time) * ch.scale;
#else
        ph64[i] = std::polar(1.0, ch.omega[i] * time) * ch.scale;
#endif
        ph[i] = std::complex<R>(ph64[i]);
      }
      int numcmp = f[ch.c][1] ? 2 : 1;
      loop_ivecs_par(gv, ch.is, ch.ie, [&](ptrdiff_t idx, const Loop &L, ptrdiff_t i1, ptrdiff_t i2, ptrdiff_t i3) {
        size_t idft = (i1 * L.n[1] + i2) * L.n[2] + i3;
#ifdef MUT_FLUX_WEIGHT
        double w = 1.0;
#else
        double w = ch.include_w ? loop_weight(gv, L, i1, i2, i3, ch.s0, ch.s1, ch.e0, ch.e1, ch.dV0 + ch.dV1 * i2) : 1.0;
        if (ch.sqrt_w) w = std::sqrt(w); // src/dft.cpp:277-280
#endif
        R fv[2] = {0, 0};
        for (int cmp = 0; cmp < numcmp; ++cmp) {
          const R *F = f[ch.c][cmp];
          if (ch.avg2)
            fv[cmp] = (w * 0.25) * (F[idx] + F[idx + ch.avg1] + F[idx + ch.avg2] + F[idx + (ch.avg1 + ch.avg2)]);
          else if (ch.avg1)
            fv[cmp] = (w * 0.5) * (F[idx] + F[idx + ch.avg1]);
          else
            fv[cmp] = w * F[idx];
        }
        for (int i = 0; i < Nw; ++i) {
          if (ch.acc == MP_DFT_F64) {
            ch.dft64[Nw * idft + i] += ph64[i] * cdouble(double(fv[0]), double(numcmp == 2 ? fv[1] : 0));
          }
          else if (numcmp == 2)
            ch.dft[Nw * idft + i] += ph[i] * std::complex<R>(fv[0], fv[1]);
          else
            ch.dft[Nw * idft + i] += std::complex<R>{fv[0] * ph[i].real(), fv[0] * ph[i].imag()};
        }
      });
    }
  }

  // ---- point probes (src/monitor.cpp:93-130) ----
  cdouble get_field_iloc(int c, const ivec &orig) const {
    ivec il = orig;
    cdouble kph = 1.0;
    locate_point_in_user_volume(&il, &kph);
    if (gv.owns(il)) {
      if (!f[c][0]) return kph * 0.0;
      ptrdiff_t i = gv.index(c, il);
      cdouble v = f[c][1] ? cdouble(f[c][0][i], f[c][1][i]) : cdouble(f[c][0][i], 0);
      return kph * v;
    }
    return 0.0;
  }
  // fields::get_eps / get_mu at frequency 0 (src/monitor.cpp:135-195): the number of electric
  // (magnetic) components over the sum of their chi1inv interpolated at p, each point located
  // in the cell, vacuum outside it
  double get_eps_mu(const vec3 &p, bool magnetic) const {
    double tr = 0;
    int nc = 0;
    for (int c = magnetic ? 5 : 0; c < (magnetic ? 10 : 5); ++c) {
      if (!has_field(c)) continue;
      ivec locs[8];
      double w[8];
      gv.interpolate(c, p, locs, w);
      double v = 0;
      for (int q = 0; q < 8 && w[q] != 0; ++q) {
        ivec il = locs[q];
        cdouble ph = 1.0;
        locate_point_in_user_volume(&il, &ph);
        double u = 1.0;
        if (gv.owns(il)) u = s->chi1inv[c].empty() ? 1.0 : double(s->chi1inv[c][gv.index(c, il)]);
        v += w[q] * u;
      }
      tr += v;
      ++nc;
    }
    return nc / tr;
  }
  cdouble get_field(int c, const vec3 &p) const {
    ivec locs[8];
    double w[8];
    gv.interpolate(c, p, locs, w);
    cdouble res = 0.0;
    for (int i = 0; i < 8 && w[i]; ++i) res += w[i] * get_field_iloc(c, locs[i]);
    return res;
  }

  // ---- fields::step (src/step.cpp:35-139) ----
  // wall time per phase under the reference's time sinks (src/meep.hpp:1595-1618, the
  // timing scopes of src/step.cpp:66-122): 0 FieldUpdateB, 1 FieldUpdateH, 2 FieldUpdateD,
  // 3 FieldUpdateE (with the polarisations), 4 BoundarySteppingB..E, 5 FourierTransforming,
  // 6 the whole step
  static double now() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }
  void step() {
    const double t0 = now();
    double ta = t0, tb;
    auto sink = [&](int k) {
      tb = now();
      times[k] += tb - ta;
      ta = tb;
    };
    update_condinv(*s); // src/step.cpp:60-62
    calc_sources(t * dt);
    if (step_db(B_stuff)) connections_valid = false;
    step_source(B_stuff);
    sink(0);
    step_boundaries(B_stuff);
    sink(4);
    calc_sources(t * dt + 0.5 * dt);
    if (update_eh(H_stuff)) connections_valid = false;
    sink(1);
    step_boundaries(WH_stuff);
    sink(4);
    if (update_pols(H_stuff)) connections_valid = false;
    sink(1);
    step_boundaries(PH_stuff);
    step_boundaries(H_stuff);
    sink(4);
    for (auto &fv : flux_vols) // flux_vol::update_half (src/meep.hpp:2330-2333, src/step.cpp:93)
      fv->cur_flux_half = flux_in_box_wrongH(fv->d, fv->wmin, fv->wmax);
    calc_sources(t * dt + 0.5 * dt);
    if (step_db(D_stuff)) connections_valid = false;
    step_source(D_stuff);
    sink(2);
    step_boundaries(D_stuff);
    sink(4);
    calc_sources(t * dt + dt);
    if (update_eh(E_stuff)) connections_valid = false;
    sink(3);
    step_boundaries(WE_stuff);
    sink(4);
    if (update_pols(E_stuff)) connections_valid = false;
    sink(3);
    step_boundaries(PE_stuff);
    step_boundaries(E_stuff);
    sink(4);
    for (auto &fv : flux_vols) // flux_vol::update (src/meep.hpp:2334-2337, src/step.cpp:124)
#ifdef MUT_FLUXT_HALF
      fv->cur_flux = flux_in_box_wrongH(fv->d, fv->wmin, fv->wmax);
#else
      fv->cur_flux = (flux_in_box_wrongH(fv->d, fv->wmin, fv->wmax) + fv->cur_flux_half) * 0.5;
#endif
    t += 1;
    update_dfts();
    sink(5);
    // get_field(D_EnergyDensity, gv.center()) must be finite (src/step.cpp:137-138)
    vec3 cen = gv.loc(gv.icenter());
    double sum = 0;
    for (int c = 0; c < 5; ++c)
      if (has_field(c) && component_direction(c) != NODIR)
        sum += real(conj(get_field(c, cen)) * get_field(c + 10, cen));
    if (!std::isfinite(sum)) fail("simulation fields are NaN or Inf");
    times[6] += now() - t0;
  }

  // ---- field energy (src/energy_and_flux.cpp:45-178, src/integrate.cpp:46-205) ----
  // 0.5 * integral of Re(conj(F) G) over [wmin, wmax] for F = E (H) and G = D (B), per
  // component (field_energy_in_box(c, where), :67-82); E and D share a grid, so the
  // integration grid is the component's own.  Summed over components in long double (:84-95).
  long double energy_in_box(int first, const vec3 &wmin, const vec3 &wmax) {
    long double sum = 0;
    for (int c = first; c < first + 5; ++c) {
      if (!has_field(c) || !f[c][0]) continue;
      sum += real(integrate_dot(c, c + 10, wmin, wmax)) * 0.5;
    }
    return sum;
  }
  long double energy_of(int first) {
    vec3 wmin, wmax;
    for (int d = 0; d < 3; ++d) wmax.t[d] = gv.has(d) ? gv.num[d] * gv.inva : 0.0;
    return energy_in_box(first, wmin, wmax);
  }
  // fields::field_energy_in_box (src/energy_and_flux.cpp:54-59)
  double field_energy_in_box(const vec3 &wmin, const vec3 &wmax) {
    const double magnetic = with_synchronized_magnetic([&] { return double(energy_in_box(5, wmin, wmax)); });
    return double(energy_in_box(0, wmin, wmax)) + magnetic;
  }
  // fields::integrate (src/integrate.cpp:46-120,122-180) of dot_integrand
  // (src/energy_and_flux.cpp:61-65), Re(conj(F0) F1), over [wmin, wmax]: on the two
  // components' own grid when they share it, else on the centred grid with each field the
  // mean of its four Yee points (a missing component reads as zero); each piece of the volume
  // summed in long double, the pieces in double.
  cdouble integrate_dot(int c0, int c1, const vec3 &wmin, const vec3 &wmax) {
    const ivec y0 = gv.iyee_shift(c0), y1 = gv.iyee_shift(c1);
    bool same = true;
    for (int d = 0; d < 3; ++d)
      if (y0.t[d] != y1.t[d]) same = false;
    const int cgrid = same ? c0 : 20;
    const int cs[2] = {c0, c1};
    cdouble total = 0;
    loop_in_chunks(wmin, wmax, cgrid, [&](const ivec &is, const ivec &ie, const vec3 &s0, const vec3 &s1, const vec3 &e0,
                                          const vec3 &e1, double dV0, const ivec &, cdouble sph) {
      ptrdiff_t off[2][2] = {{0, 0}, {0, 0}};
#ifndef MUT_FLUXT_CENTRE
      if (cgrid == 20)
        for (int q = 0; q < 2; ++q) gv.yee2cent_offsets(cs[q], off[q][0], off[q][1]);
#endif
      std::complex<long double> part = 0;
      loop_ivecs(gv, is, ie, [&](ptrdiff_t idx, const Loop &L, ptrdiff_t i1, ptrdiff_t i2, ptrdiff_t i3) {
        std::complex<R> fv[2];
        for (int q = 0; q < 2; ++q) {
          double fr[2] = {0, 0};
          for (int k = 0; k < 2; ++k)
            if (f[cs[q]][k]) {
              const R *p = f[cs[q]][k];
              fr[k] = 0.25 * (p[idx] + p[idx + off[q][0]] + p[idx + off[q][1]] + p[idx + off[q][0] + off[q][1]]);
            }
          fv[q] = std::complex<R>(cdouble(fr[0], fr[1]) * sph);
        }
        const cdouble integrand = real(conj(cdouble(fv[0])) * cdouble(fv[1]));
        part += std::complex<long double>(integrand * loop_weight(gv, L, i1, i2, i3, s0, s1, e0, e1, dV0));
      });
      total += cdouble(part);
    });
    return total;
  }
  // fields::flux_in_box_wrongH (src/energy_and_flux.cpp:186-221): the flux of E x H through
  // [wmin, wmax] along d with E and H as they stand (H half a step behind E)
  double flux_in_box_wrongH(int d, const vec3 &wmin, const vec3 &wmax) {
    int cE[2], cH[2];
    if (d == X)
      cE[0] = 1, cE[1] = 4, cH[0] = 9, cH[1] = 6;
    else if (d == Y)
      cE[0] = 4, cE[1] = 0, cH[0] = 5, cH[1] = 9;
    else
      cE[0] = 0, cE[1] = 1, cH[0] = 6, cH[1] = 5;
    long double sum = 0;
#ifdef MUT_FLUXT_SIGN
    for (int i = 0; i < 2; ++i) sum += real(integrate_dot(cE[i], cH[i], wmin, wmax));
#else
    for (int i = 0; i < 2; ++i) sum += real(integrate_dot(cE[i], cH[i], wmin, wmax)) * (1 - 2 * i);
#endif
    return double(sum);
  }
  // fields::flux_in_box (src/energy_and_flux.cpp:223-228): the same with H synchronised
  double flux_in_box(int d, const vec3 &wmin, const vec3 &wmax) {
    return with_synchronized_magnetic([&] { return flux_in_box_wrongH(d, wmin, wmax); });
  }
  struct FluxVol { // flux_vol (src/meep.hpp:2322-2352)
    int d;
    vec3 wmin, wmax;
    double cur_flux = 0, cur_flux_half = 0;
  };
  std::vector<std::unique_ptr<FluxVol>> flux_vols;

  // synchronize_magnetic_fields / restore_magnetic_fields (src/energy_and_flux.cpp:146-178)
  // around fn: B and H advanced half a step and averaged with their values at t - dt/2, fn
  // evaluated, every magnetic array (and its uPML and conductivity auxiliaries) restored.
  template <class FN> double with_synchronized_magnetic(FN fn) {
    std::vector<std::pair<R *, std::vector<R>>> saved, saved_aux;
    auto save = [&](std::vector<std::pair<R *, std::vector<R>>> &into, R *p) {
      if (p) into.push_back({p, std::vector<R>(p, p + gv.ntot)});
    };
    for (int cmp = 0; cmp < 2; ++cmp) {
      for (int c = 15; c < 20; ++c) {
        save(saved, f[c][cmp]);
        save(saved_aux, f_u[c][cmp]);
        save(saved_aux, f_w[c][cmp]);
        save(saved_aux, f_cond[c][cmp]);
      }
      for (int c = 5; c < 10; ++c)
        if (f[c][cmp] && f[c][cmp] != f[c + 10][cmp]) {
          save(saved, f[c][cmp]);
          save(saved_aux, f_u[c][cmp]);
          save(saved_aux, f_w[c][cmp]);
        }
    }
    update_condinv(*s);
    calc_sources(t * dt);
    if (step_db(B_stuff)) connections_valid = false;
    step_source(B_stuff);
    step_boundaries(B_stuff);
    calc_sources(t * dt + 0.5 * dt);
    if (update_eh(H_stuff)) connections_valid = false;
    step_boundaries(H_stuff);
    for (auto &sv : saved) {
      R *fc = sv.first;
      for (size_t i = 0; i < gv.ntot; ++i) fc[i] = 0.5 * (fc[i] + sv.second[i]);
    }
    const double value = fn();
    for (auto *lst : {&saved, &saved_aux})
      for (auto &sv : *lst) std::copy(sv.second.begin(), sv.second.end(), sv.first);
    return value;
  }
  // fields::field_energy (src/energy_and_flux.cpp:46-62): the magnetic part with H
  // synchronised, the electric part as it stands
  double field_energy() {
    const double magnetic = with_synchronized_magnetic([&] { return double(energy_of(5)); });
    return double(energy_of(0)) + magnetic;
  }

  // ---- loop_in_chunks for one chunk and the identity symmetry (src/loop_in_chunks.cpp) ----
  template <class CB> void loop_in_chunks(const vec3 &wmin, const vec3 &wmax, int cgrid, CB chunkloop) {
    // cgrid == 20 means Centered
    ivec iyc = gv.iyee_shift(20), iyg = gv.iyee_shift(cgrid), iyee_c;
    for (int d = 0; d < 3; ++d) iyee_c.t[d] = iyc.t[d] - iyg.t[d];
    vec3 yc = gv.loc(iyee_c);
    vec3 wcmin, wcmax;
    for (int d = 0; d < 3; ++d) {
      wcmin.t[d] = wmin.t[d] + yc.t[d];
      wcmax.t[d] = wmax.t[d] + yc.t[d];
    }
    ivec is, ie;
    for (int d = 0; d < 3; ++d)
      if (gv.has(d)) {
        int fl = 1 + 2 * int(floor(wcmin.t[d] * a - .5));
        int ce = 1 + 2 * int(ceil(wcmax.t[d] * a - .5));
        is.t[d] = fl - iyee_c.t[d];
        ie.t[d] = ce - iyee_c.t[d];
      }
    // compute_boundary_weights (src/loop_in_chunks.cpp:253-295), snap_empty_dimensions = false
    vec3 s0, s1, e0, e1;
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
        fail("bug: impossible(?) looping boundaries");
    }
    // lattice shifts for periodic directions
    ivec min_ish, max_ish;
    for (int d = 0; d < 3; ++d)
      if (gv.has(d) && boundaries[High][d] == Periodic) {
        double L = gv.num[d] * gv.inva;
        double gmin = 0, gmax = gv.num[d] * (2 * 0.5 * gv.inva);
        min_ish.t[d] = int(floor((wmin.t[d] - gmax) / L));
        max_ish.t[d] = int(ceil((wmax.t[d] - gmin) / L));
      }
    ivec ish = min_ish;
    ivec isco = gv.little_owned_corner(cgrid), ieco = gv.big_owned_corner(cgrid);
    do {
      cdouble ph = 1.0;
      ivec shifti;
      for (int d = 0; d < 3; ++d)
        if (gv.has(d)) {
          shifti.t[d] = 2 * gv.num[d] * ish.t[d];
          ph *= std::pow(eikna[d], ish.t[d]);
        }
      ivec iscS, iecS;
      bool nonempty = true;
      for (int d = 0; d < 3; ++d)
        if (gv.has(d)) {
          iscS.t[d] = std::max(is.t[d] - shifti.t[d], isco.t[d]);
          iecS.t[d] = std::min(ie.t[d] - shifti.t[d], ieco.t[d]);
          if (iscS.t[d] > iecS.t[d]) nonempty = false;
        }
      if (nonempty) {
        ivec isc = iscS, iec = iecS;
        vec3 s0c, s1c, e0c, e1c;
        for (int d = 0; d < 3; ++d) s0c.t[d] = s1c.t[d] = e0c.t[d] = e1c.t[d] = 1.0;
        for (int d = 0; d < 3; ++d) {
          if (!gv.has(d)) continue;
          int a0 = iscS.t[d] + shifti.t[d], b0 = iecS.t[d] + shifti.t[d];
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
          else if (b0 == a0 + 1 * 2) {
            double w = std::min(s0c.t[d], e1c.t[d]);
            s0c.t[d] = e1c.t[d] = w;
            w = std::min(s1c.t[d], e0c.t[d]);
            s1c.t[d] = e0c.t[d] = w;
          }
          else if (b0 == a0 + 2 * 2) {
            double w = std::min(s1c.t[d], e1c.t[d]);
            s1c.t[d] = e1c.t[d] = w;
          }
        }
        double dV0 = 1.0;
        for (int d = 0; d < 3; ++d)
          if (gv.has(d) && wmax.t[d] - wmin.t[d] > 0.0) dV0 *= gv.inva;
        chunkloop(isc, iec, s0c, s1c, e0c, e1c, dV0, shifti, ph);
      }
      bool carried = true;
      for (int d = 0; d < 3; ++d) {
        if (!gv.has(d)) continue;
        if (ish.t[d] + 1 <= max_ish.t[d]) {
          ish.t[d] += 1;
          carried = false;
          break;
        }
        ish.t[d] = min_ish.t[d];
      }
      if (carried) break;
    } while (true);
  }

  int add_src_time(const SrcTime &st) {
    for (size_t i = 0; i < srcs.size(); ++i)
      if (srcs[i].equal(st)) return int(i);
    srcs.push_back(st);
    return int(srcs.size()) - 1;
  }

  // fields::add_volume_source (src/sources.cpp:462-490) + src_vol_chunkloop (:243-312)
  void add_volume_source(int c, const SrcTime &st, vec3 wmin, vec3 wmax, cdouble amp, mp_complex_func A, void *Au) {
    if (!valid_cart(c) || !has_field(c)) fail("invalid source component");
    for (int d = 0; d < 3; ++d) {
      if (!gv.has(d)) continue;
      double lo = 0, hi = gv.num[d] * gv.inva; // boundary_location(Low/High)
      double w = hi - lo;
      double wd = wmax.t[d] - wmin.t[d];
      if (wd > w + gv.inva)
        fail("Source width > cell width");
      else if (wd > w) {
        double dw = wd - w;
        wmin.t[d] = wmin.t[d] - dw * 0.5;
        wmax.t[d] = wmin.t[d] + w;
      }
    }
    cdouble damp = amp;
    for (int d = 0; d < 3; ++d)
      if (gv.has(d) && wmax.t[d] - wmin.t[d] == 0.0) damp *= a;
    int isrc = add_src_time(st);
    vec3 center;
    for (int d = 0; d < 3; ++d) center.t[d] = (wmin.t[d] + wmax.t[d]) * 0.5;
    loop_in_chunks(wmin, wmax, c, [&](const ivec &is, const ivec &ie, const vec3 &s0, const vec3 &s1, const vec3 &e0, const vec3 &e1,
                                      double, const ivec &shift, cdouble shift_phase) {
      std::vector<ptrdiff_t> index;
      std::vector<cdouble> amps;
      cdouble am = damp * conj(shift_phase);
      loop_ivecs(gv, is, ie, [&](ptrdiff_t idx, const Loop &L, ptrdiff_t i1, ptrdiff_t i2, ptrdiff_t i3) {
        ivec iloc = loop_iloc(L, i1, i2, i3);
        if (!gv.owns(iloc)) return;
        vec3 loc;
        for (int d = 0; d < 3; ++d)
          if (gv.has(d)) {
            int kk = 0;
            for (int q = 0; q < 3; ++q)
              if (L.d[q] == d) kk = q;
            ptrdiff_t ii[3] = {i1, i2, i3};
            loc.t[d] = (0.5 * L.is[kk] + ii[kk]) * gv.inva + shift.t[d] * (0.5 * gv.inva);
          }
        double rel[3];
        for (int d = 0; d < 3; ++d) rel[d] = loc.t[d] - center.t[d];
        cdouble Av = 1.0;
        if (A) {
          double re = 1, im = 0;
          A(rel, Au, &re, &im);
          Av = cdouble(re, im);
        }
        amps.push_back(loop_weight(gv, L, i1, i2, i3, s0, s1, e0, e1, 1) * am * Av);
        index.push_back(idx);
      });
      int ft = (is_magnetic(c) || is_B(c)) ? B_stuff : D_stuff;
      int cc = is_D(c) ? c - 10 : is_B(c) ? c - 10 : c;
      for (SrcVol &o : sources[ft])
        if (o.c == cc && o.src == isrc && o.index == index) {
          for (size_t j = 0; j < o.amp.size(); ++j) o.amp[j] += amps[j];
          return;
        }
      sources[ft].push_back(SrcVol{cc, isrc, std::move(index), std::move(amps)});
    });
    require_component(c);
  }

  // fields::add_volume_source with an amplitude array (src/sources.cpp:394-417): the
  // amplitude at a point is amp_file_func of its position relative to the volume's centre
  // (src/sources.cpp:345-376): each direction the volume spans mapped to [0, 1], then
  // linear_interpolate over the row-major n0 x n1 x n2 array with mirror boundaries
  // (src/fields.cpp:754-810)
  struct AmpArray {
    const double *v; // interleaved (re, im)
    int n[3];
    double size[3];
  };
  static int mirrorindex(int i, int n) { return i >= n ? 2 * n - 1 - i : (i < 0 ? -1 - i : i); }
  static double linear_interpolate(double rx, double ry, double rz, const double *data, int nx, int ny, int nz, int stride) {
    rx = rx < 0.0 ? -rx : (rx > 1.0 ? 1.0 - rx : rx);
    ry = ry < 0.0 ? -ry : (ry > 1.0 ? 1.0 - ry : ry);
    rz = rz < 0.0 ? -rz : (rz > 1.0 ? 1.0 - rz : rz);
    const int x1 = mirrorindex(int(rx * nx), nx), y1 = mirrorindex(int(ry * ny), ny), z1 = mirrorindex(int(rz * nz), nz);
#ifdef MUT_AMPARR_OFFSET
    double dx = rx * nx - x1, dy = ry * ny - y1, dz = rz * nz - z1;
#else
    double dx = rx * nx - x1 - 0.5, dy = ry * ny - y1 - 0.5, dz = rz * nz - z1 - 0.5;
#endif
    const int x2 = mirrorindex(dx >= 0.0 ? x1 + 1 : x1 - 1, nx), y2 = mirrorindex(dy >= 0.0 ? y1 + 1 : y1 - 1, ny),
              z2 = mirrorindex(dz >= 0.0 ? z1 + 1 : z1 - 1, nz);
    dx = fabs(dx);
    dy = fabs(dy);
    dz = fabs(dz);
    auto D = [&](int x, int y, int z) { return data[(((size_t(x) * ny + y) * nz) + z) * stride]; };
    return (((D(x1, y1, z1) * (1.0 - dx) + D(x2, y1, z1) * dx) * (1.0 - dy) + (D(x1, y2, z1) * (1.0 - dx) + D(x2, y2, z1) * dx) * dy) *
                (1.0 - dz) +
            ((D(x1, y1, z2) * (1.0 - dx) + D(x2, y1, z2) * dx) * (1.0 - dy) + (D(x1, y2, z2) * (1.0 - dx) + D(x2, y2, z2) * dx) * dy) *
                dz);
  }
  static void amp_file_func(const double *rel, void *user, double *re, double *im) {
    const AmpArray &A = *static_cast<const AmpArray *>(user);
    double r[3];
    for (int d = 0; d < 3; ++d) r[d] = A.size[d] == 0 ? 0 : 0.5 + rel[d] / A.size[d];
    *re = linear_interpolate(r[0], r[1], r[2], A.v, A.n[0], A.n[1], A.n[2], 2);
    *im = linear_interpolate(r[0], r[1], r[2], A.v + 1, A.n[0], A.n[1], A.n[2], 2);
  }
  void add_volume_source_array(int c, const SrcTime &st, vec3 wmin, vec3 wmax, const double *arr, const size_t *dims, cdouble amp) {
    AmpArray A;
    A.v = arr;
    for (int k = 0; k < 3; ++k) A.n[k] = int(dims[k]);
    // the volume's extent in x, y, z as amp_file_func reads it: Z alone in 1D, X and Y in 2D
    for (int d = 0; d < 3; ++d) A.size[d] = gv.has(d) ? wmax.t[d] - wmin.t[d] : 0.0;
    add_volume_source(c, st, wmin, wmax, amp, amp_file_func, &A);
  }

  // fields::add_srcdata (src/sources.cpp:314-334): a source on given points of component c's
  // array with the given amplitudes; a list marked needs_boundary_fix waits for
  // require_source_components
  std::vector<std::pair<int, size_t>> boundary_fix; // (field type, list)
  void add_srcdata(int c, const SrcTime &st, std::vector<ptrdiff_t> index, std::vector<cdouble> amp, bool fix) {
    if (c < 0 || c >= NC || !(is_electric(c) || is_magnetic(c)) || component_direction(c) == NODIR || !has_field(c))
      fail("add_srcdata: invalid component %d", c);
    for (ptrdiff_t i : index)
      if (i < 0 || size_t(i) >= gv.ntot) fail("add_srcdata: point outside the array");
    const int isrc = add_src_time(st);
    const int ft = is_magnetic(c) ? B_stuff : D_stuff;
    sources[ft].push_back(SrcVol{c, isrc, std::move(index), std::move(amp)});
    if (fix) boundary_fix.push_back({ft, sources[ft].size() - 1});
  }
  // fields::require_source_components (src/fields.cpp:508-530) and fix_boundary_sources
  // (src/fix_boundary_sources.cpp:51-197) for one chunk: each point of a marked list that the
  // chunk does not own, with a nonzero amplitude, moves to the point that owns it
  // (locate_component_point: its periodic image, the amplitude times the conjugate phase)
  // unless that point is on a metal boundary, and its amplitude in the list becomes zero.
  // The moved points form new lists, one per source time and component, sorted by them.
  void require_source_components() {
    struct Moved {
      int src, c;
      ptrdiff_t index;
      cdouble A;
    };
    std::vector<Moved> moved;
    for (auto &fx : boundary_fix) {
      SrcVol &sv = sources[fx.first][fx.second];
      for (size_t j = 0; j < sv.index.size(); ++j) {
        ivec here = gv.iloc_of(sv.c, sv.index[j]);
        if (gv.owns(here) || sv.amp[j] == 0.0) continue;
        cdouble phase = 1.0;
        if (locate_point_in_user_volume(&here, &phase) && !on_metal_boundary(here))
#ifdef MUT_FIXBND_PHASE
          moved.push_back(Moved{sv.src, sv.c, gv.index(sv.c, here), sv.amp[j] * phase});
#else
          moved.push_back(Moved{sv.src, sv.c, gv.index(sv.c, here), sv.amp[j] * conj(phase)});
#endif
        sv.amp[j] = 0.0;
      }
    }
    boundary_fix.clear();
    std::stable_sort(moved.begin(), moved.end(), [](const Moved &a, const Moved &b) { return a.src != b.src ? a.src < b.src : a.c < b.c; });
    for (size_t q0 = 0; q0 < moved.size();) {
      size_t q1 = q0;
      SrcVol sv{moved[q0].c, moved[q0].src, {}, {}};
      while (q1 < moved.size() && moved[q1].src == sv.src && moved[q1].c == sv.c) {
        sv.index.push_back(moved[q1].index);
        sv.amp.push_back(moved[q1].A);
        ++q1;
      }
      sources[is_magnetic(sv.c) ? B_stuff : D_stuff].push_back(std::move(sv));
      q0 = q1;
    }
    for (int ft : {D_stuff, B_stuff})
      for (const SrcVol &sv : sources[ft]) require_component(sv.c);
  }

  // fields::add_dft (src/dft.cpp:175-235) for one component, on the centred grid or (centered
  // false) on the component's own grid without the centring average (src/dft.cpp:91-94,233)
  void add_dft(std::vector<std::unique_ptr<DftChunk<R>>> &out_list, int c, vec3 wmin, vec3 wmax, const std::vector<double> &freq,
               bool include_w, cdouble stored_weight, int decimation, bool centered = true, int vc = -1) {
    loop_in_chunks(wmin, wmax, centered ? 20 : c, [&](const ivec &is, const ivec &ie, const vec3 &s0, const vec3 &s1, const vec3 &e0, const vec3 &e1,
                                       double dV0, const ivec &shift, cdouble shift_phase) {
      if (!f[c][0]) return;
      std::unique_ptr<DftChunk<R>> ch(new DftChunk<R>);
      ch->c = c;
      ch->vc = vc;
      ch->stored_weight = stored_weight;
      ch->shift = shift;
      for (int d = 0; d < 3; ++d) ch->empty_dim[d] = gv.has(d) && wmin.t[d] == wmax.t[d];
      ch->is = is;
      ch->ie = ie;
      ch->s0 = s0;
      ch->s1 = s1;
      ch->e0 = e0;
      ch->e1 = e1;
      ch->dV0 = dV0;
      ch->dV1 = 0;
      if (centered)
        gv.yee2cent_offsets(c, ch->avg1, ch->avg2);
      else
        ch->avg1 = ch->avg2 = 0;
#ifdef MUT_DFT_SCALE
      ch->scale = stored_weight * shift_phase * double(decimation);
#else
      ch->scale = stored_weight * shift_phase * (dt / sqrt(2.0 * pi) * decimation);
#endif
      ch->include_w = include_w;
      ch->decimation = decimation;
      for (double fr : freq) ch->omega.push_back(2 * pi * fr);
      size_t N = 1;
      for (int d = 0; d < 3; ++d)
        if (gv.has(d)) N *= (ie.t[d] - is.t[d]) / 2 + 1;
      ch->N = N;
      ch->acc = dft_acc;
      if (dft_acc == MP_DFT_F64)
        ch->dft64.assign(N * freq.size(), 0.0);
      else
        ch->dft.assign(N * freq.size(), std::complex<R>(0, 0));
      out_list.push_back(std::move(ch));
    });
  }
  int auto_decimation(const std::vector<double> &freq) const {
    int dec = 0;
    double src_freq_max = 0;
    for (const SrcTime &st : srcs) {
      if (st.get_fwidth() == 0)
        dec = 1;
      else
        src_freq_max = std::max(src_freq_max, std::abs(st.freq) + 0.5 * st.get_fwidth());
    }
    double freq_max = 0;
    for (double fr : freq) freq_max = std::max(freq_max, std::abs(fr));
    if (freq_max > 0 && src_freq_max > 0)
      dec = std::max(1, int(std::floor(1 / (dt * (freq_max + src_freq_max)))));
    else
      dec = 1;
    return dec;
  }
};

// ---- DFT arrays: fields::get_dft_array (src/dft.cpp:1259-1290) ---------------------------

// dft_chunk::process_dft_component (src/dft.cpp:931-1057), its field_array branch: each point's
// accumulator at frequency num_freq over the chunk's stored weight, over the point's
// integration weight when the chunk carries weights, times the interpolation weight along the
// monitor's empty directions (retain_interp_weights, true by default, src/meep.hpp:2103-2109),
// stored at (iloc + shift - min_corner) / 2 of the non-degenerate directions ds.
template <class R>
void process_dft_chunk(const GV &gv, const DftChunk<R> &ch, int rank, const int *ds, const ivec &min_corner,
                       const size_t *array_count, int num_freq, std::complex<R> *field_array) {
  vec3 s0i = ch.s0, s1i = ch.s1, e0i = ch.e0, e1i = ch.e1;
  for (int d = 0; d < 3; ++d)
    if (!ch.empty_dim[d]) s0i.t[d] = s1i.t[d] = e0i.t[d] = e1i.t[d] = 1.0;
  const Loop L = make_loop(gv, ch.is, ch.ie);
  const size_t Nw = ch.omega.size();
  size_t chunk_idx = 0;
  for (ptrdiff_t i1 = 0; i1 < L.n[0]; ++i1)
    for (ptrdiff_t i2 = 0; i2 < L.n[1]; ++i2)
      for (ptrdiff_t i3 = 0; i3 < L.n[2]; ++i3) {
        const double w = loop_weight(gv, L, i1, i2, i3, ch.s0, ch.s1, ch.e0, ch.e1, ch.dV0 + ch.dV1 * i2);
#ifdef MUT_DFTARR_INTERP
        const double interp_w = 1.0;
#else
        const double interp_w = loop_weight(gv, L, i1, i2, i3, s0i, s1i, e0i, e1i, 1.0);
#endif
        const size_t k = Nw * (chunk_idx++) + num_freq;
        cdouble v = ch.acc == MP_DFT_F64 ? ch.dft64[k] : cdouble(ch.dft[k]);
#ifndef MUT_DFTARR_SIGN
        v /= ch.stored_weight;
#endif
#ifndef MUT_DFTARR_WEIGHT
#ifdef MUT_DFTARR_SQRT
        if (ch.include_w && v != 0.0) v /= w;
#else
        if (ch.include_w && v != 0.0) v /= ch.sqrt_w ? std::sqrt(w) : w; // src/dft.cpp:1013-1014
#endif
#endif
        ivec iloc = loop_iloc(L, i1, i2, i3);
        for (int d = 0; d < 3; ++d) {
#ifndef MUT_DFTARR_SHIFT
          iloc.t[d] += ch.shift.t[d];
#endif
          iloc.t[d] -= min_corner.t[d];
        }
        size_t idx2 = 0, stride = 1;
        for (int i = rank - 1; i >= 0; stride *= array_count[i--]) {
          if (iloc.t[ds[i]] < 0 || size_t(iloc.t[ds[i]] / 2) >= array_count[i]) fail("internal: DFT array point out of range");
          idx2 += stride * size_t(iloc.t[ds[i]] / 2);
        }
        field_array[idx2] = std::complex<R>(interp_w * v);
      }
}

// fields::process_dft_component (src/dft.cpp:1130-1256) with a field array, after
// fields::get_dft_component_dims (:1061-1110): the corners start from the cell's rounded
// corners, swapped and pushed one step out, and take the min and max over the chunks of
// component c; each direction with more than one point is an array dimension.  Then
// collapse_array (src/array_slice.cpp:554-596, reduce_array_dimensions :524-551) sums the
// array over the dimensions in which the monitor volume is empty.
template <class R>
std::vector<std::complex<R>> get_dft_array(const GV &gv, const std::vector<std::vector<void *>> &lists, int c, int num_freq,
                                           size_t nfreq, const vec3 &wmin, const vec3 &wmax, int *rank_out,
                                           size_t *dims_out) {
  if (num_freq < 0 || num_freq > int(nfreq) - 1)
    fail("process_dft_component: frequency index %d is outside the range of the frequency array of size %zu", num_freq,
         nfreq);
  ivec min_corner, max_corner;
  for (int d = 0; d < 3; ++d) {
    min_corner.t[d] = 2 * gv.num[d] + 1;
    max_corner.t[d] = -1;
  }
  std::vector<const DftChunk<R> *> mine;
  for (const auto &lst : lists)
    for (void *p : lst) {
      const DftChunk<R> *ch = reinterpret_cast<const DftChunk<R> *>(p);
      if (ch->c != c) continue;
      mine.push_back(ch);
      for (int d = 0; d < 3; ++d)
        if (gv.has(d)) {
          const int a = ch->is.t[d] + ch->shift.t[d], b = ch->ie.t[d] + ch->shift.t[d];
          min_corner.t[d] = std::min(min_corner.t[d], std::min(a, b));
          max_corner.t[d] = std::max(max_corner.t[d], std::max(a, b));
        }
    }
  int rank = 0, ds[3] = {0, 0, 0};
  size_t dims[3] = {1, 1, 1}, array_size = 1;
  for (int d = 0; d < 3; ++d) {
    if (!gv.has(d)) continue;
    const int n = std::max(0, (max_corner.t[d] - min_corner.t[d]) / 2 + 1);
    if (n > 1) {
      ds[rank] = d;
      dims[rank++] = size_t(n);
      array_size *= size_t(n);
    }
  }
  *rank_out = 0;
  dims_out[0] = dims_out[1] = dims_out[2] = 1;
  if (rank == 0) return {};
  std::vector<std::complex<R>> arr(array_size, std::complex<R>(0, 0));
  for (const DftChunk<R> *ch : mine) process_dft_chunk(gv, *ch, rank, ds, min_corner, dims, num_freq, arr.data());
  // collapse_array: sum over the dimensions in which the monitor is empty
  int rrank = 0;
  size_t rdims[3] = {1, 1, 1};
  bool keep[3] = {true, true, true};
  for (int r = 0; r < rank; ++r) {
    keep[r] = !(wmax.t[ds[r]] - wmin.t[ds[r]] == 0.0);
    if (keep[r]) rdims[rrank++] = dims[r];
  }
  *rank_out = rrank;
  for (int r = 0; r < rrank; ++r) dims_out[r] = rdims[r];
  if (rrank == rank) {
    for (int r = 0; r < rank; ++r) dims_out[r] = dims[r];
    return arr;
  }
  const size_t rsize = rdims[0] * rdims[1] * rdims[2];
  std::vector<R> red(2 * rsize, R(0));
  size_t n[3] = {0, 0, 0};
  for (size_t full = 0; full < array_size; ++full) {
    size_t rem = full;
    for (int r = rank - 1; r >= 0; --r) {
      n[r] = rem % dims[r];
      rem /= dims[r];
    }
    size_t rindex = 0;
    for (int r = 0; r < rank; ++r)
      if (keep[r]) rindex = rindex * dims[r] + n[r];
    red[2 * rindex] += arr[full].real();
    red[2 * rindex + 1] += arr[full].imag();
  }
  std::vector<std::complex<R>> out(rsize);
  for (size_t q = 0; q < rsize; ++q) out[q] = std::complex<R>(red[2 * q], red[2 * q + 1]);
  return out;
}

// ---- C-ABI glue ----------------------------------------------------------------------
struct AnyStructure {
  bool single;
  std::unique_ptr<Structure<double>> d;
  std::unique_ptr<Structure<float>> s;
};
struct AnyFields {
  bool single;
  std::unique_ptr<Fields<double>> d;
  std::unique_ptr<Fields<float>> s;
};
struct AnyFlux {
  AnyFields *owner;
  const std::vector<void *> *diag = nullptr;  // dft_force: its diag list; dft_energy: D
  const std::vector<void *> *extra = nullptr; // dft_energy: B
  std::vector<double> freq;
  int decimation;
  // chunk lists in the reference's order: E list then H list (src/dft.cpp:623-641)
  std::vector<void *> E, H;
  vec3 wmin, wmax; // the monitor volume (dft_flux::where, dft_fields::where)
};

template <class F> int guard(F fn) {
  try {
    fn();
    return 0;
  } catch (const std::exception &e) {
    last_error = e.what();
    return -1;
  }
}

template <class R> void fill_structure(Structure<R> &s, int dim, double xs, double ys, double zs, double a, mp_scalar_func eps, void *eu, int n_pml,
                                       const int *pd, const int *ps, const double *pt, const double *pR, const PmlTerm *terms,
                                       double courant) {
  auto np = [&](double sz) { return sz == 0 ? 1 : int(sz * a + 0.5); };
  if (dim == 1)
    s.gv.init(1, a, 0, 0, int(zs * a + 0.5));
  else if (dim == 2)
    s.gv.init(2, a, np(xs), np(ys), 0);
  else if (dim == 3)
    s.gv.init(3, a, np(xs), np(ys), np(zs));
  else
    fail("dim must be 1, 2 or 3");
  s.courant = courant;
  s.a = a;
  s.dt = courant / a;
  const GV &gv = s.gv;
  // boundary_region::check_ok / apply (src/structure.cpp:206-250)
  for (int i = 0; i < n_pml; ++i) {
    int d = pd[i];
    if (d < 0 || d > 2 || !gv.has(d) || gv.num[d] <= 1) continue;
    use_pml(s, d, pt[i], gv.boundary_location(ps[i], d), pR && pR[i] > 0 ? pR[i] : 1e-15, terms[i]);
  }
  // set_chi1inv without averaging: 1/eps at the Yee point (src/anisotropic_averaging.cpp:90-98)
  if (eps)
    for (int c = 0; c < 5; ++c) {
      if (component_direction(c) == NODIR) continue;
      bool hasf = gv.dim == 1 ? c == 0 : true;
      if (!hasf) continue;
      std::vector<R> arr(gv.ntot, R(1));
      bool trivial = true;
      ivec s0 = gv.iyee_shift(c), e0 = gv.big_corner();
      for (int d = 0; d < 3; ++d)
        if (gv.has(d)) e0.t[d] += s0.t[d];
      loop_ivecs(gv, s0, e0, [&](ptrdiff_t i, const Loop &L, ptrdiff_t i1, ptrdiff_t i2, ptrdiff_t i3) {
        ivec here = loop_iloc(L, i1, i2, i3);
        vec3 h = gv.loc(here), mn, mx, cen;
        double hinva = 0.5 * gv.inva;
        for (int d = 0; d < 3; ++d)
          if (gv.has(d)) {
            mn.t[d] = h.t[d] - hinva;
            mx.t[d] = h.t[d] + hinva;
            cen.t[d] = (mn.t[d] + mx.t[d]) * 0.5;
          }
        double r[3] = {cen.t[0], cen.t[1], cen.t[2]};
        arr[i] = R(1 / eps(r, eu));
        if (arr[i] != R(1)) trivial = false;
      });
      if (!trivial) s.chi1inv[c] = std::move(arr);
    }
}

template <class R> AnyFields *make_fields(Structure<R> *st) {
  AnyFields *af = new AnyFields;
  af->single = sizeof(R) == sizeof(float);
  if (af->single)
    af->s.reset(new Fields<float>(reinterpret_cast<Structure<float> *>(st)));
  else
    af->d.reset(new Fields<double>(reinterpret_cast<Structure<double> *>(st)));
  return af;
}

template <class FN> auto with_fields(AnyFields *f, FN fn) {
  if (f->single) return fn(*f->s);
  return fn(*f->d);
}
template <class FN> auto with_structure(AnyStructure *s, FN fn) {
  if (s->single) return fn(*s->s);
  return fn(*s->d);
}

vec3 to_vec(const double *p) {
  vec3 v;
  for (int d = 0; d < 3; ++d) v.t[d] = p ? p[d] : 0;
  return v;
}

} // namespace

extern "C" {

const char *mp_last_error(void) { return last_error.c_str(); }
const char *mp_engine_name(void) { return "cpu-oracle"; }

mp_structure *mp_structure_new_pml(int dim, double xsize, double ysize, double zsize, double a, mp_scalar_func eps,
                                   void *eps_user, int n_pml, const int *pml_direction, const int *pml_side,
                                   const double *pml_thickness, const double *pml_R, const double *pml_mean_stretch,
                                   const mp_scalar_func *pml_profile, void *const *pml_profile_user,
                                   const double *pml_integral, const double *pml_integral_u, double courant,
                                   int single_precision) {
  AnyStructure *s = new AnyStructure;
  s->single = single_precision != 0;
  int rc = guard([&] {
    std::vector<PmlTerm> terms(n_pml > 0 ? n_pml : 0);
    for (int i = 0; i < n_pml; ++i) {
      if (pml_mean_stretch) terms[i].mean_stretch = pml_mean_stretch[i];
      if (pml_profile) terms[i].profile = pml_profile[i];
      if (pml_profile_user) terms[i].user = pml_profile_user[i];
      if (pml_integral && pml_integral[i] > 0) terms[i].integral = pml_integral[i];
      if (pml_integral_u && pml_integral_u[i] > 0) terms[i].integral_u = pml_integral_u[i];
    }
    if (s->single) {
      s->s.reset(new Structure<float>);
      fill_structure(*s->s, dim, xsize, ysize, zsize, a, eps, eps_user, n_pml, pml_direction, pml_side, pml_thickness, pml_R,
                     terms.data(), courant);
    }
    else {
      s->d.reset(new Structure<double>);
      fill_structure(*s->d, dim, xsize, ysize, zsize, a, eps, eps_user, n_pml, pml_direction, pml_side, pml_thickness, pml_R,
                     terms.data(), courant);
    }
  });
  if (rc) {
    delete s;
    return nullptr;
  }
  return reinterpret_cast<mp_structure *>(s);
}
mp_structure *mp_structure_new(int dim, double xsize, double ysize, double zsize, double a, mp_scalar_func eps, void *eps_user, int n_pml,
                               const int *pml_direction, const int *pml_side, const double *pml_thickness, const double *pml_R,
                               double courant, int single_precision) {
  return mp_structure_new_pml(dim, xsize, ysize, zsize, a, eps, eps_user, n_pml, pml_direction, pml_side, pml_thickness, pml_R,
                              nullptr, nullptr, nullptr, nullptr, nullptr, courant, single_precision);
}
void mp_structure_delete(mp_structure *s) { delete reinterpret_cast<AnyStructure *>(s); }
size_t mp_structure_ntot(const mp_structure *s) {
  auto *as = const_cast<AnyStructure *>(reinterpret_cast<const AnyStructure *>(s));
  return with_structure(as, [](auto &st) { return st.gv.ntot; });
}
int mp_structure_num(const mp_structure *s, int *n) {
  auto *as = const_cast<AnyStructure *>(reinterpret_cast<const AnyStructure *>(s));
  return with_structure(as, [&](auto &st) {
    for (int d = 0; d < 3; ++d) n[d] = st.gv.num[d];
    return 0;
  });
}
double mp_structure_dt(const mp_structure *s) {
  auto *as = const_cast<AnyStructure *>(reinterpret_cast<const AnyStructure *>(s));
  return with_structure(as, [](auto &st) { return st.dt; });
}
int mp_structure_set_chi1inv(mp_structure *s, int c, const double *arr) {
  return guard([&] {
    if (c < 0 || c >= 5 || component_direction(c) == NODIR) fail("chi1inv: electric component expected");
    with_structure(reinterpret_cast<AnyStructure *>(s), [&](auto &st) {
      typedef typename std::remove_reference<decltype(st.chi1inv[0][0])>::type R;
      st.chi1inv[c].assign(st.gv.ntot, R(1));
      for (size_t i = 0; i < st.gv.ntot; ++i) st.chi1inv[c][i] = R(arr[i]);
      return 0;
    });
  });
}
int mp_structure_get_chi1inv(const mp_structure *s, int c, double *arr) {
  return guard([&] {
    with_structure(const_cast<AnyStructure *>(reinterpret_cast<const AnyStructure *>(s)), [&](auto &st) {
      for (size_t i = 0; i < st.gv.ntot; ++i) arr[i] = st.chi1inv[c].empty() ? 1.0 : double(st.chi1inv[c][i]);
      return 0;
    });
  });
}
int mp_structure_add_susceptibility(mp_structure *s, double omega_0, double gamma, int drude, mp_scalar_func sigma, void *su) {
  return mp_structure_add_susceptibility_ft(s, E_stuff, omega_0, gamma, drude, sigma, su);
}
int mp_structure_add_susceptibility_ft(mp_structure *s, int ft, double omega_0, double gamma, int drude, mp_scalar_func sigma,
                                       void *su) {
  int idx = -1;
  int rc = guard([&] {
    if (ft != E_stuff && ft != H_stuff) fail("add_susceptibility: field type must be E_stuff or H_stuff");
    with_structure(reinterpret_cast<AnyStructure *>(s), [&](auto &st) {
      typedef typename std::remove_reference<decltype(st.chi1inv[0][0])>::type R;
      Sus<R> sus;
      sus.omega_0 = R(omega_0);
      sus.gamma = R(gamma);
      sus.drude = drude != 0;
      sus.ft = ft;
      const GV &gv = st.gv;
      for (int c = 0; c < NC; ++c) sus.trivial[c] = true;
      // structure_chunk::add_susceptibility (src/anisotropic_averaging.cpp:330-427): sigma on
      // the components of field type ft
      const int c0 = ft == E_stuff ? 0 : 5;
      for (int c = c0; c < c0 + 5; ++c) {
        if (component_direction(c) == NODIR) continue;
        if (gv.dim == 1 && c != (ft == E_stuff ? 0 : 6)) continue;
        std::vector<R> arr(gv.ntot, R(0));
        bool trivial = true;
        ivec s0 = gv.iyee_shift(c), e0 = gv.big_corner();
        for (int d = 0; d < 3; ++d)
          if (gv.has(d)) e0.t[d] += s0.t[d];
        loop_ivecs(gv, s0, e0, [&](ptrdiff_t i, const Loop &L, ptrdiff_t i1, ptrdiff_t i2, ptrdiff_t i3) {
          vec3 here = gv.loc(loop_iloc(L, i1, i2, i3));
          double r[3] = {here.t[0], here.t[1], here.t[2]};
          arr[i] = R(sigma ? sigma(r, su) : 0.0);
          if (arr[i] != R(0)) trivial = false;
        });
        sus.trivial[c] = trivial;
        if (!trivial) sus.sigma[c] = std::move(arr);
      }
      st.sus.push_back(std::move(sus));
      idx = int(st.sus.size()) - 1;
      return 0;
    });
  });
  return rc ? rc : idx;
}
int mp_structure_set_sigma(mp_structure *s, int isus, int c, const double *sigma) {
  return guard([&] {
    with_structure(reinterpret_cast<AnyStructure *>(s), [&](auto &st) {
      typedef typename std::remove_reference<decltype(st.chi1inv[0][0])>::type R;
      if (isus < 0 || isus >= int(st.sus.size())) fail("no such susceptibility");
      auto &su = st.sus[isus];
      su.sigma[c].assign(st.gv.ntot, R(0));
      bool trivial = true;
      for (size_t i = 0; i < st.gv.ntot; ++i) {
        su.sigma[c][i] = R(sigma[i]);
        if (sigma[i] != 0) trivial = false;
      }
      su.trivial[c] = trivial;
      if (trivial) su.sigma[c].clear();
      return 0;
    });
  });
}

// structure_chunk::set_conductivity (src/structure.cpp:868-905): conductivity[c_C][d] of the D
// (or B) component c_C along c's direction, sampled at c_C's points over the whole array and,
// for an E (or H) component, multiplied by that component's chi1inv at the time of the call
int mp_structure_set_conductivity(mp_structure *s, int c, mp_scalar_func C, void *user) {
  return guard([&] {
    if (c < 0 || c >= NC || component_direction(c) == NODIR) fail("invalid component for conductivity");
    with_structure(reinterpret_cast<AnyStructure *>(s), [&](auto &st) {
      typedef typename std::remove_reference<decltype(st.chi1inv[0][0])>::type R;
      const GV &gv = st.gv;
      const int d = component_direction(c);
      const int cC = is_electric(c) ? direction_component(10, d) : is_magnetic(c) ? direction_component(15, d) : c;
      if (gv.dim == 1 && cC != 10 && cC != 16) return 0; // !gv.has_field(c)
      const R *multby = (is_electric(c) || is_magnetic(c)) && !st.chi1inv[c].empty() ? st.chi1inv[c].data() : nullptr;
#ifdef MUT_COND_CHI
      multby = nullptr;
#endif
      std::vector<R> arr(gv.ntot, R(0));
      bool trivial = true;
      ivec s0 = gv.iyee_shift(cC), e0 = gv.big_corner();
      for (int dd = 0; dd < 3; ++dd)
        if (gv.has(dd)) e0.t[dd] += s0.t[dd];
      loop_ivecs(gv, s0, e0, [&](ptrdiff_t i, const Loop &L, ptrdiff_t i1, ptrdiff_t i2, ptrdiff_t i3) {
        vec3 here = gv.loc(loop_iloc(L, i1, i2, i3));
        double r[3] = {here.t[0], here.t[1], here.t[2]};
        arr[i] = multby ? R(C(r, user) * multby[i]) : R(C(r, user));
        trivial = trivial && (arr[i] == 0.0);
      });
      if (trivial)
        st.cond[cC].clear();
      else
        st.cond[cC] = std::move(arr);
      st.condinv_stale = true;
      return 0;
    });
  });
}

mp_fields *mp_fields_new(mp_structure *s) {
  AnyStructure *as = reinterpret_cast<AnyStructure *>(s);
  AnyFields *af = nullptr;
  int rc = guard([&] {
    if (as->single)
      af = make_fields(as->s.get());
    else
      af = make_fields(as->d.get());
  });
  return rc ? nullptr : reinterpret_cast<mp_fields *>(af);
}
void mp_fields_delete(mp_fields *f) { delete reinterpret_cast<AnyFields *>(f); }
int mp_fields_use_real_fields(mp_fields *f) {
  return guard([&] {
    with_fields(reinterpret_cast<AnyFields *>(f), [&](auto &F) {
      for (int d = 0; d < 3; ++d)
        if (F.boundaries[High][d] == Periodic && F.k[d] != 0.0) fail("Can't use real fields with bloch boundary conditions!");
      F.is_real = true;
      for (int c = 0; c < NC; ++c) F.f[c][1] = F.f_u[c][1] = F.f_w[c][1] = F.fmp[c][1] = nullptr;
      F.connections_valid = false;
      return 0;
    });
  });
}
int mp_fields_use_bloch(mp_fields *f, const double *k) {
  return guard([&] {
    with_fields(reinterpret_cast<AnyFields *>(f), [&](auto &F) {
      int j = 0;
      for (int d = 0; d < 3; ++d)
        if (F.gv.has(d)) F.use_bloch(d, k[j++]);
      return 0;
    });
  });
}
int mp_fields_use_bloch_dir(mp_fields *f, int d, double kr, double ki) {
  return guard([&] {
    with_fields(reinterpret_cast<AnyFields *>(f), [&](auto &F) {
      F.use_bloch(d, cdouble(kr, ki));
      return 0;
    });
  });
}
int mp_fields_set_boundary(mp_fields *f, int side, int d, int cond) {
  return guard([&] {
    with_fields(reinterpret_cast<AnyFields *>(f), [&](auto &F) {
      F.set_boundary(side, d, cond);
      return 0;
    });
  });
}
int mp_fields_add_point_source_legacy(mp_fields *f, int c, double freq, double width, double peaktime, double cutoff, const double *p,
                                      double ar, double ai, int is_c) {
  return guard([&] {
    with_fields(reinterpret_cast<AnyFields *>(f), [&](auto &F) {
      // src/sources.cpp:193-213
      width /= freq;
      SrcTime st;
      if (is_c) {
        st = SrcTime::continuous(freq, width, F.t * F.dt, INFINITY, cutoff);
      }
      else {
        cutoff = F.gv.inva + cutoff * width;
        if (peaktime <= 0.0) peaktime = F.t * F.dt + cutoff;
        peaktime += is_magnetic(c) ? -F.dt * 0.5 : F.dt;
        st = SrcTime::gaussian(freq, width, peaktime - cutoff, peaktime + cutoff);
      }
      st.is_integrated = !is_magnetic(c);
      vec3 pv = to_vec(p);
      F.add_volume_source(c, st, pv, pv, cdouble(ar, ai), nullptr, nullptr);
      return 0;
    });
  });
}
int mp_fields_add_volume_source(mp_fields *f, int c, const mp_src_time *src, const double *vmin, const double *vmax, double ar, double ai,
                                mp_complex_func A, void *Au) {
  return guard([&] {
    with_fields(reinterpret_cast<AnyFields *>(f), [&](auto &F) {
      SrcTime st = src->kind == 0 ? SrcTime::gaussian(src->freq, src->width, src->start_time, src->end_time)
                                  : SrcTime::continuous(src->freq, src->width, src->start_time, src->end_time, src->slowness);
      st.is_integrated = src->is_integrated != 0;
      F.add_volume_source(c, st, to_vec(vmin), to_vec(vmax), cdouble(ar, ai), A, Au);
      return 0;
    });
  });
}
int mp_fields_step(mp_fields *f) {
  return guard([&] {
    with_fields(reinterpret_cast<AnyFields *>(f), [&](auto &F) {
      F.step();
      return 0;
    });
  });
}
int mp_fields_step_n(mp_fields *f, long n) {
  return guard([&] {
    with_fields(reinterpret_cast<AnyFields *>(f), [&](auto &F) {
      for (long i = 0; i < n; ++i) F.step();
      return 0;
    });
  });
}
// the oracle has no device memory: the allocation-free entry point is step_n
int mp_fields_step_n_noalloc(mp_fields *f, long n) { return mp_fields_step_n(f, n); }
double mp_fields_time(const mp_fields *f) {
  return with_fields(const_cast<AnyFields *>(reinterpret_cast<const AnyFields *>(f)), [](auto &F) { return F.t * F.dt; });
}
double mp_fields_round_time(const mp_fields *f) {
  return with_fields(const_cast<AnyFields *>(reinterpret_cast<const AnyFields *>(f)), [](auto &F) { return double(float(F.t * F.dt)); });
}
long mp_fields_t(const mp_fields *f) {
  return with_fields(const_cast<AnyFields *>(reinterpret_cast<const AnyFields *>(f)), [](auto &F) { return long(F.t); });
}
double mp_fields_dt(const mp_fields *f) {
  return with_fields(const_cast<AnyFields *>(reinterpret_cast<const AnyFields *>(f)), [](auto &F) { return F.dt; });
}
double mp_fields_last_source_time(mp_fields *f) {
  return with_fields(reinterpret_cast<AnyFields *>(f), [](auto &F) {
    double after = 0;
    for (auto &st : F.srcs) after = std::max(after, st.last_time());
    return after;
  });
}
int mp_fields_get_field(mp_fields *f, int c, const double *p, double *re, double *im) {
  return guard([&] {
    with_fields(reinterpret_cast<AnyFields *>(f), [&](auto &F) {
      cdouble v = F.get_field(c, to_vec(p));
      *re = v.real();
      *im = v.imag();
      return 0;
    });
  });
}
int mp_fields_have_component(mp_fields *f, int c) {
  return with_fields(reinterpret_cast<AnyFields *>(f), [&](auto &F) { return F.f[c][0] ? 1 : 0; });
}
int mp_fields_get_array(mp_fields *f, int c, int cmp, int kind, int ipol, double *out) {
  int missing = 0;
  int rc = guard([&] {
    with_fields(reinterpret_cast<AnyFields *>(f), [&](auto &F) {
      typedef typename std::remove_reference<decltype(*F.f[0][0])>::type R;
      const R *src = nullptr;
      if (kind == MP_ARR_F) src = F.f[c][cmp];
      if (kind == MP_ARR_F_U) src = F.f_u[c][cmp];
      if (kind == MP_ARR_F_W) src = F.f_w[c][cmp];
      if (kind == MP_ARR_F_COND) src = F.f_cond[c][cmp];
      if (kind == MP_ARR_P || kind == MP_ARR_P_PREV)
        for (auto &p : F.pols)
          if (p.isus == ipol && p.allocated) src = kind == MP_ARR_P ? p.P[c][cmp] : p.Pp[c][cmp];
      for (size_t i = 0; i < F.gv.ntot; ++i) out[i] = src ? double(src[i]) : 0.0;
      missing = src ? 0 : 1;
      return 0;
    });
  });
  return rc ? rc : missing;
}
int mp_fields_set_array(mp_fields *f, int c, int cmp, int kind, int ipol, const double *in) {
  return guard([&] {
    with_fields(reinterpret_cast<AnyFields *>(f), [&](auto &F) {
      typedef typename std::remove_reference<decltype(*F.f[0][0])>::type R;
      R *dst = nullptr;
      if (kind == MP_ARR_F) dst = F.f[c][cmp];
      if (kind == MP_ARR_F_U) dst = F.f_u[c][cmp];
      if (kind == MP_ARR_F_W) dst = F.f_w[c][cmp];
      if (kind == MP_ARR_F_COND) dst = F.f_cond[c][cmp];
      if (kind == MP_ARR_P || kind == MP_ARR_P_PREV)
        for (auto &p : F.pols)
          if (p.isus == ipol && p.allocated) dst = kind == MP_ARR_P ? p.P[c][cmp] : p.Pp[c][cmp];
      if (!dst) fail("array not allocated");
      for (size_t i = 0; i < F.gv.ntot; ++i) dst[i] = R(in[i]);
      return 0;
    });
  });
}
int mp_fields_get_times(mp_fields *f, double *t7) {
  return with_fields(reinterpret_cast<AnyFields *>(f), [&](auto &F) {
    for (int i = 0; i < 7; ++i) t7[i] = F.times[i];
    return 0;
  });
}
// the oracle has no kernels to time
int mp_fields_set_kernel_timing(mp_fields *, int) { return 0; }
int mp_fields_get_kernel_timing(mp_fields *, int, double *ms, long *n) {
  *ms = 0;
  *n = 0;
  return 0;
}
int mp_fields_reset_kernel_timing(mp_fields *) { return 0; }
int mp_fields_set_dft_accumulator(mp_fields *f, int kind) {
  return guard([&] {
    if (kind != MP_DFT_REALNUM && kind != MP_DFT_F64) fail("oracle supports realnum and FP64 DFT accumulators");
    with_fields(reinterpret_cast<AnyFields *>(f), [&](auto &F) {
      F.dft_acc = kind;
      return 0;
    });
  });
}

mp_dft_flux *mp_fields_add_dft_flux(mp_fields *f, int d, const double *vmin, const double *vmax, const double *freq, int nfreq,
                                    int use_symmetry, int centered, int dec) {
  (void)use_symmetry; // the identity symmetry: nothing to reduce (src/dft.cpp:614)
  AnyFields *af = reinterpret_cast<AnyFields *>(f);
  AnyFlux *fl = new AnyFlux;
  fl->owner = af;
  fl->freq.assign(freq, freq + nfreq);
  int rc = guard([&] {
    with_fields(af, [&](auto &F) {
      typedef typename std::remove_reference<decltype(*F.f[0][0])>::type R;
      if (!F.components_allocated) fail("allocate field components (by adding sources) before adding dft objects");
      int decim = dec > 0 ? dec : F.auto_decimation(fl->freq);
      fl->decimation = decim;
      int cE[2], cH[2];
      if (d == X) {
        cE[0] = 1; cE[1] = 4; cH[0] = 9; cH[1] = 6;
      }
      else if (d == Y) {
        cE[0] = 4; cE[1] = 0; cH[0] = 5; cH[1] = 9;
      }
      else {
        cE[0] = 0; cE[1] = 1; cH[0] = 6; cH[1] = 5;
      }
      vec3 mn = to_vec(vmin), mx = to_vec(vmax);
      fl->wmin = mn;
      fl->wmax = mx;
      // the reference prepends each new chunk list, so the E list is [cE1 chunks, cE0 chunks]
      std::vector<std::unique_ptr<DftChunk<R>>> Elist[2], Hlist[2];
      for (int i = 0; i < 2; ++i) {
        F.add_dft(Elist[i], cE[i], mn, mx, fl->freq, true, cdouble(double(1 - 2 * i), 0), decim, centered != 0);
        F.add_dft(Hlist[i], cH[i], mn, mx, fl->freq, false, cdouble(1.0, 0), decim, centered != 0);
      }
      for (int i = 1; i >= 0; --i) {
        for (int j = int(Elist[i].size()) - 1; j >= 0; --j) {
          fl->E.push_back(Elist[i][j].get());
          F.dft_chunks.push_back(std::move(Elist[i][j]));
        }
        for (int j = int(Hlist[i].size()) - 1; j >= 0; --j) {
          fl->H.push_back(Hlist[i][j].get());
          F.dft_chunks.push_back(std::move(Hlist[i][j]));
        }
      }
      return 0;
    });
  });
  if (rc) {
    delete fl;
    return nullptr;
  }
  return reinterpret_cast<mp_dft_flux *>(fl);
}
void mp_dft_flux_delete(mp_dft_flux *fl) { delete reinterpret_cast<AnyFlux *>(fl); }
int mp_dft_flux_flux(mp_dft_flux *flp, double *out) {
  AnyFlux *fl = reinterpret_cast<AnyFlux *>(flp);
  return guard([&] {
    with_fields(fl->owner, [&](auto &F) {
      typedef typename std::remove_reference<decltype(*F.f[0][0])>::type R;
      size_t Nf = fl->freq.size();
      for (size_t i = 0; i < Nf; ++i) out[i] = 0;
      for (size_t q = 0; q < fl->E.size() && q < fl->H.size(); ++q) {
        auto *E = reinterpret_cast<DftChunk<R> *>(fl->E[q]);
        auto *H = reinterpret_cast<DftChunk<R> *>(fl->H[q]);
        if (E->N != H->N) fail("flux of E and H chunks with different points (a monitor off the centred grid)");
        for (size_t k = 0; k < E->N; ++k)
          for (size_t i = 0; i < Nf; ++i) {
            cdouble e = E->acc == MP_DFT_F64 ? E->dft64[k * Nf + i] : cdouble(E->dft[k * Nf + i]);
            cdouble h = H->acc == MP_DFT_F64 ? H->dft64[k * Nf + i] : cdouble(H->dft[k * Nf + i]);
            if (E->acc == MP_DFT_F64)
              out[i] += real(e * conj(h));
            else
              out[i] += double(real(E->dft[k * Nf + i] * conj(H->dft[k * Nf + i])));
          }
      }
      return 0;
    });
  });
}
int mp_dft_flux_num_chunks(mp_dft_flux *flp, int which) {
  AnyFlux *fl = reinterpret_cast<AnyFlux *>(flp);
  return int(which ? fl->H.size() : fl->E.size());
}
long mp_dft_flux_chunk_points(mp_dft_flux *flp, int which, int i) {
  AnyFlux *fl = reinterpret_cast<AnyFlux *>(flp);
  return with_fields(fl->owner, [&](auto &F) {
    typedef typename std::remove_reference<decltype(*F.f[0][0])>::type R;
    auto *ch = reinterpret_cast<DftChunk<R> *>(which ? fl->H[i] : fl->E[i]);
    return long(ch->N);
  });
}
int mp_dft_flux_decimation(mp_dft_flux *flp) { return reinterpret_cast<AnyFlux *>(flp)->decimation; }
int mp_dft_flux_get_dft(mp_dft_flux *flp, int which, int i, double *out) {
  AnyFlux *fl = reinterpret_cast<AnyFlux *>(flp);
  return guard([&] {
    with_fields(fl->owner, [&](auto &F) {
      typedef typename std::remove_reference<decltype(*F.f[0][0])>::type R;
      auto &lst = which ? fl->H : fl->E;
      if (i < 0 || i >= int(lst.size())) fail("no such dft chunk");
      auto *ch = reinterpret_cast<DftChunk<R> *>(lst[i]);
      size_t n = ch->N * fl->freq.size();
      for (size_t j = 0; j < n; ++j) {
        cdouble v = ch->acc == MP_DFT_F64 ? ch->dft64[j] : cdouble(ch->dft[j]);
        out[2 * j] = v.real();
        out[2 * j + 1] = v.imag();
      }
      return 0;
    });
  });
}

int mp_dft_flux_set_dft(mp_dft_flux *flp, int which, int i, const double *in) {
  AnyFlux *fl = reinterpret_cast<AnyFlux *>(flp);
  return guard([&] {
    with_fields(fl->owner, [&](auto &F) {
      typedef typename std::remove_reference<decltype(*F.f[0][0])>::type R;
      auto &lst = which ? fl->H : fl->E;
      if (i < 0 || i >= int(lst.size())) fail("no such dft chunk");
      auto *ch = reinterpret_cast<DftChunk<R> *>(lst[i]);
      size_t n = ch->N * fl->freq.size();
      for (size_t j = 0; j < n; ++j) {
        if (ch->acc == MP_DFT_F64)
          ch->dft64[j] = cdouble(in[2 * j], in[2 * j + 1]);
        else
          ch->dft[j] = std::complex<R>(R(in[2 * j]), R(in[2 * j + 1]));
      }
      return 0;
    });
  });
}
// dft_flux::scale_dfts (src/dft.cpp:597-600) over dft_chunk::scale_dft (:402-406): every
// accumulator of both lists times scale, in the accumulator's type
int mp_dft_flux_scale_dfts(mp_dft_flux *flp, double re, double im) {
  AnyFlux *fl = reinterpret_cast<AnyFlux *>(flp);
  return guard([&] {
    with_fields(fl->owner, [&](auto &F) {
      typedef typename std::remove_reference<decltype(*F.f[0][0])>::type R;
#ifdef MUT_SCALE_DFTS
      for (auto *lst : {&fl->E})
#else
      for (auto *lst : std::vector<const std::vector<void *> *>{&fl->E, &fl->H, fl->diag, fl->extra})
#endif
        for (void *p : lst ? *lst : std::vector<void *>()) {
          auto *ch = reinterpret_cast<DftChunk<R> *>(p);
          for (auto &x : ch->dft) x *= std::complex<R>(R(re), R(im));
          for (auto &x : ch->dft64) x *= cdouble(re, im);
        }
      return 0;
    });
  });
}

static long dft_array(AnyFlux *fl, bool flux, int c, int num_freq, int *rank, size_t *dims, double *out) {
  long count = 0;
  const int rc = guard([&] {
    with_fields(fl->owner, [&](auto &F) {
      typedef typename std::remove_reference<decltype(*F.f[0][0])>::type R;
      if (c < 0 || c >= NC) fail("get_dft_array: component %d is not a field component", c);
      std::vector<std::vector<void *>> lists{fl->E};
      if (flux) lists.push_back(fl->H);
      if (fl->diag) lists.push_back(*fl->diag); // dft_force (src/dft.cpp:1270-1280)
      std::vector<std::complex<R>> a =
          get_dft_array<R>(F.gv, lists, c, num_freq, fl->freq.size(), fl->wmin, fl->wmax, rank, dims);
      if (out)
        for (size_t q = 0; q < a.size(); ++q) {
          out[2 * q] = double(a[q].real());
          out[2 * q + 1] = double(a[q].imag());
        }
      count = long(a.size());
      return 0;
    });
  });
  return rc ? -1 : count;
}
long mp_dft_flux_get_dft_array(mp_dft_flux *flp, int c, int num_freq, int *rank, size_t *dims, double *out) {
  return dft_array(reinterpret_cast<AnyFlux *>(flp), true, c, num_freq, rank, dims, out);
}

// ---- time-domain flux (src/energy_and_flux.cpp:180-241, src/meep.hpp:2322-2352) ----------
struct AnyFluxVol {
  AnyFields *owner;
  void *fv;
};
static int flux_direction(const GV &gv, int d, const vec3 &mn, const vec3 &mx) {
  if (d >= 0) { // coordinate_mismatch rejects only the cylindrical directions here
    if (d > Z) fail("invalid direction in add_flux_vol");
    return d;
  }
  // volume::normal_direction (src/vec.cpp:227-257): Z in 1D, else the one direction of zero
  // extent when the others have positive extent
  if (gv.dim == 1) return Z;
  int nd = -1, nzero = 0, npos = 0;
  for (int k = 0; k < 3; ++k)
    if (gv.has(k)) {
      if (mx.t[k] - mn.t[k] == 0) nd = k, ++nzero;
      else if (mx.t[k] - mn.t[k] > 0) ++npos;
    }
  if (nzero != 1 || npos != gv.dim - 1) fail("invalid direction in add_flux_vol");
  return nd;
}
mp_flux_vol *mp_fields_add_flux_vol(mp_fields *f, int d, const double *vmin, const double *vmax) {
  AnyFields *af = reinterpret_cast<AnyFields *>(f);
  AnyFluxVol *h = new AnyFluxVol{af, nullptr};
  int rc = guard([&] {
    with_fields(af, [&](auto &F) {
      typedef typename std::remove_reference<decltype(F)>::type FT;
      std::unique_ptr<typename FT::FluxVol> fv(new typename FT::FluxVol);
      fv->wmin = to_vec(vmin);
      fv->wmax = to_vec(vmax);
      fv->d = flux_direction(F.gv, d, fv->wmin, fv->wmax);
      h->fv = fv.get();
      F.flux_vols.push_back(std::move(fv));
      return 0;
    });
  });
  if (rc) {
    delete h;
    return nullptr;
  }
  return reinterpret_cast<mp_flux_vol *>(h);
}
void mp_flux_vol_delete(mp_flux_vol *h) { delete reinterpret_cast<AnyFluxVol *>(h); }
int mp_flux_vol_flux(mp_flux_vol *hp, double *out) {
  AnyFluxVol *h = reinterpret_cast<AnyFluxVol *>(hp);
  return guard([&] {
    with_fields(h->owner, [&](auto &F) {
      typedef typename std::remove_reference<decltype(F)>::type FT;
      *out = static_cast<typename FT::FluxVol *>(h->fv)->cur_flux;
      return 0;
    });
  });
}
static int flux_in_box(mp_fields *f, int d, const double *vmin, const double *vmax, bool sync, double *out) {
  return guard([&] {
    with_fields(reinterpret_cast<AnyFields *>(f), [&](auto &F) {
      const vec3 mn = to_vec(vmin), mx = to_vec(vmax);
      const int dd = flux_direction(F.gv, d, mn, mx);
      *out = sync ? F.flux_in_box(dd, mn, mx) : F.flux_in_box_wrongH(dd, mn, mx);
      return 0;
    });
  });
}
int mp_fields_flux_in_box(mp_fields *f, int d, const double *vmin, const double *vmax, double *out) {
  return flux_in_box(f, d, vmin, vmax, true, out);
}
int mp_fields_flux_in_box_wrongH(mp_fields *f, int d, const double *vmin, const double *vmax, double *out) {
  return flux_in_box(f, d, vmin, vmax, false, out);
}
int mp_fields_energy_in_box(mp_fields *f, int which, const double *vmin, const double *vmax, double *out) {
  return guard([&] {
    with_fields(reinterpret_cast<AnyFields *>(f), [&](auto &F) {
      if (which < 0 || which > 2) fail("energy_in_box: which must be 0 (field), 1 (electric) or 2 (magnetic)");
      const vec3 mn = to_vec(vmin), mx = to_vec(vmax);
      *out = which == 0 ? F.field_energy_in_box(mn, mx) : double(F.energy_in_box(which == 1 ? 0 : 5, mn, mx));
      return 0;
    });
  });
}

static SrcTime src_time_of(const mp_src_time *src) {
  SrcTime st = src->kind == 0 ? SrcTime::gaussian(src->freq, src->width, src->start_time, src->end_time)
                              : SrcTime::continuous(src->freq, src->width, src->start_time, src->end_time, src->slowness);
  st.is_integrated = src->is_integrated != 0;
  return st;
}
int mp_fields_add_volume_source_array(mp_fields *f, int c, const mp_src_time *src, const double *vmin, const double *vmax,
                                      const double *arr, size_t dim1, size_t dim2, size_t dim3, double amp_re,
                                      double amp_im) {
  return guard([&] {
    if (!arr || dim1 == 0 || dim2 == 0 || dim3 == 0) fail("add_volume_source: empty amplitude array");
    const size_t dims[3] = {dim1, dim2, dim3};
    with_fields(reinterpret_cast<AnyFields *>(f), [&](auto &F) {
      F.add_volume_source_array(c, src_time_of(src), to_vec(vmin), to_vec(vmax), arr, dims, cdouble(amp_re, amp_im));
      return 0;
    });
  });
}
int mp_fields_add_srcdata(mp_fields *f, int c, const mp_src_time *src, long n, const long *index, const double *amp,
                          int needs_boundary_fix) {
  return guard([&] {
    if (n < 0) fail("add_srcdata: negative point count");
    std::vector<ptrdiff_t> idx(index, index + n);
    std::vector<cdouble> a(static_cast<size_t>(n));
    for (long q = 0; q < n; ++q) a[size_t(q)] = cdouble(amp[2 * q], amp[2 * q + 1]);
    with_fields(reinterpret_cast<AnyFields *>(f), [&](auto &F) {
      F.add_srcdata(c, src_time_of(src), std::move(idx), std::move(a), needs_boundary_fix != 0);
      return 0;
    });
  });
}
int mp_fields_require_source_components(mp_fields *f) {
  return guard([&] {
    with_fields(reinterpret_cast<AnyFields *>(f), [&](auto &F) {
      F.require_source_components();
      return 0;
    });
  });
}

// fields::add_dft_fields (src/dft.cpp:913-927): one chunk list over all components, no
// integration weights, stored weight 1, centred grid; each component's chunks are prepended.
mp_dft_fields *mp_fields_add_dft_fields(mp_fields *f, const int *comps, int ncomp, const double *vmin, const double *vmax,
                                        const double *freq, int nfreq, int centered, int dec) {
  AnyFields *af = reinterpret_cast<AnyFields *>(f);
  AnyFlux *fl = new AnyFlux;
  fl->owner = af;
  fl->freq.assign(freq, freq + nfreq);
  int rc = guard([&] {
    with_fields(af, [&](auto &F) {
      typedef typename std::remove_reference<decltype(*F.f[0][0])>::type R;
      if (!F.components_allocated) fail("allocate field components (by adding sources) before adding dft objects");
      int decim = dec > 0 ? dec : F.auto_decimation(fl->freq);
      fl->decimation = decim;
      vec3 mn = to_vec(vmin), mx = to_vec(vmax);
      fl->wmin = mn;
      fl->wmax = mx;
      for (int q = 0; q < ncomp; ++q) {
        std::vector<std::unique_ptr<DftChunk<R>>> lst;
        F.add_dft(lst, comps[q], mn, mx, fl->freq, false, cdouble(1.0, 0), decim, centered != 0);
        std::vector<void *> front;
        for (int j = int(lst.size()) - 1; j >= 0; --j) {
          front.push_back(lst[j].get());
          F.dft_chunks.push_back(std::move(lst[j]));
        }
        fl->E.insert(fl->E.begin(), front.begin(), front.end());
      }
      return 0;
    });
  });
  if (rc) {
    delete fl;
    return nullptr;
  }
  return reinterpret_cast<mp_dft_fields *>(fl);
}
void mp_dft_fields_delete(mp_dft_fields *h) { delete reinterpret_cast<AnyFlux *>(h); }
int mp_dft_fields_num_chunks(mp_dft_fields *h) { return mp_dft_flux_num_chunks(reinterpret_cast<mp_dft_flux *>(h), 0); }
long mp_dft_fields_chunk_points(mp_dft_fields *h, int i) {
  return mp_dft_flux_chunk_points(reinterpret_cast<mp_dft_flux *>(h), 0, i);
}
int mp_dft_fields_chunk_component(mp_dft_fields *h, int i) {
  AnyFlux *fl = reinterpret_cast<AnyFlux *>(h);
  if (i < 0 || i >= int(fl->E.size())) {
    last_error = "no such dft chunk";
    return -1;
  }
  return with_fields(fl->owner, [&](auto &F) {
    typedef typename std::remove_reference<decltype(*F.f[0][0])>::type R;
    return reinterpret_cast<DftChunk<R> *>(fl->E[i])->c;
  });
}
int mp_dft_fields_get_dft(mp_dft_fields *h, int i, double *out) {
  return mp_dft_flux_get_dft(reinterpret_cast<mp_dft_flux *>(h), 0, i, out);
}

long mp_dft_fields_get_dft_array(mp_dft_fields *h, int c, int num_freq, int *rank, size_t *dims, double *out) {
  return dft_array(reinterpret_cast<AnyFlux *>(h), false, c, num_freq, rank, dims, out);
}

} // extern "C"

namespace {

// ---- near-to-far transformation (src/near2far.cpp) ----------------------------------------
// The fields EH = (Ex, Ey, Ez, Hx, Hy, Hz) at x radiated by a point current f0 of component c0
// at x0 in the homogeneous medium (eps, mu), at frequency freq, for the e^{-i omega t}
// convention of the DFT accumulators.

// green3d (src/near2far.cpp:133-186): the dipole fields written for a current rather than a
// dipole moment (the moment's fields divided by -i omega).  The radiated field of a current
// of one kind (E for J, H for a magnetic current) is pre (t1 p + t2 (p.r) r), the other kind
// is pre t3 (r x p) divided (J) or multiplied (magnetic current, negated) by the impedance.
void near2far_green3d(cdouble EH[6], const vec3 &x, double freq, double eps, double mu, const vec3 &x0, int c0,
                      cdouble f0) {
  double rh[3], r2 = 0;
  for (int d = 0; d < 3; ++d) {
    rh[d] = x.t[d] - x0.t[d];
    r2 += rh[d] * rh[d];
  }
  const double r = std::sqrt(r2);
  for (int d = 0; d < 3; ++d) rh[d] /= r;
  const double n = std::sqrt(eps * mu), k = 2 * pi * freq * n, kr = k * r, imp = std::sqrt(mu / eps);
  const cdouble ikr(0.0, kr);
  const double ikr2 = -kr * kr;
  cdouble pre = f0 * std::polar(k * n / (4 * pi * r), kr + 0.5 * pi);
  const int pd = component_direction(c0);
  if (pd == NODIR) fail("unrecognized source type");
  const double p[3] = {pd == X ? 1.0 : 0.0, pd == Y ? 1.0 : 0.0, pd == Z ? 1.0 : 0.0};
  const double prh = rh[pd];
  const double rxp[3] = {rh[1] * p[2] - rh[2] * p[1], rh[2] * p[0] - rh[0] * p[2], rh[0] * p[1] - rh[1] * p[0]};
  const cdouble t1 = 1.0 - 1.0 / ikr + 1.0 / ikr2, t2 = (-1.0 + 3.0 / ikr - 3.0 / ikr2) * prh, t3 = 1.0 - 1.0 / ikr;
  const bool electric = is_electric(c0);
  if (!electric && !is_magnetic(c0)) fail("unrecognized source type");
  pre /= electric ? eps : mu;
  cdouble *same = electric ? EH : EH + 3, *other = electric ? EH + 3 : EH;
  for (int d = 0; d < 3; ++d) {
    same[d] = pre * (t1 * p[d] + t2 * rh[d]);
    other[d] = electric ? pre * t3 * rxp[d] / imp : -pre * t3 * rxp[d] * imp;
  }
}

// hankel (src/near2far.cpp:188-204): H_n^(1)(x) = J_n(x) + i Y_n(x), from the C library's jn/yn
cdouble near2far_hankel(int n, double x) { return cdouble(jn(n, x), yn(n, x)); }

// green2d (src/near2far.cpp:207-267): the line-current fields from the Hankel functions of
// order 0, 1 and 2 of k r.  An out-of-plane current (Ez or Hz) radiates its own component
// -omega (mu or eps) H0 / 4 and the in-plane field of the other kind along z x r; an in-plane
// current radiates the out-of-plane field of the other kind (r x p) i k H1 / 4 and its own
// in-plane field from H1 (along r) and H0 - H2 (across r).
void near2far_green2d(cdouble EH[6], const vec3 &x, double freq, double eps, double mu, const vec3 &x0, int c0,
                      cdouble f0) {
  const double dx = x.t[0] - x0.t[0], dy = x.t[1] - x0.t[1], r = std::sqrt(dx * dx + dy * dy);
  const double rx = dx / r, ry = dy / r;
  const double omega = 2 * pi * freq, k = omega * std::sqrt(eps * mu), kr = k * r, imp = std::sqrt(mu / eps);
  const cdouble H0 = near2far_hankel(0, kr) * f0, H1 = near2far_hankel(1, kr) * f0;
  const cdouble ikH1 = 0.25 * cdouble(0.0, k) * H1;
  for (int j = 0; j < 6; ++j) EH[j] = 0.0;
  const int pd = component_direction(c0);
  const bool electric = is_electric(c0);
  if (pd == Z) {
    if (electric) {
      EH[2] = (-0.25 * omega * mu) * H0;
      EH[3] = -ry * ikH1;
      EH[4] = rx * ikH1;
    }
    else {
      EH[0] = ry * ikH1;
      EH[1] = -rx * ikH1;
      EH[5] = (-0.25 * omega * eps) * H0;
    }
    return;
  }
  const cdouble H2 = near2far_hankel(2, kr) * f0;
  const double px = pd == X ? 1.0 : 0.0, py = pd == Y ? 1.0 : 0.0;
  const double prh = px * rx + py * ry, rxp = rx * py - ry * px;
  // the in-plane pair of the source's own kind, and the out-of-plane field of the other kind
  const double radial = prh / r * 0.25 * (electric ? imp : 1.0 / imp), across = rxp * omega * (electric ? mu : eps) * 0.125;
  cdouble *own = electric ? EH : EH + 3;
  own[0] = -(rx * radial) * H1 + (ry * across) * (H0 - H2);
  own[1] = -(ry * radial) * H1 - (rx * across) * (H0 - H2);
  if (electric)
    EH[5] = -rxp * ikH1;
  else
    EH[2] = rxp * ikH1;
}

struct AnyN2F : AnyFlux {
  double eps = 0, mu = 0;
  int periodic_d[2] = {NODIR, NODIR}, periodic_n[2] = {0, 0};
  double periodic_k[2] = {0, 0}, period[2] = {0, 0};
};

// dft_near2far::farfield_lowlevel (src/near2far.cpp:350-397): every accumulator point of every
// chunk (its location plus the chunk's periodic-image shift) is a point current of the chunk's
// equivalent-source component with the accumulator as amplitude; the periodic copies within
// periodic_n periods add with the Bloch phase of their offset.
template <class R> void near2far_point(const AnyN2F &n2f, const GV &gv, const vec3 &xin, cdouble *EH) {
  if (gv.dim != 2 && gv.dim != 3) fail("only 2d or 3d or cylindrical far-field computation is supported");
  vec3 x = xin;
  if (gv.dim == 2) x.t[Z] = 0;
  const size_t Nf = n2f.freq.size();
  for (size_t i = 0; i < 6 * Nf; ++i) EH[i] = 0.0;
  for (void *p : n2f.E) {
    const DftChunk<R> &ch = *reinterpret_cast<const DftChunk<R> *>(p);
    std::vector<vec3> locs;
    loop_ivecs(gv, ch.is, ch.ie, [&](ptrdiff_t, const Loop &L, ptrdiff_t i1, ptrdiff_t i2, ptrdiff_t i3) {
      ivec at = ch.shift;
      const ptrdiff_t ik[3] = {i1, i2, i3};
      for (int k = 0; k < 3; ++k)
        if (L.pres[k]) at.t[L.d[k]] += int(L.is[k] + 2 * ik[k]);
      locs.push_back(gv.loc(at));
    });
    for (size_t i = 0; i < Nf; ++i)
      for (size_t q = 0; q < locs.size(); ++q) {
        const cdouble f0 = ch.acc == MP_DFT_F64 ? ch.dft64[Nf * q + i] : cdouble(ch.dft[Nf * q + i]);
        vec3 xs = locs[q];
        for (int i0 = -n2f.periodic_n[0]; i0 <= n2f.periodic_n[0]; ++i0) {
          if (n2f.periodic_d[0] != NODIR) xs.t[n2f.periodic_d[0]] = locs[q].t[n2f.periodic_d[0]] + i0 * n2f.period[0];
          for (int i1 = -n2f.periodic_n[1]; i1 <= n2f.periodic_n[1]; ++i1) {
            if (n2f.periodic_d[1] != NODIR) xs.t[n2f.periodic_d[1]] = locs[q].t[n2f.periodic_d[1]] + i1 * n2f.period[1];
            const cdouble cphase = std::polar(1.0, i0 * n2f.periodic_k[0] + i1 * n2f.periodic_k[1]);
            cdouble EH6[6];
            if (gv.dim == 2)
              near2far_green2d(EH6, x, n2f.freq[i], n2f.eps, n2f.mu, xs, ch.vc, f0);
            else
              near2far_green3d(EH6, x, n2f.freq[i], n2f.eps, n2f.mu, xs, ch.vc, f0);
#ifdef MUT_N2F_PHASE
            for (int j = 0; j < 6; ++j) EH[i * 6 + j] += EH6[j];
#else
            for (int j = 0; j < 6; ++j) EH[i * 6 + j] += EH6[j] * cphase;
#endif
          }
        }
      }
  }
}

// The output grid of dft_near2far::get_farfields_array (src/near2far.cpp:411-427): floor(extent
// resolution) points along each direction of the cell's dimensionality, spaced extent / (n - 1)
// from the volume's low corner (one point at the low corner when n <= 1).
struct FarGrid {
  int rank = 0, dirs[3] = {X, Y, Z};
  size_t dims[3] = {1, 1, 1}, N = 1;
  double dx[3] = {0, 0, 0};
};
FarGrid far_grid(const GV &gv, const vec3 &mn, const vec3 &mx, double resolution) {
  FarGrid g;
  for (int d = 0; d < 3; ++d) {
    if (!gv.has(d)) continue;
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

// dft_near2far::get_farfields_array (src/near2far.cpp:411-483): the fields at every grid point
// as a 6 x 2 x N x Nfreq row-major array (component, real/imaginary part, point, frequency),
// the singleton dimensions then dropped from dims.
template <class R> std::vector<double> farfields_array(const AnyN2F &n2f, const GV &gv, const vec3 &mn, const vec3 &mx,
                                                       double resolution, int *rank, size_t *dims) {
  const FarGrid g = far_grid(gv, mn, mx, resolution);
  const size_t Nf = n2f.freq.size(), N = g.N;
  std::vector<double> out(6 * 2 * N * Nf, 0.0);
  std::vector<cdouble> EH1(6 * Nf);
  for (size_t i0 = 0; i0 < g.dims[0]; ++i0)
    for (size_t i1 = 0; i1 < g.dims[1]; ++i1)
      for (size_t i2 = 0; i2 < g.dims[2]; ++i2) {
        vec3 x;
        const size_t ii[3] = {i0, i1, i2};
        for (int r = 0; r < g.rank; ++r) x.t[g.dirs[r]] = mn.t[g.dirs[r]] + double(ii[r]) * g.dx[r];
        near2far_point<R>(n2f, gv, x, EH1.data());
        const size_t idx = (i0 * g.dims[1] + i1) * g.dims[2] + i2;
        for (size_t i = 0; i < Nf; ++i)
          for (int k = 0; k < 6; ++k) {
            out[((k * 2 + 0) * N + idx) * Nf + i] = EH1[i * 6 + k].real();
            out[((k * 2 + 1) * N + idx) * Nf + i] = EH1[i * 6 + k].imag();
          }
      }
  int rr = 0;
  dims[0] = dims[1] = dims[2] = 1;
  for (int r = 0; r < g.rank; ++r)
    if (g.dims[r] > 1) dims[rr++] = g.dims[r];
  *rank = rr;
  return out;
}

// dft_near2far::flux (src/near2far.cpp:517-565): the far-field Poynting flux along df summed
// over the grid, Re(E1 conj(H1)) - Re(E2 conj(H2)) with the pair of transverse components in
// cyclic order, times the grid's cell volume (the spacing of each direction of more than one
// point).
void near2far_flux_sum(const std::vector<double> &EH, const FarGrid &g, size_t Nf, int df, const GV &gv, const vec3 &mn,
                       const vec3 &mx, double resolution, double *F) {
  int cE[2], cH[2];
  if (df == X)
    cE[0] = 1, cE[1] = 2, cH[0] = 5, cH[1] = 4;
  else if (df == Y)
    cE[0] = 2, cE[1] = 0, cH[0] = 3, cH[1] = 5;
  else
    cE[0] = 0, cE[1] = 1, cH[0] = 4, cH[1] = 3;
  const size_t N = g.N;
  for (size_t i = 0; i < Nf; ++i) F[i] = 0;
  auto at = [&](int k, size_t idx, size_t i) {
    return cdouble(EH[((k * 2 + 0) * N + idx) * Nf + i], EH[((k * 2 + 1) * N + idx) * Nf + i]);
  };
  for (size_t idx = 0; idx < N; ++idx)
    for (size_t i = 0; i < Nf; ++i)
      for (int j = 0; j < 2; ++j) F[i] += std::real(at(cE[j], idx, i) * std::conj(at(cH[j], idx, i))) * (1 - 2 * j);
  double dV = 1;
  for (int d = 0; d < 3; ++d)
    if (gv.has(d)) {
      const int n = int(std::floor((mx.t[d] - mn.t[d]) * resolution));
      if (n > 1) dV *= (mx.t[d] - mn.t[d]) / (n - 1);
    }
  for (size_t i = 0; i < Nf; ++i) F[i] *= dV;
}

bool n2f_approxeq(double a, double b) { return std::fabs(a - b) < 0.5e-11 * (std::fabs(a) + std::fabs(b)); }

}  // namespace

extern "C" {

// fields::add_dft_near2far (src/near2far.cpp:569-654): for each surface (normal nd from its
// direction or, when that is -1, volume::normal_direction) the tangential E and H components
// of the two directions transverse to nd in cyclic order are accumulated with their
// integration weights, off the centred grid, as the equivalent currents n x H (electric,
// component of the other tangential direction) and -n x E (magnetic): sign +1 for the first
// tangential direction and -1 for the second, negated for E, times the surface's (complex)
// weight, given as (re, im) pairs.  The
// medium is eps, mu at the surface's centre and must be the same for all surfaces; with
// nperiods > 1 a tangential direction that is periodic and spanned by the surface is summed
// over nperiods copies each way.
mp_near2far *mp_fields_add_dft_near2far(mp_fields *f, int nsurf, const int *dirs, const double *vmins,
                                        const double *vmaxs, const double *weights, const double *freq, int nfreq,
                                        int dec, int nperiods) {
  AnyFields *af = reinterpret_cast<AnyFields *>(f);
  AnyN2F *n2f = new AnyN2F;
  n2f->owner = af;
  n2f->freq.assign(freq, freq + nfreq);
  int rc = guard([&] {
    with_fields(af, [&](auto &F) {
      typedef typename std::remove_reference<decltype(*F.f[0][0])>::type R;
      if (!F.components_allocated) fail("allocate field components (by adding sources) before adding dft objects");
      if (nsurf < 1) fail("add_dft_near2far: no surfaces");
      const int decim = dec > 0 ? dec : F.auto_decimation(n2f->freq);
      n2f->decimation = decim;
      // dft_near2far::where is the first surface's volume (the union in the loop takes where->v)
      n2f->wmin = to_vec(vmins);
      n2f->wmax = to_vec(vmaxs);
      std::vector<void *> front;
      for (int w = 0; w < nsurf; ++w) {
        const vec3 mn = to_vec(vmins + 3 * w), mx = to_vec(vmaxs + 3 * w);
        int nd = dirs[w];
        if (nd < 0) nd = flux_direction(F.gv, -1, mn, mx);
        if (nd > Z) fail("unknown dft_near2far normal");
        vec3 cen;
        for (int d = 0; d < 3; ++d) cen.t[d] = 0.5 * (mn.t[d] + mx.t[d]);
        const double weps = F.get_eps_mu(cen, false), wmu = F.get_eps_mu(cen, true);
        if (w > 0 && !(n2f_approxeq(n2f->eps, weps) && n2f_approxeq(n2f->mu, wmu)))
          fail("dft_near2far requires surfaces in a homogeneous medium");
        n2f->eps = weps;
        n2f->mu = wmu;
        const int fd[2] = {(nd + 1) % 3, (nd + 2) % 3};
        if (nperiods > 1)
          for (int i = 0; i < 2; ++i) {
            const double user_width = F.gv.num[fd[i]] * F.gv.inva;
            if (F.gv.has(fd[i]) && F.boundaries[High][fd[i]] == Periodic && F.boundaries[Low][fd[i]] == Periodic &&
                float(mx.t[fd[i]] - mn.t[fd[i]]) >= float(user_width)) {
              n2f->periodic_d[i] = fd[i];
              n2f->periodic_n[i] = nperiods;
              n2f->period[i] = user_width;
              n2f->periodic_k[i] = 2 * pi * F.k[fd[i]].real() * user_width;
            }
          }
        for (int i = 0; i < 2; ++i)
          for (int j = 0; j < 2; ++j) {
            const int c = direction_component(i == 0 ? 0 : 5, fd[j]);
            const int c0 = direction_component(i == 0 ? 5 : 0, fd[1 - j]);
            double s = j == 0 ? 1 : -1;
#ifndef MUT_N2F_SIGN
            if (is_electric(c)) s = -s;
#endif
            std::vector<std::unique_ptr<DftChunk<R>>> lst;
            F.add_dft(lst, c, mn, mx, n2f->freq, true, s * cdouble(weights[2 * w], weights[2 * w + 1]), decim, false, c0);
            // each call's chunks go in front of the list built so far
            std::vector<void *> mine;
            for (int q = int(lst.size()) - 1; q >= 0; --q) {
              mine.push_back(lst[q].get());
              F.dft_chunks.push_back(std::move(lst[q]));
            }
            front.insert(front.begin(), mine.begin(), mine.end());
          }
      }
      n2f->E = front;
      return 0;
    });
  });
  if (rc) {
    delete n2f;
    return nullptr;
  }
  return reinterpret_cast<mp_near2far *>(n2f);
}
void mp_near2far_delete(mp_near2far *h) { delete reinterpret_cast<AnyN2F *>(h); }
int mp_near2far_info(const mp_near2far *h, double *eps_mu, int *periodic, long *npoints) {
  const AnyN2F *n = reinterpret_cast<const AnyN2F *>(h);
  eps_mu[0] = n->eps, eps_mu[1] = n->mu;
  for (int i = 0; i < 2; ++i) periodic[2 * i] = n->periodic_d[i] == NODIR ? -1 : n->periodic_d[i], periodic[2 * i + 1] = n->periodic_n[i];
  *npoints = with_fields(n->owner, [&](auto &F) {
    typedef typename std::remove_reference<decltype(*F.f[0][0])>::type R;
    long np = 0;
    for (void *p : n->E) np += long(reinterpret_cast<const DftChunk<R> *>(p)->N);
    return np;
  });
  return int(n->freq.size());
}
int mp_near2far_farfield(mp_near2far *h, const double *x, double *EH) {
  AnyN2F *n = reinterpret_cast<AnyN2F *>(h);
  return guard([&] {
    with_fields(n->owner, [&](auto &F) {
      typedef typename std::remove_reference<decltype(*F.f[0][0])>::type R;
      std::vector<cdouble> v(6 * n->freq.size());
      near2far_point<R>(*n, F.gv, to_vec(x), v.data());
      for (size_t q = 0; q < v.size(); ++q) EH[2 * q] = v[q].real(), EH[2 * q + 1] = v[q].imag();
      return 0;
    });
  });
}
long mp_near2far_get_farfields_array(mp_near2far *h, const double *wmin, const double *wmax, double resolution, int *rank,
                                     size_t *dims, double *out) {
  AnyN2F *n = reinterpret_cast<AnyN2F *>(h);
  long count = -1;
  const int rc = guard([&] {
    with_fields(n->owner, [&](auto &F) {
      typedef typename std::remove_reference<decltype(*F.f[0][0])>::type R;
      const vec3 mn = to_vec(wmin), mx = to_vec(wmax);
      const FarGrid g = far_grid(F.gv, mn, mx, resolution);
      count = long(g.N);
      if (!out) {
        int rr = 0;
        dims[0] = dims[1] = dims[2] = 1;
        for (int r = 0; r < g.rank; ++r)
          if (g.dims[r] > 1) dims[rr++] = g.dims[r];
        *rank = rr;
        return 0;
      }
      std::vector<double> a = farfields_array<R>(*n, F.gv, mn, mx, resolution, rank, dims);
      std::copy(a.begin(), a.end(), out);
      return 0;
    });
  });
  return rc ? -1 : count;
}
int mp_near2far_flux(mp_near2far *h, int d, const double *wmin, const double *wmax, double resolution, double *out) {
  AnyN2F *n = reinterpret_cast<AnyN2F *>(h);
  return guard([&] {
    with_fields(n->owner, [&](auto &F) {
      typedef typename std::remove_reference<decltype(*F.f[0][0])>::type R;
      if (d < 0 || d > Z || !F.gv.has(d)) fail("cannot get flux for near2far: co-ordinate mismatch");
      const vec3 mn = to_vec(wmin), mx = to_vec(wmax);
      int rank;
      size_t dims[3];
      std::vector<double> a = farfields_array<R>(*n, F.gv, mn, mx, resolution, &rank, dims);
      near2far_flux_sum(a, far_grid(F.gv, mn, mx, resolution), n->freq.size(), d, F.gv, mn, mx, resolution, out);
      return 0;
    });
  });
}
long mp_near2far_get_dft_array(mp_near2far *h, int c, int num_freq, int *rank, size_t *dims, double *out) {
  return dft_array(reinterpret_cast<AnyFlux *>(h), false, c, num_freq, rank, dims, out);
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
// dft_near2far::scale_dfts (src/near2far.cpp:124-126)
int mp_near2far_scale_dfts(mp_near2far *n, double re, double im) {
  return mp_dft_flux_scale_dfts(reinterpret_cast<mp_dft_flux *>(n), re, im);
}
// green2d / green3d as library routines (fields of a point current c0 of amplitude f0 at x0)
int mp_near2far_green(int dim, const double *x, double freq, double eps, double mu, const double *x0, int c0,
                      const double *f0, double *EH) {
  return guard([&] {
    cdouble v[6];
    vec3 a = to_vec(x), b = to_vec(x0);
    if (dim == 2) {
      a.t[Z] = b.t[Z] = 0;
      near2far_green2d(v, a, freq, eps, mu, b, c0, cdouble(f0[0], f0[1]));
    }
    else if (dim == 3)
      near2far_green3d(v, a, freq, eps, mu, b, c0, cdouble(f0[0], f0[1]));
    else
      fail("only 2d or 3d Green's functions");
    for (int j = 0; j < 6; ++j) EH[2 * j] = v[j].real(), EH[2 * j + 1] = v[j].imag();
    return 0;
  });
}

// The reverse mode is a product feature with no counterpart in the reference's C++ library;
// its gradients are checked against finite differences of the engine itself, so the oracle
// exports the entry points only to keep one ABI and reports that it does not provide them.
mp_adjoint *mp_adjoint_new(mp_fields *, mp_dft_fields *, int, const double *, const double *, int, int *, int *) {
  last_error = "the oracle has no reverse mode";
  return nullptr;
}
void mp_adjoint_delete(mp_adjoint *) {}
int mp_adjoint_objective(mp_adjoint *, double *) {
  last_error = "the oracle has no reverse mode";
  return -1;
}
int mp_adjoint_gradient(mp_adjoint *, double *) {
  last_error = "the oracle has no reverse mode";
  return -1;
}

// Batches are a product feature with no counterpart in the reference; batched runs are checked
// against the same objects stepped one at a time.
mp_batch *mp_batch_new(mp_fields *const *, int) {
  last_error = "the oracle has no batches";
  return nullptr;
}
void mp_batch_delete(mp_batch *) {}
int mp_batch_step_n(mp_batch *, long) {
  last_error = "the oracle has no batches";
  return -1;
}
int mp_batch_set_options(mp_batch *, int, int, int) {
  last_error = "the oracle has no batches";
  return -1;
}
int mp_batch_stats(mp_batch *, int *, double *) {
  last_error = "the oracle has no batches";
  return -1;
}

// fields::field_energy (src/energy_and_flux.cpp:48)
int mp_fields_field_energy(mp_fields *f, double *out) {
  return guard([&] {
    with_fields(reinterpret_cast<AnyFields *>(f), [&](auto &F) {
      *out = F.field_energy();
      return 0;
    });
  });
}


// ---- dft_energy (src/dft.cpp:654-750) ------------------------------------------------------
struct AnyEnergy : AnyFlux {
  std::vector<void *> D, B;
};
// fields::add_dft_energy (src/dft.cpp:725-750): for every volume and every field direction of
// the grid, E and H with integration weights and D and B without, on the centred grid, each
// call's chunks in front of its list
mp_dft_energy *mp_fields_add_dft_energy(mp_fields *f, int nvol, const double *vmins, const double *vmaxs, const double *freq,
                                        int nfreq, int dec) {
  AnyFields *af = reinterpret_cast<AnyFields *>(f);
  AnyEnergy *en = new AnyEnergy;
  en->owner = af;
  en->diag = &en->D;
  en->extra = &en->B;
  en->freq.assign(freq, freq + nfreq);
  int rc = guard([&] {
    with_fields(af, [&](auto &F) {
      typedef typename std::remove_reference<decltype(*F.f[0][0])>::type R;
      if (!F.components_allocated) fail("allocate field components (by adding sources) before adding dft objects");
      if (nvol < 1) fail("add_dft_energy: no volumes");
      const int decim = dec > 0 ? dec : F.auto_decimation(en->freq);
      en->decimation = decim;
      en->wmin = to_vec(vmins);
      en->wmax = to_vec(vmaxs);
      std::vector<void *> *lists[4] = {&en->E, &en->D, &en->H, &en->B};
      const int base[4] = {0, 10, 5, 15};
      for (int w = 0; w < nvol; ++w) {
        const vec3 mn = to_vec(vmins + 3 * w), mx = to_vec(vmaxs + 3 * w);
        // LOOP_OVER_FIELD_DIRECTIONS (src/meep/vec.hpp:147-149): X, Y, Z; absent components add
        // no chunks
        for (int d = 0; d < 3; ++d) {
          for (int q = 0; q < 4; ++q) {
            const int c = direction_component(base[q], d);
            std::vector<std::unique_ptr<DftChunk<R>>> lst;
#ifdef MUT_DFTENERGY_WEIGHT
            F.add_dft(lst, c, mn, mx, en->freq, true, cdouble(1.0, 0), decim, true);
#else
            F.add_dft(lst, c, mn, mx, en->freq, q == 0 || q == 2, cdouble(1.0, 0), decim, true);
#endif
            std::vector<void *> mine;
            for (int j = int(lst.size()) - 1; j >= 0; --j) {
              mine.push_back(lst[j].get());
              F.dft_chunks.push_back(std::move(lst[j]));
            }
            lists[q]->insert(lists[q]->begin(), mine.begin(), mine.end());
          }
        }
      }
      return 0;
    });
  });
  if (rc) {
    delete en;
    return nullptr;
  }
  return reinterpret_cast<mp_dft_energy *>(en);
}
void mp_dft_energy_delete(mp_dft_energy *e) { delete reinterpret_cast<AnyEnergy *>(e); }
// dft_energy::electric / magnetic / total (src/dft.cpp:681-723): 0.5 Re(conj(F) G) summed over
// the paired chunks of the E and D (H and B) lists, in the accumulator's type
int mp_dft_energy_get(mp_dft_energy *ep, int which, double *out) {
  AnyEnergy *en = reinterpret_cast<AnyEnergy *>(ep);
  return guard([&] {
    if (which < 0 || which > 2) fail("dft_energy: which must be 0 (total), 1 (electric) or 2 (magnetic)");
    with_fields(en->owner, [&](auto &F) {
      typedef typename std::remove_reference<decltype(*F.f[0][0])>::type R;
      const size_t Nf = en->freq.size();
      std::vector<double> e(Nf, 0.0), m(Nf, 0.0);
      auto part = [&](const std::vector<void *> &A, const std::vector<void *> &B, std::vector<double> &acc) {
        for (size_t q = 0; q < A.size() && q < B.size(); ++q) {
          auto *a = reinterpret_cast<DftChunk<R> *>(A[q]);
          auto *b = reinterpret_cast<DftChunk<R> *>(B[q]);
          for (size_t k = 0; k < a->N; ++k)
            for (size_t i = 0; i < Nf; ++i) {
              if (a->acc == MP_DFT_F64)
                acc[i] += 0.5 * std::real(std::conj(a->dft64[k * Nf + i]) * b->dft64[k * Nf + i]);
              else
                acc[i] += 0.5 * double(std::real(std::conj(a->dft[k * Nf + i]) * b->dft[k * Nf + i]));
            }
        }
      };
      part(en->E, en->D, e);
      part(en->H, en->B, m);
      for (size_t i = 0; i < Nf; ++i) out[i] = which == 1 ? e[i] : which == 2 ? m[i] : e[i] + m[i];
      return 0;
    });
  });
}

// ---- dft_force (src/stress.cpp:90-200) -----------------------------------------------------
struct AnyForce : AnyFlux {
  std::vector<void *> D; // the diag list (E list: offdiag1, H list: offdiag2)
};
// fields::add_dft_force (src/stress.cpp:153-198): for each volume, normal nd from the volume
// and force direction fd from its component; fd != nd: offdiag1 takes E and H along fd with
// integration weights and the volume's weight, offdiag2 E and H along nd unweighted, both on
// the centred grid; fd == nd: diag takes E and H of every field direction with square-root
// weights on their own grids, the readout factor +-0.5 times the weight (+ along fd).
mp_dft_force *mp_fields_add_dft_force(mp_fields *f, int nvol, const int *comps, const double *vmins, const double *vmaxs,
                                      const double *weights, const double *freq, int nfreq, int dec) {
  AnyFields *af = reinterpret_cast<AnyFields *>(f);
  AnyForce *fo = new AnyForce;
  fo->owner = af;
  fo->diag = &fo->D;
  fo->freq.assign(freq, freq + nfreq);
  int rc = guard([&] {
    with_fields(af, [&](auto &F) {
      typedef typename std::remove_reference<decltype(*F.f[0][0])>::type R;
      if (!F.components_allocated) fail("allocate field components (by adding sources) before adding dft objects");
      if (nvol < 1) fail("add_dft_force: no volumes");
      const int decim = dec > 0 ? dec : F.auto_decimation(fo->freq);
      fo->decimation = decim;
      fo->wmin = to_vec(vmins);
      fo->wmax = to_vec(vmaxs);
      auto push = [&](std::vector<void *> &to, std::vector<std::unique_ptr<DftChunk<R>>> &lst) {
        std::vector<void *> mine;
        for (int j = int(lst.size()) - 1; j >= 0; --j) {
          mine.push_back(lst[j].get());
          F.dft_chunks.push_back(std::move(lst[j]));
        }
        to.insert(to.begin(), mine.begin(), mine.end());
      };
      for (int w = 0; w < nvol; ++w) {
        const vec3 mn = to_vec(vmins + 3 * w), mx = to_vec(vmaxs + 3 * w);
        const int nd = flux_direction(F.gv, -1, mn, mx);
        const int fd = comps[w] >= 100 ? (comps[w] == 100 ? X : comps[w] == 101 ? Y : comps[w] == 104 ? Z : NODIR)
                                       : component_direction(comps[w]);
        if (fd == NODIR) fail("NO_DIRECTION dft_force is invalid");
        const cdouble wt(weights[2 * w], weights[2 * w + 1]);
        if (fd != nd) {
          for (int base : {0, 5}) {
            std::vector<std::unique_ptr<DftChunk<R>>> l1, l2;
            F.add_dft(l1, direction_component(base, fd), mn, mx, fo->freq, true, wt, decim, true);
            F.add_dft(l2, direction_component(base, nd), mn, mx, fo->freq, false, cdouble(1.0, 0), decim, true);
            push(fo->E, l1);
            push(fo->H, l2);
          }
        }
        else
          for (int d = 0; d < 3; ++d) {
#ifdef MUT_FORCE_DIAG
            const cdouble w1 = wt * 0.5;
#else
            const cdouble w1 = wt * (d == fd ? 0.5 : -0.5);
#endif
            for (int base : {0, 5}) {
              std::vector<std::unique_ptr<DftChunk<R>>> l;
              F.add_dft(l, direction_component(base, d), mn, mx, fo->freq, true, cdouble(1.0, 0), decim, false);
              for (auto &ch : l) ch->sqrt_w = true, ch->extra_weight = w1;
              push(fo->D, l);
            }
          }
      }
      return 0;
    });
  });
  if (rc) {
    delete fo;
    return nullptr;
  }
  return reinterpret_cast<mp_dft_force *>(fo);
}
void mp_dft_force_delete(mp_dft_force *h) { delete reinterpret_cast<AnyForce *>(h); }
// dft_force::scale_dfts (src/stress.cpp:144-148) and dft_energy::scale_dfts (src/dft.cpp:785-790)
int mp_dft_force_scale_dfts(mp_dft_force *h, double re, double im) {
  return mp_dft_flux_scale_dfts(reinterpret_cast<mp_dft_flux *>(h), re, im);
}
int mp_dft_energy_scale_dfts(mp_dft_energy *h, double re, double im) {
  return mp_dft_flux_scale_dfts(reinterpret_cast<mp_dft_flux *>(h), re, im);
}
// fields::get_dft_array(dft_force, component, num_freq, rank, dims) (src/dft.cpp:1270-1280)
long mp_dft_force_get_dft_array(mp_dft_force *h, int c, int num_freq, int *rank, size_t *dims, double *out) {
  return dft_array(reinterpret_cast<AnyFlux *>(h), true, c, num_freq, rank, dims, out);
}
// dft_force::force (src/stress.cpp:90-114): stress_sum of (offdiag1, offdiag2) and (diag, diag),
// Re(extra_weight F1 conj(F2)) with the accumulators widened to complex<double>
int mp_dft_force_force(mp_dft_force *h, double *out) {
  AnyForce *fo = reinterpret_cast<AnyForce *>(h);
  return guard([&] {
    with_fields(fo->owner, [&](auto &F) {
      typedef typename std::remove_reference<decltype(*F.f[0][0])>::type R;
      const size_t Nf = fo->freq.size();
      for (size_t i = 0; i < Nf; ++i) out[i] = 0;
      auto sum = [&](const std::vector<void *> &A, const std::vector<void *> &B) {
        for (size_t q = 0; q < A.size() && q < B.size(); ++q) {
          auto *a = reinterpret_cast<DftChunk<R> *>(A[q]);
          auto *b = reinterpret_cast<DftChunk<R> *>(B[q]);
          for (size_t k = 0; k < a->N; ++k)
            for (size_t i = 0; i < Nf; ++i) {
              const cdouble x = a->acc == MP_DFT_F64 ? a->dft64[k * Nf + i] : cdouble(a->dft[k * Nf + i]);
              const cdouble y = b->acc == MP_DFT_F64 ? b->dft64[k * Nf + i] : cdouble(b->dft[k * Nf + i]);
              out[i] += std::real(a->extra_weight * x * std::conj(y));
            }
        }
      };
      sum(fo->E, fo->H);
      sum(fo->D, fo->D);
      return 0;
    });
  });
}
} // extern "C"
