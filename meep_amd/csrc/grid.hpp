// grid.hpp — Yee-grid geometry of one chunk (host side).
//
// Integer coordinates are in half pixels with the origin at the cell's low corner, as the
// reference's ivec / grid_volume (src/meep/vec.hpp:751-1177, src/vec.cpp:262-605).  Only the
// Cartesian D1/D2/D3 cases exist here.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>

namespace mx {

constexpr int NC = 20; // component enum size (src/meep/vec.hpp:28)
constexpr int NODIR = -1;

inline bool is_electric(int c) { return c >= 0 && c < 5; }
inline bool is_magnetic(int c) { return c >= 5 && c < 10; }
inline bool is_D(int c) { return c >= 10 && c < 15; }
inline bool is_B(int c) { return c >= 15 && c < 20; }
inline int field_type(int c) { return is_electric(c) ? 0 : is_magnetic(c) ? 1 : is_D(c) ? 2 : 3; }
inline int comp_dir(int c) {
  switch (c % 5) {
    case 0: return 0;
    case 1: return 1;
    case 4: return 2;
  }
  return NODIR;
}
inline int dir_comp(int c, int d) { return (c / 5) * 5 + (d == 2 ? 4 : d); }
inline int cycle_dir(int d, int shift) { return (d + shift + 99) % 3; }

struct IV {
  int t[3] = {0, 0, 0};
};
struct V3 {
  double t[3] = {0, 0, 0};
};

struct Geom {
  int dim = 3;
  double a = 1, inva = 1;
  int num[3] = {0, 0, 0};
  int64_t st[3] = {0, 0, 0}; // strides of this layout
  size_t ntot = 1;           // points of this layout (padding included)
  int64_t rst[3] = {0, 0, 0}; // strides of the reference layout
  size_t rntot = 1;           // points of the reference layout
  int fast = 2;               // the stride-1 direction

  bool has(int d) const { return dim == 1 ? d == 2 : (dim == 2 ? d != 2 : true); }
  int ydir(int k) const { return dim == 2 ? (k + 2) % 3 : k; }

  void init(int d, double aa, int nx, int ny, int nz) {
    dim = d;
    a = aa;
    inva = 1.0 / aa;
    num[0] = has(0) ? nx : 0;
    num[1] = has(1) ? ny : 0;
    num[2] = has(2) ? nz : 0;
    ntot = 1;
    for (int q = 0; q < 3; ++q)
      if (has(q)) ntot *= size_t(num[q] + 1);
    st[0] = has(0) ? int64_t(num[2] + 1) * (num[1] + 1) : 0;
    st[1] = has(1) ? int64_t(num[2] + 1) : 0;
    st[2] = has(2) ? 1 : 0;
    fast = dim == 2 ? 1 : 2;
    for (int q = 0; q < 3; ++q) rst[q] = st[q];
    rntot = ntot;
  }
  // Device layout: the stride-1 direction's rows padded to a multiple of `mult` points
  // (the padding belongs to no component).  Index arithmetic is unchanged otherwise.
  void pad_rows(int mult) {
    const int64_t row = (int64_t(num[fast] + 1) + mult - 1) / mult * mult;
    int64_t s = 1;
    st[0] = st[1] = st[2] = 0;
    st[fast] = 1;
    s = row;
    // remaining directions, fastest first (3D: Y then X; 2D: X)
    for (int d = 2; d >= 0; --d)
      if (has(d) && d != fast) {
        st[d] = s;
        s *= num[d] + 1;
      }
    ntot = size_t(s);
  }
  // reference-layout index of the point with array coordinates j
  int64_t ref_of(const int j[3]) const { return j[0] * rst[0] + j[1] * rst[1] + j[2] * rst[2]; }
  int64_t dev_of(const int j[3]) const { return j[0] * st[0] + j[1] * st[1] + j[2] * st[2]; }
  // visit every point as (reference index, device index)
  template <class F> void for_each_point(F fn) const {
    int hi[3];
    for (int d = 0; d < 3; ++d) hi[d] = has(d) ? num[d] : 0;
    int j[3];
    for (j[0] = 0; j[0] <= hi[0]; ++j[0])
      for (j[1] = 0; j[1] <= hi[1]; ++j[1])
        for (j[2] = 0; j[2] <= hi[2]; ++j[2]) fn(ref_of(j), dev_of(j));
  }
  int ysh(int c, int d) const {
    if (!has(d)) return 0;
    if (c == 20) return 1;
    bool s = ((is_electric(c) || is_D(c)) && d == comp_dir(c)) ||
             ((is_magnetic(c) || is_B(c)) && d != comp_dir(c));
    return s ? 1 : 0;
  }
  // owned array-index range of component c along d: [1 - ysh, n - ysh] (src/vec.cpp:432-463)
  int own_lo(int c, int d) const { return has(d) ? 1 - ysh(c, d) : 0; }
  int own_hi(int c, int d) const { return has(d) ? num[d] - ysh(c, d) : 0; }
  bool owns(const IV &p) const {
    for (int d = 0; d < 3; ++d)
      if (has(d) && !(p.t[d] > 0 && p.t[d] <= 2 * num[d])) return false;
    return true;
  }
  int64_t index(int c, const IV &p) const {
    int64_t idx = 0;
    for (int d = 0; d < 3; ++d)
      if (has(d)) idx += (p.t[d] - ysh(c, d)) / 2 * st[d];
    return idx;
  }
  IV iloc(int c, const int j[3]) const {
    IV r;
    for (int d = 0; d < 3; ++d)
      if (has(d)) r.t[d] = 2 * j[d] + ysh(c, d);
    return r;
  }
  V3 loc(const IV &p) const {
    V3 r;
    for (int d = 0; d < 3; ++d)
      if (has(d)) r.t[d] = p.t[d] * (0.5 * inva);
    return r;
  }
  V3 center() const {
    IV c;
    for (int d = 0; d < 3; ++d)
      if (has(d)) c.t[d] = num[d] - num[d] % 2;
    return loc(c);
  }
  // yee2cent_offsets (src/vec.cpp:333-344)
  void yee2cent(int c, int64_t &o1, int64_t &o2) const {
    o1 = o2 = 0;
    for (int d = 0; d < 3; ++d)
      if (has(d) && !ysh(c, d)) {
        if (o1)
          o2 = st[d];
        else
          o1 = st[d];
      }
  }
  IV round_vec(const V3 &p) const {
    IV r;
    for (int d = 0; d < 3; ++d)
      if (has(d)) {
        double x = p.t[d] * 2 * a;
        r.t[d] = int(std::floor(std::fabs(x) + 0.5) * (x < 0 ? -1 : 1));
      }
    return r;
  }
  // interpolate(component, vec, ivec[8], double[8]) — src/vec.cpp:558-605
  bool interpolate(int c, const V3 &pc, IV locs[8], double w[8]) const;
};

} // namespace mx
