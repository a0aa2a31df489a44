// engine.hpp — parameter blocks shared by the host plan (fields.cpp) and the CDNA4 kernels
// (kernels.hip).  Field arrays keep the reference's organisation — one array per component
// and real/imaginary part over the (n+1)^d points of the chunk, in the reference's loop order
// (src/vec.cpp:293-296,482-494) — with the stride-1 rows padded (see Grid).  A thread of the
// generic kernel handles every component of one point; the interior kernel handles a short
// run of points along the stride-1 axis.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mx {

constexpr int MAXPOLE = 4;  // Lorentz/Drude poles per field type
constexpr int SRC_TAB = 16; // entries of a short source list table

// Geometry of one chunk in the device layout, plus the box a launch covers.  Loop axis k
// (0 slowest, 2 stride-1) walks direction axdir[k] (-1 when the grid has no such direction);
// this is the reference's yucky_direction order (src/vec.cpp:262-270): X,Y,Z in 3D,
// Z(empty),X,Y in 2D, Z in 1D.  The stride-1 axis is padded to a multiple of 4 points so
// every row starts 16-byte aligned; padding points belong to no component.
struct Grid {
  int64_t st[3];   // device strides per direction X,Y,Z (0 when absent)
  int axdir[3];    // direction of each loop axis
  int num[3];      // cells per direction X,Y,Z (0 when absent)
  int hi_metal;    // bit d set: the High wall normal to d is Metallic (src/boundaries.cpp:191-205)
  int lo_magnetic; // bit d set: the Low wall normal to d is Magnetic
  int blo[3];      // launch box: lowest index per direction
  int bhi[3];      // launch box: highest index per direction
  // The launch box by loop axis, derived from the above just before a launch (finalize_args
  // in fields.cpp).  Kernels index only these and the per-slot loop-axis data, with constant
  // indices: a run-time index into the kernel's parameter block makes the compiler copy the
  // whole block into scratch memory.
  int alo[3];      // lowest index along loop axis k (0 when absent)
  int an[3];       // points along loop axis k (1 when absent)
  int64_t ast[3];  // device stride along loop axis k (0 when absent)
  int nax[3];      // cells along loop axis k
  // Periodic halos folded into the kernels along loop axis k when the Bloch phase there is
  // real: +1 or -1 (the phase), 0 when not folded.  A component with no Yee shift along the
  // axis copies its owned point n into the halo point 0, one with a shift copies point 0
  // into the halo point n (the reference's connections, src/boundaries.cpp:368-600), and a
  // point that wraps along several folded axes writes every combination.
  int wrap[3];
  int mhi_bits, mlo_bits; // bit k: loop axis k's High wall is Metallic / Low wall Magnetic
  // indices along loop axis k where no uPML profile of that axis is active for either Yee
  // shift (empty, qlo > qhi, when not known): the kernels skip the profile loads there
  int qlo[3], qhi[3];
};

// Per-slot data by loop axis (derived by finalize_args): owned index range, and the index of
// the plane zeroed by a High metallic wall (mhi) or a Low magnetic wall (mlo), -1 for none
// (on_metal_boundary, src/boundaries.cpp:191-205), and the Yee shift along each loop axis.
struct AxisSlot {
  int lo[3], hi[3];
  int mhi[3], mlo[3];
  int ysh[3];
  int ybits; // bit k: ysh[k] (the vectorised kernel derives the rest from it and the Grid)
};

template <class R> struct PmlProfile {
  const R *sig[3], *kap[3], *siginv[3]; // per direction; null when that direction has no PML
};

// One B or D component of the curl update (step_curl, src/step_generic.cpp:65-249).
template <class R> struct CurlComp {
  R *f;
  const R *g1, *g2; // curl partners after the reference's swap (g1 never null unless both are)
  int64_t s1, s2;   // partner strides (negated for the D update, src/step_db.cpp:81-84)
  R dtdx;
  double *fu; // PML auxiliary (dsigu), FP64 whatever R is; null when absent
  // conductivity (null when none) and 1 / (1 + cnd dt / 2); fcnd is the conductivity
  // auxiliary where uPML acts on f (src/step_generic.cpp:85-248, src/step_db.cpp:66-69).
  // Chunks with conductivity run the generic kernel only.
  const R *cnd, *cndinv;
  double *fcnd; // FP64 whatever R is
  R dt2; // dt / 2 in realnum
  int dsig;  // PML direction acting on f, -1 if none
  int dsigu; // PML direction acting through fu, -1 if none
  int lo[3], hi[3]; // owned index range per direction (inclusive)
  int ysh[3];       // Yee shift per direction (PML index k = 2*j + ysh)
  // derived by finalize_args: loop axis of dsig / dsigu (-1 if none), the Yee shift along it,
  // and that direction's uPML profiles
  AxisSlot ax;
  int ksig, ksigu, ysig, ysigu;
  const R *sig, *kap, *siginv, *sigu, *kapu, *siginvu;
};

// One E or H component of update_eh (src/update_eh.cpp:66-215, src/step_generic.cpp:566-785)
// plus its Lorentzian polarisations (src/susceptibility.cpp:188-281).
template <class R> struct UpdComp {
  R *f;       // E (or H)
  const R *g; // D (or B)
  const R *u; // diagonal chi1inv; null means 1
  double *fw; // PML auxiliary W, FP64 whatever R is; null when this direction has no PML
  int dsigw;  // -1 when none
  int lo[3], hi[3], ysh[3];
  R *P[MAXPOLE], *Pp[MAXPOLE];
  const R *sigma[MAXPOLE];
  // bit p: sigma of pole p equals sigval[p] at every point of the pole's box (and is zero
  // outside it), so the vectorised kernels take the value instead of loading the array
  int sigone;
  R sigval[MAXPOLE];
  // derived by finalize_args, as for CurlComp
  AxisSlot ax;
  int ksigw, ysigw;
  const R *sigw, *kapw;
};

template <class R> struct PoleCoef {
  R gamma1inv, gamma1, omega0dtsqr, omega0dtsqr_denom;
  // loop-axis box of the points where some component's sigma is non-zero; outside it P and
  // P_prev stay exactly zero, so the kernels skip the pole there (empty when lo > hi)
  int lo[3], hi[3];
};

// Short source lists: f[idx[s]] of slot comp[s] loses the value vrow[vbase + s] of the step's
// value row (the real or imaginary part of amp * current * dt, or of amp * dipole, formed on
// the host in double), subtracted in double as the reference does (src/step.cpp:310-315,
// src/update_eh.cpp:128-138).
struct SrcTab {
  int n;
  int64_t lo, hi; // index range of the entries (cells outside it skip the table)
  int comp[SRC_TAB];
  int64_t idx[SRC_TAB];
  int vbase;
};

// Sources too long for a table, carried as dense index boxes: source s covers array indices
// lo[s][k]..hi[s][k] along loop axis k of slot comp[s], with one complex amplitude per box
// point in amp[s] (box-local, strides bst[s], zero where the source has no point).  The value
// subtracted at a point is the real (cmp 0) or imaginary (cmp 1) part of (amp * scal) * mult,
// formed in double as the reference forms sv.current(j) * dt (src/step.cpp:307-315) and
// sv.dipole(j) (src/update_eh.cpp:128-138); scal is (vrow[vbase + 2 s], vrow[vbase + 2 s + 1]).
constexpr int SRC_BOX = 2;
struct SrcBoxes {
  int n;
  int comp[SRC_BOX];
  int lo[SRC_BOX][3], hi[SRC_BOX][3];
  int64_t bst[SRC_BOX][3];
  const double2 *amp[SRC_BOX];
  double mult[SRC_BOX];
  int vbase;
};

// The sources of one launch, in device memory: the kernels read them only in the few groups
// inside the box StepArgs::sbox_lo..sbox_hi (loop-axis indices), so the tables cost the other groups
// no registers.
struct SrcSet {
  SrcTab cur;          // currents, after the curl
  SrcTab dip;          // integrated dipoles, inside f_minus_p
  SrcBoxes curb, dipb; // the same for long source lists
};

template <class R> struct StepArgs {
  Grid g;
  PmlProfile<R> pml;
  int nc;
  CurlComp<R> cc[3];
  UpdComp<R> uc[3];
  int npole;
  PoleCoef<R> pole[MAXPOLE];
  const SrcSet *src;   // null: no sources in this launch
  int sbox_lo[3], sbox_hi[3]; // loop-axis box holding every source point of src
  const double *vrow;  // this step's source values (batched launches pass theirs per step)
  int64_t vofs;        // batched launches: this chunk's offset in the step's value row
  int cmp;             // 0: real parts, 1: imaginary parts of the fields this launch moves
  // Slots are by direction: cc[d] / uc[d] is the component along d (f == nullptr when
  // absent).  For the vectorised kernel the curl partners are given by direction too: F[d]
  // is the partner field's component along d; slot d's g1 is F[(d+2)%3] differenced along
  // d+1 and its g2 is F[(d+1)%3] along d+2 (the Cartesian plan of src/fields.cpp:428-456),
  // when hasp[d] / hasm[d].  sgn is +1 for forward (B) and -1 for backward (D) differences.
  const R *F[3];
  int hasp[3], hasm[3];
  int sgn;
  R dtdx; // Courant number (the vectorised kernel applies the partner signs itself)
  int rev; // visit logical blocks in descending order (backward differences)
  // pswap: the new P of each pole goes into the P_prev array (which then holds P, while the
  // P array holds P_prev: the host swaps the two pointers after the step) instead of
  // P_prev = P; P = new P, saving one store per pole and point
  int pswap;
};

// Halo connection (connect_the_chunks / process_incoming_chunk_data,
// src/boundaries.cpp:368-600, src/step.cpp:172-223): non-owned point dst of array slot
// `slot` takes the owned point src, copied (cls 2), negated (cls 1) or multiplied by the
// Bloch phase ph[entry] (cls 0).  12 bytes per entry; phases live in a side array.
struct Conn {
  uint32_t dst, src;
  uint16_t slot;
  uint16_t cls;
};

// DFT accumulation (dft_chunk::update_dft, src/dft.cpp:266-308) over npts points.
// The per-frequency phase exp(i omega t) * scale is formed on the device (src/dft.cpp:270-271).
template <class R> struct DftArgs {
  const R *f0, *f1; // real / imaginary field arrays (f1 null for real fields)
  const int64_t *idx;
  const double *w;
  const double *omega;
  double time, scale_re, scale_im;
  int64_t avg1, avg2;
  int npts, nfreq;
  void *acc;    // accumulators, frequency fastest
  void *comp;   // Kahan compensation (float2), or null
  int acc_kind; // MP_DFT_* of meep_amd.h
};

// The flux of E x H through a plane at one instant (fields::flux_in_box_wrongH,
// src/energy_and_flux.cpp:186-221, over fields::integrate, src/integrate.cpp:46-120): the sum
// over the plane's centred-grid points of w * sign_p * Re(conj(E_p) H_p) for the component
// pairs p, each field the mean of its four Yee points around the point.  Each block writes its
// partial sum; the host adds the FLUX_BLOCKS partials.
constexpr int FLUX_BLOCKS = 128;
template <class R> struct FluxPlaneArgs {
  const int64_t *idx;
  const double *w;
  int n, npair;
  const R *e[2][2], *h[2][2]; // [pair][real, imaginary]; null where a part is absent
  int64_t eo[2][2], ho[2][2]; // [pair][the two centring offsets]
  double sign[2];
};

// Deferred DFT accumulation: the averaged, weighted field values of a chunk's points are kept
// for up to DFT_DEFER accumulation steps (snap0/snap1, step-major), then one pass applies
// every step's term to each accumulator, in step order and with the arithmetic of one step at
// a time, so the accumulators are read and written once per DFT_DEFER steps.
constexpr int DFT_DEFER = 8;
constexpr int DFT_FLUSH_TAB = 2048; // phases (steps x frequencies) a flush keeps in LDS
template <class R> struct DftDefer {
  const R *snap0, *snap1; // [step][point]; snap1 null for real fields
  int nsteps;
  double time[DFT_DEFER]; // each step's time (as the per-step accumulation would use it)
};

// ---- batched steps (many chunks in one launch) ---------------------------------------------
// One launch boxes of many independent chunks: item k is one launch of the single-chunk path
// (its argument block, thread count and row groups), placed in device memory; a block finds
// its item in blk_item and its first thread in blk_first.  Source values come from the step's
// value row (SrcTab::vbase, SrcBoxes::vbase).
template <class R> struct BatchItem {
  StepArgs<R> a;
  uint32_t count;
  int g0, ngroups; // vectorised items: first row group and groups per row
};
// DFT accumulation of one chunk in a batch; the time of step s is
// (t_start + s + 1) dt - half dt / 2, and steps where that index is not a multiple of the
// decimation are skipped.
template <class R> struct BatchDft {
  DftArgs<R> a;
  double dt;
  int half;
  int decimation;
  int64_t t_start;
  R *snap0, *snap1; // deferred accumulation (DftDefer): this chunk's point values by step slot
};
// kind: 0 generic kernel, 1 vectorised, 2 vectorised with per-point uPML, 3 and 4 the same two
// updating Lorentz/Drude poles
template <class R>
void launch_step_batch(const BatchItem<R> *items, const uint32_t *blk_item, const uint32_t *blk_first, unsigned nblk, int kind,
                       bool curl, bool upd, int block, const double *vrow, hipStream_t s);
template <class R>
void launch_dft_batch(const BatchDft<R> *items, const uint32_t *blk_item, const uint32_t *blk_first, unsigned nblk, long step,
                      hipStream_t s);
// deferred accumulation of a batch whose chunks all accumulate every step: step `step`'s point
// values into slot `slot` (block tables over points), and the steps first .. first + nsteps - 1
// into the accumulators (block tables over points x frequency quads)
template <class R>
void launch_dft_snap_batch(const BatchDft<R> *items, const uint32_t *blk_item, const uint32_t *blk_first, unsigned nblk, int slot,
                           hipStream_t s);
template <class R>
void launch_dft_flush_batch(const BatchDft<R> *items, const uint32_t *blk_item, const uint32_t *blk_first, unsigned nblk, long first,
                            int nsteps, hipStream_t s);
// halo lists of many chunks: item k is (conns, phases, n, slots, ncmp)
struct BatchConn {
  const Conn *conns;
  const double2 *phases;
  int n;
  void *const *slots;
  int ncmp;
};
template <class R>
void launch_conn_batch(const BatchConn *items, const uint32_t *blk_item, const uint32_t *blk_first, unsigned nblk, hipStream_t s);
constexpr int BATCH_BLOCKS[4] = {64, 128, 256, 512}; // block sizes the batch autotuner tries

// ---- reverse-mode step (adjoint.hip) -------------------------------------------------------
// The adjoint of one step applies the transpose of each forward update, in reverse order, to
// cotangent arrays ("bar" arrays, one per forward state array, aliased as the forward arrays
// are).  Slots mirror the forward StepArgs slots.
template <class R> struct AdjSlots {
  R *fb[3];   // cotangent of the slot's field (B or D for a curl slot, E or H for an update slot)
  R *fub[3];  // curl: cotangent of f_u
  R *gb[3];   // update: cotangent of the source field (D or B)
  R *fwb[3];  // update: cotangent of f_w
  R *cbar[3]; // curl: cotangent of the slot's curl term, zero off the owned points
};
// design box (loop-axis indices) of one update slot: the gradient of chi1inv there collects
// (cotangent of u * D) * D over every step, with D read from the forward history
template <class R> struct AdjDesign {
  int slot; // -1: none
  int lo[3], hi[3];
  const R *hist; // D at the box's points for this step, box-local (loop order)
  double *grad;  // box-local accumulator
};
// Gather of curl-term cotangents into one partner array: gb[j] += sum over entries e of
// cbar_e[j - s_e] (when that point exists along axis ax_e) - cbar_e[j] for a "plus" entry,
// and cbar_e[j] - cbar_e[j - s_e] for a "minus" entry.
template <class R> struct AdjGather {
  int n;
  R *gb;
  const R *cbar[4];
  int64_t s[4];
  int ax[4];   // loop axis of s
  int step[4]; // +1 or -1: coordinate step of s along ax
  int plus[4]; // 1: g1 (the i + s1 partner), 0: g2
};

// launchers (adjoint.hip)
template <class R> void launch_adj_upd(const StepArgs<R> &a, const AdjSlots<R> &b, const AdjDesign<R> &dz, hipStream_t s);
template <class R> void launch_adj_curl(const StepArgs<R> &a, const AdjSlots<R> &b, hipStream_t s);
template <class R> void launch_adj_gather(const Grid &g, const AdjGather<R> &ga, hipStream_t s);
// halo transpose: for each owned source point src[q], its halo images dst[first[q] .. first[q+1])
// with signs sgn: bar[src] += sgn * bar[dst]; bar[dst] = 0 (slot pointers in bars)
template <class R>
void launch_adj_halo(const int *first, const uint32_t *src, const uint32_t *dst, const int8_t *sgn, const uint16_t *slot, int nsrc,
                     R *const *bars, hipStream_t s);
// DFT objective J = sum |a|^2: G[k] = w_k * sum_f 2 Re(conj(a_f(k)) ph_f) with ph_f =
// exp(i omega_f t) * scale, then bar[idx_k + off] += coef * G[k]
template <class R>
void launch_adj_dft(const int64_t *idx, const double *w, const double2 *acc, const double *omega, int npts, int nfreq, double time,
                    double scale_re, double scale_im, double *G, hipStream_t s);
template <class R> void launch_adj_dft_scatter(R *bar, const int64_t *idx, const double *G, int npts, int64_t off, double coef, hipStream_t s);
// copy D at the design box points (loop-axis box, device layout) into hist (box-local)
template <class R> void launch_record(R *hist, const R *D, const Grid &g, const int *lo, const int *hi, hipStream_t s);

// launchers (kernels.hip)
// generic scalar kernel over the box a.g.blo..bhi (any cells, uPML included)
template <class R> void launch_step(const StepArgs<R> &a, bool curl, bool upd, hipStream_t s);
// vectorised kernel (16-byte accesses); pml: the uPML branches per point (else the box must
// have no active uPML profile)
template <class R> void launch_fast(const StepArgs<R> &a, bool curl, bool upd, bool pml, hipStream_t s);
// the 3D row kernel for a box with no active x or y uPML profile and no folded images; false
// when the box or its slots do not qualify (the caller then uses launch_fast).
// MEEP_AMD_NO_ROW3 (read once) turns it off.
template <class R> bool launch_row3(const StepArgs<R> &a, bool curl, bool upd, hipStream_t s);
template <class R>
void launch_conn(const Conn *conns, const double2 *phases, int n, R *const *slot_ptrs, int ncmp, hipStream_t s);
// f[idx[j]] -= Re or Im of (amp[j] * scal) * mult for the n points of one source
template <class R>
void launch_src(R *f, const int64_t *idx, const double2 *amp, int n, double2 scal, double mult, int cmp,
                hipStream_t s);
// f[i] = (f[i] + b[i]) / 2 over n points (fields_chunk::average_with_backup)
template <class R> void launch_average(R *f, const R *b, size_t n, hipStream_t s);
// dst[i] = src[i] over n points, realnum to FP64 (an auxiliary array's first values)
template <class R> void launch_widen(double *dst, const R *src, size_t n, hipStream_t s);
// f[dst[q]] = sgn[q] * f[idx[k[q]]]: the folded periodic images of source points
template <class R>
void launch_src_images(R *f, const int64_t *idx, const int *k, const int64_t *dst, const int8_t *sgn, int n, hipStream_t s);
// dst[j] = f[idx[j]] and f[idx[j]] = src[j]
template <class R> void launch_gather(R *dst, const R *f, const int64_t *idx, int n, hipStream_t s);
template <class R> void launch_scatter(R *f, const int64_t *idx, const R *src, int n, hipStream_t s);
template <class R> void launch_dft(const DftArgs<R> &a, hipStream_t s);
// the values of a's points at this step into snap0 / snap1 (one step's slot)
template <class R> void launch_dft_snap(const DftArgs<R> &a, R *snap0, R *snap1, hipStream_t s);
// the deferred steps of d into a's accumulators
template <class R> void launch_dft_flush(const DftArgs<R> &a, const DftDefer<R> &d, hipStream_t s);
template <class R>
void launch_check(const R *const *ptrs, const int64_t *idx, const double *w, int n, int *flag,
                  hipStream_t s);
// FLUX_BLOCKS partial sums of the plane flux into partial
template <class R> void launch_flux_plane(const FluxPlaneArgs<R> &a, double *partial, hipStream_t s);

// The near-to-far field sums (near2far.hip): nsrc equivalent currents at (sx, sy, sz) with
// amplitudes f0[frequency][current], sorted by class (Ex Ey Ez Hx Hy Hz: class c holds
// [cls_lo[c], cls_lo[c + 1])), seen at npts far points (px, py, pz), in the medium eps, mu of
// dimension dim (2 or 3).  Each (point, frequency group) is split into segs segments of the
// current list.
struct FarArgs {
  const double *sx, *sy, *sz;
  int cls_lo[7];
  const double2 *f0;
  int nsrc;
  const double *px, *py, *pz;
  int npts;
  const double *freq;
  int nfreq;
  double eps, mu;
  int dim, segs;
  int uniform; // freq[i] = freq[0] + i (freq[1] - freq[0]) to rounding
};
// frequencies a block of the far-field kernel takes together in 3D: 2 and 3 measured alike
// (4 runs out of registers and halves the rate); 2 divides the usual even frequency counts
constexpr int FF_GROUP3 = 2;
inline int farfield_group(int dim) { return dim == 3 ? FF_GROUP3 : 1; }
// partial: npts * nfreq * segs * 12 doubles of scratch; out: npts * nfreq * 12 doubles, the
// (re, im) of Ex Ey Ez Hx Hy Hz per (point, frequency), point-major
void launch_farfield(const FarArgs &a, double *partial, double *out, hipStream_t s);

} // namespace mx
