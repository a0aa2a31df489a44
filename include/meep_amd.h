/* meep_amd.h — C-ABI of the MI355X FDTD engine, the drop-in for Meep's time-step path.
 *
 * Every entry point replaces one method of the reference C++ API (cited as file:line in
 * NanoComp/meep).  The reference exposes that API to Python through SWIG (python/meep.i);
 * a ctypes / SWIG binding of the functions below is the same shape.  The C++ classes in
 * meep_amd.hpp (namespace meep) are a thin layer over these functions, so a C++ caller of
 * meep::fields::step() links against this library unchanged.
 *
 * Conventions
 *  - Enumerations carry the reference's numeric values (src/meep/vec.hpp:28-79,
 *    src/meep.hpp:1594), so component/direction integers pass through unchanged.
 *  - Field arrays are exchanged in the reference layout: one array of ntot = prod(n_d + 1)
 *    values per component and real/imaginary part, Z fastest in 3D, Y fastest in 2D
 *    (src/vec.cpp:262-296,482-494).  Values cross the ABI as double whatever the engine's
 *    arithmetic type.
 *  - Functions returning int return 0 on success and a negative code on error; the message
 *    is in mp_last_error().  The C++ layer turns a failure into meep::abort (a
 *    std::runtime_error "meep: ..."), matching src/mympi.cpp:238-256.
 *  - A handle is not thread-safe; one host thread drives one fields object (as in the
 *    reference, src/meep.hpp:1621-1622).
 */
#ifndef MEEP_AMD_H
#define MEEP_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* component (src/meep/vec.hpp:31-54) */
enum {
  MP_Ex = 0, MP_Ey = 1, MP_Ez = 4, MP_Hx = 5, MP_Hy = 6, MP_Hz = 9,
  MP_Dx = 10, MP_Dy = 11, MP_Dz = 14, MP_Bx = 15, MP_By = 16, MP_Bz = 19
};
/* direction (src/meep/vec.hpp:78) and boundary_side (:77) */
enum { MP_X = 0, MP_Y = 1, MP_Z = 2 };
enum { MP_High = 0, MP_Low = 1 };
/* boundary_condition (src/meep.hpp:1594) */
enum { MP_Periodic = 0, MP_Metallic = 1, MP_Magnetic = 2, MP_None = 3 };
/* state arrays reachable through mp_fields_get_array / mp_fields_set_array */
enum { MP_ARR_F = 0, MP_ARR_F_U = 1, MP_ARR_F_W = 2, MP_ARR_P = 3, MP_ARR_P_PREV = 4, MP_ARR_F_COND = 5 };
/* DFT accumulator storage: the reference's complex<realnum> (src/meep.hpp:1138), FP64, or
   compensated (Kahan) complex<float>. */
enum { MP_DFT_REALNUM = 0, MP_DFT_F64 = 1, MP_DFT_KAHAN32 = 2 };

typedef struct mp_structure mp_structure;
typedef struct mp_fields mp_fields;
typedef struct mp_dft_flux mp_dft_flux;
typedef struct mp_dft_fields mp_dft_fields;
typedef struct mp_flux_vol mp_flux_vol;
typedef struct mp_near2far mp_near2far;
typedef struct mp_dft_energy mp_dft_energy;
typedef struct mp_dft_force mp_dft_force;
typedef struct mp_adjoint mp_adjoint;
typedef struct mp_batch mp_batch;

/* r points at (x, y, z) in the reference's coordinates (origin at the cell's low corner). */
typedef double (*mp_scalar_func)(const double *r, void *user);
typedef void (*mp_complex_func)(const double *r, void *user, double *re, double *im);

/* Time dependence of a source: gaussian_src_time(freq, width, start, end)
   (src/sources.cpp:85-96) when kind == 0, continuous_src_time(freq, width, start, end,
   slowness) (src/meep.hpp:1040-1044) when kind == 1.  is_integrated as src_time
   (src/meep.hpp:944-957). */
typedef struct {
  int kind;
  double freq;
  double width;
  double start_time;
  double end_time;
  double slowness;
  int is_integrated;
} mp_src_time;

/* Message of the last failed call on this thread. */
const char *mp_last_error(void);

/* Name of the engine behind this library ("hip-gfx950" for the product). */
const char *mp_engine_name(void);

/* ---- structure -------------------------------------------------------------------- */

/* structure(grid_volume, double eps(const vec&), boundary_region, identity(), 0, Courant,
   use_anisotropic_averaging = false)  — src/meep.hpp:835-839, src/structure.cpp:48-63.
   dim is 1, 2 or 3 (vol1d/vol2d/vol3d, src/vec.cpp:904-930; sizes of absent directions are
   ignored).  eps may be NULL (vacuum).  The boundary region is the sum of n_pml
   pml(thickness, direction, side, R) terms (src/structure.cpp:252-273).  single_precision
   selects float as realnum (the reference's --enable-single, src/meep.hpp:43-46). */
mp_structure *mp_structure_new(int dim, double xsize, double ysize, double zsize, double a,
                               mp_scalar_func eps, void *eps_user, int n_pml,
                               const int *pml_direction, const int *pml_side,
                               const double *pml_thickness, const double *pml_R,
                               double courant, int single_precision);
/* mp_structure_new with every parameter of a uPML term, boundary_region(PML, thickness,
   Rasymptotic, mean_stretch, pml_profile, pml_profile_data, pml_profile_integral,
   pml_profile_integral_u, d, side) — src/meep.hpp:665-672,732-736, structure_chunk::use_pml
   src/structure.cpp:630-691: sigma = -ln(R) / (4 d integral) * profile(x/d) and kappa =
   1 + (mean_stretch - 1) / integral_u * profile(x/d) * x/d at depth x into the term's layer.
   The profile is called with r[0] = u in [0, 1]; integral and integral_u are its integrals of
   profile(u) and u profile(u) over [0, 1] (the reference's Python PML passes 1/3 and 1/4
   whatever the profile, python/simulation.py:4985-5000).  Each array may be NULL, and an
   entry of profile NULL or of an integral <= 0, for the defaults: mean stretch 1, profile u^2,
   integrals 1/3 and 1/4. */
mp_structure *mp_structure_new_pml(int dim, double xsize, double ysize, double zsize, double a,
                                   mp_scalar_func eps, void *eps_user, int n_pml,
                                   const int *pml_direction, const int *pml_side,
                                   const double *pml_thickness, const double *pml_R,
                                   const double *pml_mean_stretch,
                                   const mp_scalar_func *pml_profile,
                                   void *const *pml_profile_user, const double *pml_integral,
                                   const double *pml_integral_u, double courant,
                                   int single_precision);
void mp_structure_delete(mp_structure *s);
/* grid_volume::ntot (src/meep/vec.hpp:1056) and num_direction (src/meep/vec.hpp:1028). */
size_t mp_structure_ntot(const mp_structure *s);
int mp_structure_num(const mp_structure *s, int *n_xyz);
double mp_structure_dt(const mp_structure *s);
/* Overwrite the diagonal chi1inv[c][dir(c)] of electric component c on all ntot points
   (structure_chunk::chi1inv, src/meep.hpp:597; filled by set_chi1inv,
   src/anisotropic_averaging.cpp:211-328). */
int mp_structure_set_chi1inv(mp_structure *s, int c, const double *chi1inv);
int mp_structure_get_chi1inv(const mp_structure *s, int c, double *chi1inv);
/* structure::add_susceptibility(sigma, E_stuff, lorentzian_susceptibility(omega_0, gamma,
   no_omega_0_denominator)) — src/meep.hpp:248,864-865, src/structure.cpp:471-507.
   Returns the susceptibility's index (0, 1, ...) or a negative code. */
int mp_structure_add_susceptibility(mp_structure *s, double omega_0, double gamma,
                                    int no_omega_0_denominator, mp_scalar_func sigma,
                                    void *sigma_user);
/* structure::add_susceptibility(double sigma(const vec &), field_type ft, const
   susceptibility &) — src/meep.hpp:864, src/structure.cpp:471-507 — for ft E_stuff (0) or
   H_stuff (1, a magnetic Lorentz/Drude pole: sigma on the H components, P subtracted from B
   and updated from H, src/step.cpp:76-88); mp_structure_add_susceptibility is ft = E_stuff. */
int mp_structure_add_susceptibility_ft(mp_structure *s, int ft, double omega_0, double gamma, int drude,
                                       mp_scalar_func sigma, void *user);
/* Overwrite sigma[c][dir(c)] of susceptibility isus on all ntot points. */
int mp_structure_set_sigma(mp_structure *s, int isus, int c, const double *sigma);

/* structure::set_conductivity(component c, double C(const vec &)) — src/meep.hpp:855-856,
   src/structure.cpp:425-437,868-905.  c is an E, H, D or B component; the conductivity is
   stored for the D (B) component along c's direction, sampled at that component's points,
   and for an E (H) component multiplied by c's chi1inv as it stands at this call (so set
   epsilon first, as in the reference).  A zero function removes it.  Applies to fields
   objects created afterwards. */
int mp_structure_set_conductivity(mp_structure *s, int c, mp_scalar_func C, void *user);

/* ---- fields ----------------------------------------------------------------------- */

/* fields::fields(structure*, m = 0, beta = 0) — src/meep.hpp:1740-1742, src/fields.cpp:32-86.
   Fields are complex until mp_fields_use_real_fields. */
mp_fields *mp_fields_new(mp_structure *s);
void mp_fields_delete(mp_fields *f);
/* fields::use_real_fields — src/meep.hpp:1746, src/fields.cpp:144-156 */
int mp_fields_use_real_fields(mp_fields *f);
/* fields::use_bloch(const vec &k) — src/meep.hpp:1763, src/boundaries.cpp:113-119.
   k holds one real entry per grid direction in X, Y, Z order. */
int mp_fields_use_bloch(mp_fields *f, const double *k);
/* fields::use_bloch(direction, complex k) — src/meep.hpp:1762, src/boundaries.cpp:90-111 */
int mp_fields_use_bloch_dir(mp_fields *f, int d, double k_re, double k_im);
/* fields::set_boundary — src/meep.hpp:1761, src/boundaries.cpp:78-88 */
int mp_fields_set_boundary(mp_fields *f, int side, int d, int condition);
/* Deprecated fields::add_point_source(c, freq, width, peaktime, cutoff, p, amp, is_c)
   — src/meep.hpp:1888-1889, src/sources.cpp:193-213. */
int mp_fields_add_point_source_legacy(mp_fields *f, int c, double freq, double width,
                                      double peaktime, double cutoff, const double *p,
                                      double amp_re, double amp_im, int is_continuous);
/* fields::add_volume_source(c, src, volume(vmin, vmax), A, amp)
   — src/meep.hpp:1893-1897, src/sources.cpp:462-490.  A may be NULL (A = 1). */
int mp_fields_add_volume_source(mp_fields *f, int c, const mp_src_time *src,
                                const double *vmin, const double *vmax, double amp_re,
                                double amp_im, mp_complex_func A, void *A_user);
/* fields::add_volume_source(c, src, where, complex<double> *arr, dim1, dim2, dim3, amp) —
   src/meep.hpp:1895-1896, src/sources.cpp:394-417: the amplitude at a point is the array
   interpolated at the point's place in the volume (amp_file_func, src/sources.cpp:345-376:
   each direction the volume spans mapped to [0, 1]; linear_interpolate with mirror
   boundaries, src/fields.cpp:754-810).  arr holds dim1 x dim2 x dim3 complex values, row
   major, interleaved (re, im). */
int mp_fields_add_volume_source_array(mp_fields *f, int c, const mp_src_time *src,
                                      const double *vmin, const double *vmax, const double *arr,
                                      size_t dim1, size_t dim2, size_t dim3, double amp_re,
                                      double amp_im);
/* fields::add_srcdata(sourcedata{c, index, chunk 0, amp}, src, n, amp, needs_boundary_fix) —
   src/meep.hpp:1907-1908, src/sources.cpp:314-334: a source on n points of the E or H
   component c, given by their indices in c's array (reference layout) and complex
   amplitudes (interleaved re, im) applied as the reference applies a source's (times the
   current and dt, or the dipole).  Points the chunk does not own wait for
   mp_fields_require_source_components when needs_boundary_fix, and are dropped otherwise
   (the halo exchange would overwrite them). */
int mp_fields_add_srcdata(mp_fields *f, int c, const mp_src_time *src, long n, const long *index,
                          const double *amp, int needs_boundary_fix);
/* fields::require_source_components — src/fields.cpp:508-530, with fix_boundary_sources
   (src/fix_boundary_sources.cpp:51-197): every marked point the chunk does not own moves to
   the point that owns it (its periodic image, the amplitude times the conjugate Bloch phase)
   unless that point is on a metal boundary; the source components are allocated. */
int mp_fields_require_source_components(mp_fields *f);
/* fields::step() — src/meep.hpp:1871, src/step.cpp:35-139.  Returns an error (the
   reference aborts) when the fields are not finite. */
int mp_fields_step(mp_fields *f);
/* n calls of fields::step() issued back to back with no host synchronisation between
   them; the finiteness check runs on the device and is read once at the end. */
int mp_fields_step_n(mp_fields *f, long n);
/* n steps as mp_fields_step_n with no device memory allocated or freed during the call.  No
   counterpart in the reference: it is for callers that cannot allocate (code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc).  The per-step source values
   are staged in a library-owned device array of 1 MiB per GPU, held until the process ends;
   a call stages as many steps as fit at a time, and fails if one step's values (8 bytes per
   source table entry or box component) exceed it.  Every call on one GPU shares that array:
   calls that overlap on one GPU race.  The fields must have been stepped by mp_fields_step_n
   since the last change of sources, monitors or boundaries, so that their arrays and launch
   plan exist; otherwise, or with kernel timing on, the call fails before its first step.  DFT
   monitors accumulate every step (the deferred accumulation of mp_fields_step_n keeps a
   buffer of its own); results equal mp_fields_step_n's. */
int mp_fields_step_n_noalloc(mp_fields *f, long n);
/* fields::time(), round_time() — src/meep.hpp:1876-1877; fields::t and dt. */
double mp_fields_time(const mp_fields *f);
double mp_fields_round_time(const mp_fields *f);
long mp_fields_t(const mp_fields *f);
double mp_fields_dt(const mp_fields *f);
/* fields::last_source_time — src/meep.hpp:1887 */
double mp_fields_last_source_time(mp_fields *f);
/* fields::field_energy() — src/energy_and_flux.cpp:48,54-62: the electric energy plus the
   magnetic energy of B and H advanced half a step and averaged to time t
   (synchronize_magnetic_fields, :146-178); every magnetic array is restored afterwards. */
int mp_fields_field_energy(mp_fields *f, double *energy);
/* fields::get_field(component, const vec&) — src/meep.hpp:2183-2185, src/monitor.cpp:93-108 */
int mp_fields_get_field(mp_fields *f, int c, const double *p, double *re, double *im);
/* Copy one state array (reference layout, ntot doubles) to the host: kind MP_ARR_F is
   fields_chunk::f, MP_ARR_F_U f_u, MP_ARR_F_W f_w, MP_ARR_F_COND f_cond (the conductivity
   auxiliary where uPML acts on f, src/step_db.cpp:66-69), MP_ARR_P / MP_ARR_P_PREV the
   polarisation of susceptibility ipol (lorentzian_data, src/susceptibility.cpp:98-139).
   Returns 1 (and zeros) when the array is not allocated. */
int mp_fields_get_array(mp_fields *f, int c, int cmp, int kind, int ipol, double *out);
/* Overwrite one state array (checkpoint restore, adjoint seeding). */
int mp_fields_set_array(mp_fields *f, int c, int cmp, int kind, int ipol, const double *in);
/* Whether component c is allocated (fields::have_component, src/fields.cpp:158-162). */
int mp_fields_have_component(mp_fields *f, int c);
/* Per-phase wall time in seconds under the reference's sink names (time_sink,
   src/meep.hpp:1595-1618): 0 FieldUpdateB, 1 FieldUpdateH, 2 FieldUpdateD, 3 FieldUpdateE,
   4 BoundarySteppingB..E (sum), 5 FourierTransforming, 6 Stepping (whole steps). */
int mp_fields_get_times(mp_fields *f, double *times7);
/* Kernel timing with HIP events recorded on the engine's stream around every launch (no
   counterpart in the reference; feeds the time sinks above and bench.py's roofline).
   which: MP_K_STEP_BH (curl B + H update: every launch of the phase), MP_K_STEP_DE (curl D +
   E/P update: every launch of the phase), MP_K_HALO, MP_K_DFT, MP_K_SRC, and the interior-box
   launch alone of each phase, MP_K_BH_INTERIOR and MP_K_DE_INTERIOR, and the near-to-far field
   sums, MP_K_FARFIELD.  Times accumulate in
   milliseconds until mp_fields_reset_kernel_timing. */
enum { MP_K_STEP_BH = 0, MP_K_STEP_DE = 1, MP_K_HALO = 2, MP_K_DFT = 3, MP_K_SRC = 4, MP_K_BH_INTERIOR = 5,
       MP_K_DE_INTERIOR = 6, MP_K_FARFIELD = 7 };
int mp_fields_set_kernel_timing(mp_fields *f, int on);
int mp_fields_get_kernel_timing(mp_fields *f, int which, double *total_ms, long *launches);
int mp_fields_reset_kernel_timing(mp_fields *f);
/* DFT accumulator storage for monitors added later (MP_DFT_*). */
int mp_fields_set_dft_accumulator(mp_fields *f, int kind);

/* ---- DFT flux ----------------------------------------------------------------------- */

/* fields::add_dft_flux(direction, volume, const double *freq, Nfreq, use_symmetry,
   centered_grid, decimation_factor) — src/meep.hpp:2056-2058, src/dft.cpp:602-650,835-842.
   With the identity symmetry use_symmetry changes nothing.  centered_grid 0 accumulates each
   component on its own Yee grid without averaging (src/dft.cpp:91-94,233); flux() then needs
   the E and H chunks to hold the same number of points.  decimation_factor 0 picks the
   reference's automatic factor (src/dft.cpp:196-217). */
mp_dft_flux *mp_fields_add_dft_flux(mp_fields *f, int d, const double *vmin,
                                    const double *vmax, const double *freq, int nfreq,
                                    int use_symmetry, int centered_grid, int decimation_factor);
void mp_dft_flux_delete(mp_dft_flux *fl);
/* dft_flux::flux() — src/dft.cpp:542-556.  out has nfreq entries. */
int mp_dft_flux_flux(mp_dft_flux *fl, double *out);
/* Number of dft_chunks in the E (which = 0) or H (which = 1) list and the points in chunk i. */
int mp_dft_flux_num_chunks(mp_dft_flux *fl, int which);
long mp_dft_flux_chunk_points(mp_dft_flux *fl, int which, int i);
int mp_dft_flux_decimation(mp_dft_flux *fl);
/* dft_chunk::dft of chunk i, interleaved (re, im), frequency fastest (src/meep.hpp:1138). */
int mp_dft_flux_get_dft(mp_dft_flux *fl, int which, int i, double *out);
/* fields::get_dft_array(dft_flux, component c, int num_freq, int *rank, size_t dims[3]) —
   src/meep.hpp:2121-2122, src/dft.cpp:1259-1268 (process_dft_component :931-1057,1130-1256,
   collapse_array src/array_slice.cpp:554-596).  The DFT of component c at frequency index
   num_freq on the points of the monitor, each accumulator divided by its chunk's stored weight
   and integration weight, summed over the directions in which the monitor volume is empty
   with the interpolation weights along them.  Writes rank and dims (array dimensions in
   X, Y, Z order of the non-degenerate directions, last fastest); out, if not NULL, receives
   prod(dims) complex values, interleaved (re, im), rounded to realnum as the reference's
   complex<realnum> array is.  Returns the number of complex values (0 where the reference
   returns a null array: no chunk of c, or a monitor on a single grid point), -1 on error
   (frequency index out of range). */
long mp_dft_flux_get_dft_array(mp_dft_flux *fl, int c, int num_freq, int *rank, size_t *dims,
                               double *out);

/* ---- DFT fields ---------------------------------------------------------------------- */

/* fields::add_dft_fields(components, num_components, volume, const double *freq, Nfreq,
   use_centered_grid, decimation_factor) — src/meep.hpp:2085-2090, src/dft.cpp:913-927:
   per component, accumulators without integration weights, on the centred grid (averaged
   from the Yee grid) or, use_centered_grid 0, on the component's own Yee grid.  Chunks are
   listed in the reference's order (each component's chunks prepended). */
mp_dft_fields *mp_fields_add_dft_fields(mp_fields *f, const int *components, int num_components,
                                        const double *vmin, const double *vmax,
                                        const double *freq, int nfreq, int use_centered_grid,
                                        int decimation_factor);
void mp_dft_fields_delete(mp_dft_fields *df);
int mp_dft_fields_num_chunks(mp_dft_fields *df);
long mp_dft_fields_chunk_points(mp_dft_fields *df, int i);
/* the field component accumulated by chunk i (dft_chunk::c) */
int mp_dft_fields_chunk_component(mp_dft_fields *df, int i);
/* dft_chunk::dft of chunk i, interleaved (re, im), frequency fastest. */
int mp_dft_fields_get_dft(mp_dft_fields *df, int i, double *out);
/* fields::get_dft_array(dft_fields, component, num_freq, rank, dims) — src/meep.hpp:2123-2124,
   src/dft.cpp:1281-1290; as mp_dft_flux_get_dft_array. */
long mp_dft_fields_get_dft_array(mp_dft_fields *df, int c, int num_freq, int *rank, size_t *dims,
                                 double *out);

/* Replace chunk i's accumulators with in (interleaved (re, im), frequency fastest, as
   mp_dft_flux_get_dft gives them): the per-chunk load of the Python layer's load_flux_data
   (python/simulation.py:3638-3647, mp._load_dft_data over flux.E and flux.H). */
int mp_dft_flux_set_dft(mp_dft_flux *fl, int which, int i, const double *in);
/* dft_flux::scale_dfts(complex<double>) — src/dft.cpp:597-600, dft_chunk::scale_dft
   :402-406 (load_minus_flux_data scales by -1, python/simulation.py:3651-3659). */
int mp_dft_flux_scale_dfts(mp_dft_flux *fl, double re, double im);

/* fields::add_dft_energy(const volume_list *, freq, Nfreq, decimation_factor) —
   src/dft.cpp:725-750: volume w is [vmins + 3w, vmaxs + 3w]; E and H accumulate with
   integration weights, D and B without, on the centred grid. */
mp_dft_energy *mp_fields_add_dft_energy(mp_fields *f, int nvol, const double *vmins, const double *vmaxs,
                                        const double *freq, int nfreq, int decimation_factor);
void mp_dft_energy_delete(mp_dft_energy *e);
/* dft_energy::total (which 0), electric (1), magnetic (2) — src/dft.cpp:681-723: out[Nfreq]. */
int mp_dft_energy_get(mp_dft_energy *e, int which, double *out);

/* fields::add_dft_force(const volume_list *, freq, Nfreq, decimation_factor) —
   src/stress.cpp:153-198: volume w is [vmins + 3w, vmaxs + 3w] with complex weight
   (weights[2w], weights[2w + 1]); comps[w] names the force direction (an E or D component, or
   Sx = 100, Sy = 101, Sz = 104); the normal is the volume's.  Maxwell stress-tensor terms:
   off-diagonal E/H pairs on the centred grid, diagonal ones with square-root weights on the
   components' own grids. */
mp_dft_force *mp_fields_add_dft_force(mp_fields *f, int nvol, const int *comps, const double *vmins,
                                      const double *vmaxs, const double *weights, const double *freq,
                                      int nfreq, int decimation_factor);
void mp_dft_force_delete(mp_dft_force *fo);
/* dft_force::force() — src/stress.cpp:101-114: out[Nfreq]. */
int mp_dft_force_force(mp_dft_force *fo, double *out);
/* fields::get_dft_array(dft_force, component, num_freq, rank, dims) — src/meep.hpp:2122,
   src/dft.cpp:1270-1280; as mp_dft_flux_get_dft_array over offdiag1, offdiag2 and diag. */
long mp_dft_force_get_dft_array(mp_dft_force *fo, int c, int num_freq, int *rank, size_t *dims, double *out);
/* dft_force::scale_dfts — src/stress.cpp:144-148; dft_energy::scale_dfts — src/dft.cpp:785-790
   (load_minus_force_data / load_minus_energy_data scale by -1). */
int mp_dft_force_scale_dfts(mp_dft_force *fo, double re, double im);
int mp_dft_energy_scale_dfts(mp_dft_energy *e, double re, double im);

/* ---- time-domain flux ----------------------------------------------------------------- */

/* fields::add_flux_vol(direction, const volume &) — src/meep.hpp:2205-2207,
   src/energy_and_flux.cpp:230-235; d < 0 is fields::add_flux_plane(volume) (:238-241), whose
   direction is the volume's normal (volume::normal_direction, src/vec.cpp:227-257).  The
   flux of E x H through the volume, taken by every step after its H update and after its E
   update (src/step.cpp:93,124; flux_vol::update_half / update, src/meep.hpp:2330-2337).  The
   plane stays with the fields as the reference's list does; mp_flux_vol_delete releases the
   handle. */
mp_flux_vol *mp_fields_add_flux_vol(mp_fields *f, int d, const double *vmin, const double *vmax);
void mp_flux_vol_delete(mp_flux_vol *fv);
/* flux_vol::flux() — src/meep.hpp:2339: the mean of the last step's two fluxes. */
int mp_flux_vol_flux(mp_flux_vol *fv, double *out);
/* fields::flux_in_box(direction, const volume &) — src/energy_and_flux.cpp:223-228: E x H with
   H synchronised to E's time (synchronize_magnetic_fields, :146-164); d < 0 as above. */
int mp_fields_flux_in_box(mp_fields *f, int d, const double *vmin, const double *vmax, double *out);
/* fields::flux_in_box_wrongH — src/energy_and_flux.cpp:186-221: E x H as they stand. */
int mp_fields_flux_in_box_wrongH(mp_fields *f, int d, const double *vmin, const double *vmax,
                                 double *out);
/* fields::field_energy_in_box(const volume &) (which 0: the magnetic part with H synchronised,
   src/energy_and_flux.cpp:54-59), electric_energy_in_box (1) and magnetic_energy_in_box (2,
   H as it stands) (:84-95), over [vmin, vmax]. */
int mp_fields_energy_in_box(mp_fields *f, int which, const double *vmin, const double *vmax,
                            double *out);

/* ---- near-to-far transformation --------------------------------------------------------- */

/* fields::add_dft_near2far(const volume_list *, const double *freq, size_t Nfreq,
   int decimation_factor, int Nperiods) — src/meep.hpp:2161-2168, src/near2far.cpp:569-654.
   Surface w is [vmins + 3w, vmaxs + 3w] with the complex weight (weights[2w], weights[2w + 1])
   (volume_list::weight) and normal dirs[w] (X, Y, Z, or
   -1 for the volume's normal direction; the reference's volume_list::c is the flux component
   Sx, Sy or Sz of that direction).  Its tangential E and H are accumulated off the centred grid
   with integration weights as the equivalent currents of the other kind.  The medium is
   eps, mu at each surface's centre and must agree between surfaces.  dec <= 0 picks the
   decimation as add_dft_flux does. */
mp_near2far *mp_fields_add_dft_near2far(mp_fields *f, int nsurf, const int *dirs, const double *vmins,
                                        const double *vmaxs, const double *weights, const double *freq,
                                        int nfreq, int dec, int nperiods);
void mp_near2far_delete(mp_near2far *n);
/* dft_near2far::eps, mu (src/meep.hpp:1468) into eps_mu[2], periodic_d / periodic_n of the
   two tangential directions (-1 when not periodic) into periodic[4], and the number of
   accumulator points over its chunks into *npoints; returns Nfreq. */
int mp_near2far_info(const mp_near2far *n, double *eps_mu, int *periodic, long *npoints);
/* dft_near2far::farfield(const vec &x) — src/near2far.cpp:399-408: EH[12 * Nfreq] as
   (re, im) of Ex, Ey, Ez, Hx, Hy, Hz for each frequency in turn. */
int mp_near2far_farfield(mp_near2far *n, const double *x, double *EH);
/* dft_near2far::get_farfields_array(const volume &, int &rank, size_t *dims, size_t &N,
   double resolution) — src/near2far.cpp:411-483: a 6 x 2 x N x Nfreq row-major array
   (component, real/imaginary part, point, frequency) on floor(extent * resolution) points per
   direction.  Returns N; out == NULL returns N, rank and dims only. */
long mp_near2far_get_farfields_array(mp_near2far *n, const double *wmin, const double *wmax,
                                     double resolution, int *rank, size_t *dims, double *out);
/* dft_near2far::flux(direction, const volume &, double resolution) — src/near2far.cpp:517-565. */
int mp_near2far_flux(mp_near2far *n, int d, const double *wmin, const double *wmax, double resolution,
                     double *out);
/* fields::get_dft_array(dft_near2far, component, num_freq, rank, dims) — src/dft.cpp:1282-1290;
   as mp_dft_flux_get_dft_array, collapsed over the first surface's empty directions. */
long mp_near2far_get_dft_array(mp_near2far *n, int c, int num_freq, int *rank, size_t *dims,
                               double *out);
/* The near2far monitor's chunk list (dft_near2far::F) as mp_dft_flux_* give a flux's E list:
   get_near2far_data / load_near2far_data (python/simulation.py:3355-3371), and
   dft_near2far::scale_dfts (src/near2far.cpp:124-126; load_minus_near2far_data scales by -1). */
int mp_near2far_num_chunks(mp_near2far *n);
long mp_near2far_chunk_points(mp_near2far *n, int i);
int mp_near2far_get_dft(mp_near2far *n, int i, double *out);
int mp_near2far_set_dft(mp_near2far *n, int i, const double *in);
int mp_near2far_scale_dfts(mp_near2far *n, double re, double im);
/* green2d / green3d (src/near2far.cpp:133-186, 207-267): EH[12] of a point current of
   component c0 and amplitude f0[2] at x0, seen at x, in the medium eps, mu; dim is 2 or 3. */
int mp_near2far_green(int dim, const double *x, double freq, double eps, double mu, const double *x0,
                      int c0, const double *f0, double *EH);

/* ---- reverse mode (gradient of a DFT objective) ------------------------------------- */

/* Differentiation of J = sum over the accumulators a of `objective` of |a|^2 with respect to
   chi1inv of electric component c at its points inside [vmin, vmax] (the design region), by
   the transposed step run backwards in time.  Its counterpart in the reference is the adjoint
   solver's gradient (python/adjoint/optimization_problem.py, src/meepgeom.cpp:2835-3011),
   which is formed from forward and adjoint DFT fields in the frequency domain; this one is
   the exact gradient of the discrete time-stepped J.
   Create it before the steps to be differentiated: they then record D at the design points,
   every step when checkpoint_interval is 0, or the whole state every checkpoint_interval steps
   (recording must then start at t = 0; the backward sweep recomputes each segment from its
   checkpoint).  lo_xyz and n_xyz (either may be NULL) receive the design box's first array
   index and size per direction (0 and 1 for absent directions).  Real fields, no dispersive materials, no integrated sources, real Bloch phases
   only.  One per fields object. */
mp_adjoint *mp_adjoint_new(mp_fields *f, mp_dft_fields *objective, int c, const double *vmin,
                           const double *vmax, int checkpoint_interval, int *lo_xyz,
                           int *n_xyz);
void mp_adjoint_delete(mp_adjoint *a);
/* J of the accumulators as they stand. */
int mp_adjoint_objective(mp_adjoint *a, double *J);
/* dJ / dchi1inv over the design box, in the reference layout order restricted to the box
   (X slowest, the stride-1 direction fastest).  The fields end in the state they had. */
int mp_adjoint_gradient(mp_adjoint *a, double *grad);

/* ---- many independent simulations stepped together -------------------------------- */

/* A batch of n fields objects of one precision (each with its own structure, sources and
   monitors, any sizes) that advance together: each launch group of a step (B phase, its
   halos, D phase, its halos, DFT) covers every object in one kernel, source values for the
   steps are computed on the host before the loop, and the loop is captured in a hipGraph.
   No counterpart in the reference, which runs one simulation per process; the objects remain
   usable on their own between batch calls.  Objects whose next step would allocate arrays
   take that step on their own first (it counts towards n).  Long source lists that fall back
   to separate launches are not supported in a batch. */
mp_batch *mp_batch_new(mp_fields *const *fields, int n);
void mp_batch_delete(mp_batch *b);
/* n steps of every object. */
int mp_batch_step_n(mp_batch *b, long n);
/* use_graph: capture the step loop in a hipGraph (default 1); autotune: time the first steps
   with each block size (64, 128, 256, 512) and keep the fastest (default 1); block: fix the
   block size instead (0 keeps the current choice). */
int mp_batch_set_options(mp_batch *b, int use_graph, int autotune, int block);
/* the block size in use and the wall time per step of the last call. */
int mp_batch_stats(mp_batch *b, int *block, double *ms_per_step);

#ifdef __cplusplus
}
#endif

#endif /* MEEP_AMD_H */
