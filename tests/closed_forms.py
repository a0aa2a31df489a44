"""Closed forms that pin the parts of the FDTD step the reference's golden values leave open.

Each check runs one small simulation through the C-ABI of any library (the CPU oracle, a
planted-mistake build of it, or the MI355X engine) and returns the relative error of the
run against a value the mathematics fixes; the tests assert on these errors.

* ``line_source_1d``: a current sheet in 1D vacuum.  On the infinite Yee lattice with Courant
  number S the update equations are solved exactly in the frequency domain: with
  theta = omega dt / 2 the field is E_j = A exp(i K |j - j0|), sin(K / 2) = sin(theta) / S, and
  the jump at the sheet gives A = -sin(theta) dt J / (S^2 sin K), where J is the transform of
  the current density sequence the step subtracts from D.  H between two E points follows
  from the B update, H_{j+1/2} = S (E_{j+1} - E_j) / (2 i sin theta).  The reference's DFT of a
  field is dt / sqrt(2 pi) times the sum of its samples times exp(i omega t) at the sample
  times (E at t dt, H half a step earlier; src/dft.cpp:98,253,270), on the centred grid (the
  mean of the two E points around a centred point), and the flux is Re(E conj(H)) there.
  The continuum limit A -> -I / 2 is Maxwell's sheet radiating half its current each way.
  This pins the DFT scale, its phase convention and the E/H time offset, the current
  injection's dt and resolution factors, and the flux of one point.
* ``plane_wave_2d``: the same sheet as a periodic line current in 2D TM; the flux through a
  transverse line of length Ly is Ly times the 1D flux, which pins the integration weights.
* ``source_spectrum``: the discrete Gaussian current's transform against the Gaussian
  integral (the pulse's Fourier transform in closed form).
* ``pole_free_response``: a uniform field in a periodic cell of a Lorentz or Drude medium
  with D held fixed (or, for a magnetic pole, the same on H with B held fixed).  The reference's continuum model, chi(omega) = sigma omega0^2 /
  (omega0^2 - omega^2 - i gamma omega) with omega0 = 2 pi f0 and gamma in 2 pi units
  (src/susceptibility.cpp:297-307), makes P relax to sigma D / (1 + sigma) (Lorentz) or to D
  (Drude: no restoring term), oscillating at sqrt(omega0^2 (1 + sigma) - (pi gamma)^2)
  (Lorentz) or sqrt(omega0^2 sigma - (pi gamma)^2) (Drude) with amplitude decay rate
  pi gamma; the discrete recurrence approaches this as (omega dt)^2.
* ``dft_array_plane_wave``: fields::get_dft_array on the periodic line current of
  ``plane_wave_2d``, whose field is uniform along the line.  From a flux monitor every point's
  value is the lattice field (the E accumulators lose their stored weight, the flux sign, and
  their integration weight; H has neither), from a field monitor on a line between two
  centred planes it is the linear interpolation of the two planes' values (the interpolation
  weights of the empty direction, then the sum over it), and from a monitor that crosses the
  periodic boundary every point of both pieces holds the field (each piece placed by its
  shift).
* ``flux_time_domain``: the instantaneous flux and energy integrals (flux_in_box_wrongH,
  flux_in_box, flux_vol, the energy in a box) on fields set by hand in a 2D cell.  The
  integrals are sums over loop_in_chunks points with weights that add up to the volume's
  length or area, so a uniform field gives field x extent; a field linear in x, read on the
  centred grid as the mean of its two Yee neighbours, gives its value at the centred line.
  Under uniform electric and magnetic conductivities the curl-free fields of a periodic cell
  decay by fixed factors per step (``conductive_decay``), so after one step the flux plane's
  mean of the fluxes after the H and the E updates, the half step that synchronises H, and
  the synchronised magnetic energy all have closed forms.
* ``pml_reflection``: a uPML layer of asymptotic reflection R (a custom profile u^3 with its
  integrals 1/4 and 1/5, and the default u^2) reflects power R in the continuum: the
  attenuation of each pass is exp(-int sigma) with int sigma = -ln(R) / 4 (structure_chunk::
  use_pml, src/structure.cpp:630-691).  The reflected wave is the difference between a run
  with the layer (metal behind it) and one with a thick layer of R = 1e-15.
* ``pml_stretch``: a lossless (R = 1) layer of thickness d and mean stretch s is, in the
  continuum, the coordinate map x -> int kappa dx, i.e. a layer of vacuum s d thick: the
  fields in front of it equal those of a cell longer by (s - 1) d, up to the lattice's
  O(dx^2) error.
* ``source_array``: a line source whose amplitudes come from an array (add_volume_source with
  arr, src/sources.cpp:394-417) of values 0, 1, ..., n - 1 equals the same line with the
  amplitude function the array's linear interpolation defines: the position mapped to
  r = 0.5 + rel / size, the value clamp(n r - 1/2, 0, n - 1) (cell-centred samples, mirror
  boundaries).
* ``srcdata_moves``: add_srcdata at the point and amplitude add_volume_source gives a point
  source (amplitude a^2 in 2D) reproduces it; a point on the periodic halo marked for the
  boundary fix equals the point that owns it with the amplitude times exp(i k L) (the
  conjugate of the phase that locates it), and one not marked, or with no owner behind a
  metal wall, is dropped.
* ``near2far_green``: green2d / green3d (src/near2far.cpp:133-267) against the dyadic Green's
  function of the homogeneous medium.  With G the scalar Green's function ((i/4) H0(k r) in 2D,
  exp(i k r) / (4 pi r) in 3D) and p the current's direction, an electric current radiates
  E = i omega mu (G p + grad grad.(G p) / k^2) and H = G'(r) r x p, a magnetic current
  H = i omega eps (G p + grad grad.(G p) / k^2) and E = -G'(r) r x p (Maxwell's equations with
  e^{-i omega t}); the Hessian of the radial G is G'' r r + G' / r (P - r r), P the projector
  onto the directions G varies in.
* ``near2far_equivalence``: the equivalence principle on the lattice.  A point current inside
  a closed box of near2far surfaces in a homogeneous medium: the far-field transform of the
  box's tangential fields at points outside the box equals the fields the simulation itself
  accumulates there (dft_fields at those points), and the transform's flux through a large
  box far away equals the DFT flux through the near box, both to the lattice's discretisation
  error (the reference's own near2far test allows 0.05 * 30 / resolution).
* ``near2far_periodic``: with Nperiods copies of a surface that spans a Bloch-periodic cell,
  the transform at x is the sum over n of exp(i n 2 pi k L) times the single-period transform
  at x - n L (a copy of every current sits n L away with the Bloch phase of its offset).
* ``dft_data_reload``: the normalisation-run pattern of the Python layer (get_flux_data, then
  load_minus_flux_data into an identical run before it steps): the second run's flux and
  near2far transform are zero up to roundoff, since its accumulators are the first run's
  minus themselves; and scale_dfts(s) multiplies a flux by |s|^2 (E and H both scaled) and a
  transform by s.
* ``dft_energy_plane_wave``: dft_energy in a 1D vacuum line: a wave travelling away from a
  current sheet carries the same electric and magnetic DFT energy, and the total over a
  stretch of length L equals L times the DFT flux through a plane (|E| = |H| and S = u in the
  continuum), up to the lattice's O(dx^2) dispersion, checked at two resolutions.
* ``dft_force_stress``: dft_force (the Maxwell stress tensor T = E E* + H H* - |F|^2 / 2) on
  travelling waves: in a 1D line the force through a plane along its normal is minus the DFT
  flux (T_zz = -|E|^2 with |H| = |E|); for an oblique wave in 2D (a line current with a Bloch
  phase k_y, at angle sin(theta) = k_y / f) the off-diagonal force along y through an x-plane
  is -(k_y / f) times the flux and the diagonal one along x is -cos(theta) times the flux,
  all to the lattice's O(dx^2).
* ``magnetic_cavity``: a 1D cavity with a magnetic Low wall and a metallic High wall.  The
  magnetic wall zeroes H at the first half pixel (src/boundaries.cpp:197-199) and the metallic
  one E at the last point, so E_j = cos(K (j - 1/2)) with K (n - 1/2) = (m + 1/2) pi is an
  exact eigenvector of the discrete update: started from it, E stays proportional to it, with
  an amplitude a_k obeying a_{k+1} + a_{k-1} = 2 cos(omega dt) a_k at the lattice frequency
  sin(omega dt / 2) = S sin(K / 2).
"""

import math

import numpy as np

from meep_amd.capi import (ARR_F, ARR_P, Bx, By, Dx, Dz, E_stuff, Ex, Ey, Ez, H_stuff, Hx, Hy, Hz, High, Low, Magnetic, Metallic, X, Y,
                           Z, continuous_src_time, fields, gaussian_src_time, structure)

S = 0.5  # Courant number of every check


def _dipole(fcen, width, t0, cutoff):
    """gaussian_src_time::dipole (src/sources.cpp:98-108): the source's time dependence."""
    def d(t):
        tt = t - t0
        v = np.exp(-tt * tt / (2 * width * width)) * np.exp(-2j * math.pi * fcen * tt) / (-2j * math.pi * fcen)
        return np.where(np.abs(tt).astype(np.float32) > np.float32(cutoff), 0.0, v)
    return d


def current_transform(fcen, fwidth, dt, nsteps, freqs):
    """sum_n Re(current((n + 1/2) dt)) exp(i omega (n + 1/2) dt): the transform of the current
    the D update at step n subtracts (times dt), current(t) = (dipole(t + dt) - dipole(t)) / dt
    (src/meep.hpp:989-991) taken at t = (n + 1/2) dt (src/step.cpp:92)."""
    width = 1.0 / fwidth
    d = _dipole(fcen, width, 5 * width, float(np.float32(5 * width)))
    t = (np.arange(nsteps) + 0.5) * dt
    cur = ((d(t + dt) - d(t)) / dt).real
    return np.array([np.sum(cur * np.exp(2j * math.pi * fr * t)) for fr in freqs])


def lattice_field(freqs, dt, J, j, j0, half=False, loss=0.0):
    """The exact frequency-domain field of the 1D Yee lattice driven by the current transform
    J at point j0: E at integer j, or H at j + 1/2 (half=True).  loss = sigma dt / 2 of a
    uniform electric conductivity sigma (eps = 1): the D update (1 + loss) D' - (1 - loss) D =
    -S curl H - J dt turns one factor sin(theta) of the dispersion relation into
    sin(theta) + i loss cos(theta), so sin^2(K / 2) = sin(theta) (sin(theta) + i loss
    cos(theta)) / S^2 with Im K > 0, while the jump at the sheet keeps its form."""
    theta = math.pi * np.asarray(freqs) * dt
    s2 = np.sin(theta) * (np.sin(theta) + 1j * loss * np.cos(theta)) / (S * S)
    K = 2 * np.arcsin(np.sqrt(s2.astype(complex)))
    K = np.where(K.imag < 0, -K, K)
    A = -np.sin(theta) * dt * J / (S * S * np.sin(K))

    def E(jj):
        return A * np.exp(1j * K * abs(jj - j0))
    if not half:
        return E(j)
    return S * (E(j + 1) - E(j)) / (2j * np.sin(theta))


def line_source_1d(lib, a=20, freqs=None, sigma=0.0, pml=0.0, T=32.0):
    """Errors of the 1D current sheet's DFT fields and flux against the lattice solution.
    sigma: a uniform electric conductivity (set through set_conductivity, scaled by
    chi1inv = 1); pml: uPML of that thickness at both ends (the metallic walls are far
    enough that nothing returns from them within T = 32)."""
    if freqs is None:
        freqs = np.linspace(0.88, 1.12, 7)
    L, z0, fcen, fwidth = 80.0, 30.0, 1.0, 0.5
    s = structure(lib, dim=1, size=(0, 0, L), a=a, pml=[(Z, side, pml) for side in (High, Low)] if pml else ())
    if sigma:
        s.set_conductivity(Ex, lambda x, y, z: sigma)
    f = fields(s)
    f.use_real_fields()
    f.add_volume_source(Ex, gaussian_src_time(fcen, fwidth), (0, 0, z0), (0, 0, z0))
    # centred points (i + 1/2) / a after and before the sheet: 5 units away each side
    zs = {"fwd": z0 + 5.0 + 0.5 / a, "bwd": z0 - 5.0 + 0.5 / a}
    mons = {k: f.add_dft_fields([Ex, Hy], (0, 0, z), (0, 0, z), freqs, decimation_factor=1) for k, z in zs.items()}
    flx = f.add_dft_flux(Z, (0, 0, zs["fwd"]), (0, 0, zs["fwd"]), freqs, decimation_factor=1)
    dt = s.dt
    nsteps = int(round(T / dt))
    f.step_n(nsteps)
    J = a * current_transform(fcen, fwidth, dt, nsteps, freqs)  # the sheet's density: amp * a
    j0 = int(round(z0 * a))
    scale = dt / math.sqrt(2 * math.pi)
    loss = sigma * dt / 2
    err = {}
    for k, z in zs.items():
        i = int(round(z * a - 0.5))
        e_exp = scale * 0.5 * (lattice_field(freqs, dt, J, i, j0, loss=loss) + lattice_field(freqs, dt, J, i + 1, j0, loss=loss))
        h_exp = scale * lattice_field(freqs, dt, J, i, j0, half=True, loss=loss)
        got = {}
        for q in range(mons[k].num_chunks()):
            d = mons[k].dft(q)
            assert d.shape[0] == 1, "a centred point is one DFT point"
            got[mons[k].component(q)] = d[0]
        err[f"E_{k}"] = float(np.abs(got[Ex] - e_exp).max() / np.abs(e_exp).max())
        err[f"H_{k}"] = float(np.abs(got[Hy] - h_exp).max() / np.abs(h_exp).max())
        if k == "fwd":
            flux_exp = np.real(e_exp * np.conj(h_exp))
            err["flux"] = float(np.abs(flx.flux() - flux_exp).max() / np.abs(flux_exp).max())
    return err


def plane_wave_2d(lib, a=10, freqs=None):
    """Error of the flux of a periodic line current in 2D TM through a transverse line of
    length Ly against Ly times the 1D lattice solution's flux."""
    if freqs is None:
        freqs = np.linspace(0.88, 1.12, 7)
    Lx, Ly, x0, fcen, fwidth, T = 44.0, 1.3, 14.0, 1.0, 0.5, 28.0
    s = structure(lib, dim=2, size=(Lx, Ly), a=a)
    f = fields(s)
    f.use_real_fields()
    f.use_bloch((0.0, 0.0))
    f.set_boundary(Low, X, Metallic)
    f.set_boundary(High, X, Metallic)
    f.add_volume_source(Ez, gaussian_src_time(fcen, fwidth), (x0, 0, 0), (x0, Ly, 0))
    x1 = x0 + 5.0 + 0.5 / a
    flx = f.add_dft_flux(X, (x1, 0, 0), (x1, Ly, 0), freqs, decimation_factor=1)
    dt = s.dt
    nsteps = int(round(T / dt))
    f.step_n(nsteps)
    J = a * current_transform(fcen, fwidth, dt, nsteps, freqs)
    j0, i = int(round(x0 * a)), int(round(x1 * a - 0.5))
    scale = dt / math.sqrt(2 * math.pi)
    e = scale * 0.5 * (lattice_field(freqs, dt, J, i, j0) + lattice_field(freqs, dt, J, i + 1, j0))
    h = scale * lattice_field(freqs, dt, J, i, j0, half=True)
    # TM along x: Ez plays 1D Ex and Hy plays -(1D Hy); the X flux is -Re(Ez conj(Hy))
    flux_exp = Ly * np.real(e * np.conj(h))
    return {"flux": float(np.abs(flx.flux() - flux_exp).max() / np.abs(flux_exp).max())}


def dft_array_plane_wave(lib, a=10, freqs=None):
    """Errors of get_dft_array on a periodic 2D TM line current (a plane wave along x,
    uniform along y) against the 1D lattice solution: the flux monitor's Ez and Hy arrays along
    a centred line, a field monitor's Ez on a line 0.3 pixel past a centred line (the linear
    interpolation of two centred lines), and a field monitor across the periodic y boundary.
    Points of a weighted (flux E) chunk whose integration weight is zero hold zero, as the
    reference's do; the others must hold the field."""
    if freqs is None:
        freqs = np.linspace(0.9, 1.1, 3)
    Lx, Ly, x0, fcen, fwidth, T = 44.0, 1.3, 14.0, 1.0, 0.5, 28.0
    s = structure(lib, dim=2, size=(Lx, Ly), a=a)
    f = fields(s)
    f.use_real_fields()
    f.use_bloch((0.0, 0.0))
    f.set_boundary(Low, X, Metallic)
    f.set_boundary(High, X, Metallic)
    f.add_volume_source(Ez, gaussian_src_time(fcen, fwidth), (x0, 0, 0), (x0, Ly, 0))
    x1 = x0 + 5.0 + 0.5 / a
    flx = f.add_dft_flux(X, (x1, 0.2, 0), (x1, 1.1, 0), freqs, decimation_factor=1)
    mid = f.add_dft_fields([Ez], (x1 + 0.3 / a, 0.25, 0), (x1 + 0.3 / a, 1.05, 0), freqs, decimation_factor=1)
    wrap = f.add_dft_fields([Ez], (x1, Ly - 0.35, 0), (x1, Ly + 0.35, 0), freqs, decimation_factor=1)
    dt = s.dt
    nsteps = int(round(T / dt))
    f.step_n(nsteps)
    J = a * current_transform(fcen, fwidth, dt, nsteps, freqs)
    j0, i = int(round(x0 * a)), int(round(x1 * a - 0.5))
    scale = dt / math.sqrt(2 * math.pi)
    def e_at(k):
        return scale * 0.5 * (lattice_field(freqs, dt, J, k, j0) + lattice_field(freqs, dt, J, k + 1, j0))
    e, e1 = e_at(i), e_at(i + 1)
    h = scale * lattice_field(freqs, dt, J, i, j0, half=True)
    err = {"E_flux": 0.0, "H_flux": 0.0, "E_interp": 0.0, "E_wrap": 0.0}
    for nf in range(len(freqs)):
        ez = flx.get_dft_array(Ez, nf)
        hy = flx.get_dft_array(Hy, nf)
        assert ez.ndim == 1 and ez.shape == hy.shape and ez.size >= int(0.9 * a) - 1, (ez.shape, hy.shape)
        held = np.abs(ez) > 0.5 * abs(e[nf])
        assert held.sum() >= ez.size - 2, np.abs(ez)
        err["E_flux"] = max(err["E_flux"], float(np.abs(ez[held] - e[nf]).max() / abs(e[nf])))
        # TM along x: Hy is minus the 1D H of the sheet
        err["H_flux"] = max(err["H_flux"], float(np.abs(hy + h[nf]).max() / abs(h[nf])))
        m = mid.get_dft_array(Ez, nf)
        assert m.ndim == 1 and m.size >= int(0.8 * a), m.shape
        m_exp = 0.7 * e[nf] + 0.3 * e1[nf]
        err["E_interp"] = max(err["E_interp"], float(np.abs(m - m_exp).max() / abs(m_exp)))
        w = wrap.get_dft_array(Ez, nf)
        assert w.ndim == 1 and w.size >= int(0.7 * a), w.shape
        err["E_wrap"] = max(err["E_wrap"], float(np.abs(w - e[nf]).max() / abs(e[nf])))
    return err


def flux_time_domain(lib, a=10):
    """Errors of the time-domain flux and energy integrals against their closed forms (see the
    module docstring)."""
    err = {}
    # a field linear in x, read on the centred line x0 through a full y period, never stepped
    s = structure(lib, dim=2, size=(2.0, 1.0), a=a)
    f = fields(s)
    f.use_real_fields()
    f.use_bloch((0.0, 0.0))
    f.add_volume_source(Ez, continuous_src_time(1.0), (0.55, 0.55, 0), (0.55, 0.55, 0), amp=0.0)
    nx, ny = s.shape
    h0 = 0.7
    f.set_array(Ez, 0, ARR_F, np.repeat(np.arange(nx)[:, None] / a, ny, axis=1))
    f.set_array(Hy, 0, ARR_F, np.full(s.shape, h0))
    f.set_array(By, 0, ARR_F, np.full(s.shape, h0))
    x0 = 12.5 / a
    got = f.flux_in_box_wrongH(X, (x0, 0, 0), (x0, 1.0, 0))
    err["linear"] = abs(got - (-x0 * h0 * 1.0)) / (x0 * h0)
    # uniform fields decaying under conductivity in a periodic cell with eps = 2
    eps, se, sm, e0, g0 = 2.0, 1.2, 0.8, 0.9, -0.4
    s = structure(lib, dim=2, size=(1.0, 1.0), a=a, eps=lambda x, y, z: eps)
    s.set_conductivity(Ez, lambda x, y, z: se)
    s.set_conductivity(Hy, lambda x, y, z: sm)
    f = fields(s)
    f.use_real_fields()
    f.use_bloch((0.0, 0.0))
    f.add_volume_source(Ez, continuous_src_time(1.0), (0.55, 0.55, 0), (0.55, 0.55, 0), amp=0.0)
    one = np.ones(s.shape)
    for c, v in ((Ez, e0), (Dz, eps * e0), (Hx, g0), (Bx, g0), (Hy, h0), (By, h0)):
        f.set_array(c, 0, ARR_F, v * one)
    box = ((0.2, 0.3, 0), (0.7, 0.9, 0))
    area = 0.5 * 0.6
    err["energy_e"] = abs(f.electric_energy_in_box(*box) - 0.5 * e0 * eps * e0 * area) / (e0 * e0 * area)
    err["energy_h"] = abs(f.magnetic_energy_in_box(*box) - 0.5 * (g0 * g0 + h0 * h0) * area) / (h0 * h0 * area)
    err["flux_y"] = abs(f.flux_in_box_wrongH(Y, (0.2, 0.45, 0), (0.7, 0.45, 0)) - e0 * g0 * 0.5) / abs(e0 * g0 * 0.5)
    plane = f.add_flux_plane((x0 - 1.0, 0, 0), (x0 - 1.0, 1.0, 0))
    f.step()
    dt = s.dt
    def rate(c):
        return (1 - c * dt / 2) / (1 + c * dt / 2)
    r_e, r_m = rate(se / eps), rate(sm)
    ref = e0 * h0
    # after the step: E at t + 1 (r_e e0), H at t + 1/2 (r_m h0); the plane averaged the flux
    # after the H update (E still e0) with the one after the E update
    err["flux_vol"] = abs(plane.flux() - (-e0 * r_m * h0 * (1 + r_e) / 2)) / ref
    err["wrongH"] = abs(f.flux_in_box_wrongH(X, (x0 - 1.0, 0, 0), (x0 - 1.0, 1.0, 0)) - (-r_e * e0 * r_m * h0)) / ref
    h_sync = r_m * h0 * (1 + r_m) / 2  # H half a step on, averaged with H now
    err["sync"] = abs(f.flux_in_box(X, (x0 - 1.0, 0, 0), (x0 - 1.0, 1.0, 0)) - (-r_e * e0 * h_sync)) / ref
    e_exp = 0.5 * eps * (r_e * e0) ** 2 * area + 0.5 * (g0 * g0 + h_sync * h_sync) * area
    err["energy_sync"] = abs(f.field_energy_in_box(*box) - e_exp) / e_exp
    return err


def _pml_line(lib, L, terms, a, zs=3.0, zm=5.0, T=80.0, freqs=(0.4, 0.5, 0.6)):
    """DFT of Ex at zm after a pulse from zs on a 1D line of length L with the uPML terms."""
    s = structure(lib, dim=1, size=(0, 0, L), a=a, pml=terms)
    f = fields(s)
    f.use_real_fields()
    f.add_volume_source(Ex, gaussian_src_time(0.5, 0.4), (0, 0, zs), (0, 0, zs))
    m = f.add_dft_fields([Ex], (0, 0, zm), (0, 0, zm), list(freqs), decimation_factor=1)
    f.step_n(int(round(T / s.dt)))
    return m.dft(0)[0]


def pml_reflection(lib, a=20, R=1e-3):
    """|r|^2 / R - 1 over three frequencies for a cubic-profile layer and a quadratic one."""
    inc = _pml_line(lib, 14.0, [(Z, Low, 2.0, 1e-15), (Z, High, 6.0, 1e-15)], a)
    out = {}
    for name, term in (("cubic", (Z, High, 2.0, R, 1.0, lambda u: u ** 3, 0.25, 0.2)), ("quadratic", (Z, High, 2.0, R))):
        tot = _pml_line(lib, 14.0, [(Z, Low, 2.0, 1e-15), term], a)
        out[name] = float(np.abs(np.abs(tot - inc) ** 2 / np.abs(inc) ** 2 / R - 1).max())
    return out


def pml_stretch(lib, a=40, d=1.0, s=2.0):
    """Relative difference of the fields in front of a lossless stretched layer and of those
    in a cell longer by (s - 1) d."""
    A = _pml_line(lib, 10.0, [(Z, Low, 2.0, 1e-15), (Z, High, d, 1.0, s)], a)
    B = _pml_line(lib, 10.0 + (s - 1) * d, [(Z, Low, 2.0, 1e-15)], a)
    return float(np.abs(A - B).max() / np.abs(B).max())


def _src_run(lib, setup, nsteps=120, k=None, metal_x=False):
    s = structure(lib, dim=2, size=(3.0, 2.5), a=10)
    f = fields(s)
    if k is None:
        f.use_real_fields()
        f.use_bloch((0.0, 0.0))
    else:
        f.use_bloch((0.0, k))
    if metal_x:
        f.set_boundary(Low, X, Metallic)
        f.set_boundary(High, X, Metallic)
    setup(s, f)
    f.step_n(nsteps)
    return f.get_array(Ez, 0)[0], (f.get_array(Ez, 1)[0] if k is not None else np.zeros(s.shape))


def source_array(lib, n=7):
    """Relative difference between a line source with an amplitude array and the same line
    with the amplitude function the array's interpolation defines."""
    src = gaussian_src_time(1.0, 0.5)
    # ends on grid points, so every source point lies inside the volume (0 <= r <= 1)
    lo, hi = (1.3, 0.3, 0), (1.3, 2.2, 0)
    data = np.arange(n, dtype=float) * (1.0 + 0.5j)
    def fn(x, y, z):
        r = 0.5 + y / (hi[1] - lo[1])
        return min(max(n * r - 0.5, 0.0), n - 1.0) * (1.0 + 0.5j)
    A = _src_run(lib, lambda s, f: f.add_volume_source_array(Ez, src, lo, hi, data.reshape(1, n, 1)), k=0.0)
    B = _src_run(lib, lambda s, f: f.add_volume_source(Ez, src, lo, hi, amp_func=fn), k=0.0)
    scale = max(np.abs(B[0]).max(), np.abs(B[1]).max())
    return float(max(np.abs(A[0] - B[0]).max(), np.abs(A[1] - B[1]).max()) / scale)


def srcdata_moves(lib, k=0.3):
    """Errors of add_srcdata against the sources it stands for (see the module docstring)."""
    src = gaussian_src_time(1.0, 0.5)
    a = 10
    err = {}
    def ref_index(s, i, j):
        return i * s.shape[1] + j  # Ez sits at integer points; Y is the stride-1 direction
    # a point source at grid point (13, 11) / a: one point of amplitude a^2
    P = _src_run(lib, lambda s, f: f.add_volume_source(Ez, src, (1.3, 1.1, 0), (1.3, 1.1, 0)), k=k)
    Q = _src_run(lib, lambda s, f: (f.add_srcdata(Ez, src, [ref_index(s, 13, 11)], [a * a]), f.require_source_components()), k=k)
    scale = max(np.abs(P[0]).max(), np.abs(P[1]).max())
    err["point"] = float(max(np.abs(P[0] - Q[0]).max(), np.abs(P[1] - Q[1]).max()) / scale)
    # the halo point y = 0 of a periodic y with Bloch k: its owner is y = L with exp(i k L)
    L = 2.5
    ph = np.exp(2j * math.pi * k * L)
    H = _src_run(lib, lambda s, f: (f.add_srcdata(Ez, src, [ref_index(s, 13, 0)], [a * a], True), f.require_source_components()), k=k)
    O = _src_run(lib, lambda s, f: (f.add_srcdata(Ez, src, [ref_index(s, 13, 25)], [a * a * ph]), f.require_source_components()), k=k)
    scale = max(np.abs(O[0]).max(), np.abs(O[1]).max())
    err["moved"] = float(max(np.abs(H[0] - O[0]).max(), np.abs(H[1] - O[1]).max()) / scale)
    # not marked: dropped; behind a metal wall (x = 0, no periodic owner): dropped
    D = _src_run(lib, lambda s, f: (f.add_srcdata(Ez, src, [ref_index(s, 13, 0)], [a * a]), f.require_source_components()), k=k)
    M = _src_run(lib, lambda s, f: (f.add_srcdata(Ez, src, [ref_index(s, 0, 11)], [a * a], True), f.require_source_components()),
                 k=k, metal_x=True)
    err["unmarked"] = float(max(np.abs(D[0]).max(), np.abs(D[1]).max()) / scale)
    err["metal"] = float(max(np.abs(M[0]).max(), np.abs(M[1]).max()) / scale)
    return err


def source_spectrum(dt=0.025, fcen=1.0, fwidth=0.5, freqs=None):
    """Error of the discrete current transform against the Gaussian integral:
    dt * sum ~ D(omega) (exp(-i omega dt) - 1) / dt with D the dipole's transform
    exp(i omega t0) w sqrt(2 pi) exp(-w^2 (omega - omega0)^2 / 2) / (-i omega0)."""
    if freqs is None:
        freqs = np.linspace(0.88, 1.12, 7)
    w = 1.0 / fwidth
    t0 = 5 * w
    n = int(round(12 * w / dt))
    got = dt * current_transform(fcen, fwidth, dt, n, freqs)
    om, om0 = 2 * math.pi * np.asarray(freqs), 2 * math.pi * fcen
    D = np.exp(1j * om * t0) * w * math.sqrt(2 * math.pi) * np.exp(-w * w * (om - om0) ** 2 / 2) / (-1j * om0)
    exp = 0.5 * D * (np.exp(-1j * om * dt) - 1) / dt  # Re(): half the positive-frequency part
    return float(np.abs(got - exp).max() / np.abs(exp).max())


def pole_free_response(lib, drude, a=100, f0=1.0, gamma=0.3, sigma=2.0, T=20.0, magnetic=False):
    """Errors of a pole's relaxation against the continuum model: the final P against its
    static value, and the oscillation frequency and decay rate of P against the model's.
    magnetic: the same pole on H (add_susceptibility with H_stuff), with B held fixed."""
    s = structure(lib, dim=1, size=(0, 0, 3.0 / a), a=a)
    s.add_susceptibility(lambda x, y, z: sigma, f0, gamma, drude=drude, ft=H_stuff if magnetic else E_stuff)
    f = fields(s)
    f.use_real_fields()
    f.use_bloch((0.0,))
    # a source of zero amplitude allocates the fields without driving them
    f.add_volume_source(Ex, continuous_src_time(1.0), (0, 0, 0.5 / a), (0, 0, 0.5 / a), amp=0.0)
    one = np.ones(s.shape)
    fc, gc = (Hy, By) if magnetic else (Ex, Dx)
    f.set_array(gc, 0, ARR_F, one)
    f.set_array(fc, 0, ARR_F, one)
    dt = s.dt
    n = int(round(T / dt))
    P = np.empty(n)
    for k in range(n):
        f.step()
        P[k] = f.get_array(fc, 0, ARR_P, 0)[0][1]
    p_inf = 1.0 if drude else sigma / (1.0 + sigma)
    om0 = 2 * math.pi * f0
    rate = math.pi * gamma
    om = math.sqrt(om0 * om0 * (sigma if drude else 1.0 + sigma) - rate * rate)
    # the model's free response about p_inf, fitted as x_{k+1} = alpha x_k + beta x_{k-1}
    x = P - p_inf
    M = np.stack([x[1:-1], x[:-2]], axis=1)
    alpha, beta = np.linalg.lstsq(M, x[2:], rcond=None)[0]
    lam = np.roots([1.0, -alpha, -beta])
    lam = lam[np.argmax(lam.imag)]
    om_fit = abs(np.angle(lam)) / dt
    rate_fit = -math.log(abs(lam)) / dt
    # after T the free response has decayed by exp(-pi gamma T) (2.6e-9 here)
    return {"p_final": float(abs(P[-1] - p_inf) / p_inf),
            "frequency": abs(om_fit - om) / om, "decay": abs(rate_fit - rate) / rate}


def magnetic_cavity(lib, a=20, L=2.0, m=2, nsteps=300):
    """Errors of a quarter-wave cavity mode between a magnetic Low wall and a metallic High
    wall: the field's departure from the mode shape and the amplitude's departure from the
    lattice frequency's recurrence (both relative to the amplitude)."""
    s = structure(lib, dim=1, size=(0, 0, L), a=a)
    f = fields(s)
    f.use_real_fields()
    f.set_boundary(Low, Z, Magnetic)
    f.set_boundary(High, Z, Metallic)
    f.add_volume_source(Ex, continuous_src_time(1.0), (0, 0, 0.5), (0, 0, 0.5), amp=0.0)
    n = int(round(L * a))
    K = (m + 0.5) * math.pi / (n - 0.5)
    j = np.arange(n + 1)
    mode = np.cos(K * (j - 0.5))
    mode[n] = 0.0
    f.set_array(Dx, 0, ARR_F, mode)
    f.set_array(Ex, 0, ARR_F, mode)
    dt = s.dt
    om = 2 * math.asin(S * math.sin(K / 2)) / dt
    amp, resid = [], 0.0
    owned = slice(1, n + 1)  # E's owned points (index 0 is the halo plane)
    for _ in range(nsteps):
        f.step()
        e = f.get_array(Ex, 0)[0]
        c = float(np.dot(e[owned], mode[owned]) / np.dot(mode[owned], mode[owned]))
        amp.append(c)
        resid = max(resid, float(np.abs(e[owned] - c * mode[owned]).max()))
    amp = np.array(amp)
    rec = np.abs(amp[2:] + amp[:-2] - 2 * math.cos(om * dt) * amp[1:-1]).max()
    return {"shape": resid / np.abs(amp).max(), "frequency": float(rec / np.abs(amp).max())}


def conductive_decay(lib, eps=3.0, sigma=0.7, sigma_m=0.4, nsteps=200, a=20):
    """Errors of uniform fields in a periodic 1D cell with electric conductivity sigma (in a
    medium of permittivity eps) and magnetic conductivity sigma_m.  The curl of a uniform
    field vanishes, so each step multiplies D by (1 - s dt / 2) / (1 + s dt / 2) with s =
    sigma / eps (the reference scales an electric conductivity by chi1inv,
    src/structure.cpp:878-888) and B likewise with s = sigma_m; E = D / eps."""
    s = structure(lib, dim=1, size=(0, 0, 3.0 / a), a=a, eps=lambda x, y, z: eps)
    s.set_conductivity(Ex, lambda x, y, z: sigma)
    s.set_conductivity(Hy, lambda x, y, z: sigma_m)
    f = fields(s)
    f.use_real_fields()
    f.use_bloch((0.0,))
    f.add_volume_source(Ex, continuous_src_time(1.0), (0, 0, 0.5 / a), (0, 0, 0.5 / a), amp=0.0)
    one = np.ones(s.shape)
    f.set_array(Dx, 0, ARR_F, one)
    f.set_array(Ex, 0, ARR_F, one / eps)
    f.set_array(By, 0, ARR_F, one)
    f.step_n(nsteps)
    dt = s.dt
    def rate(c):
        return (1 - c * dt / 2) / (1 + c * dt / 2)
    d_exp = rate(sigma / eps) ** nsteps
    b_exp = rate(sigma_m) ** nsteps
    d = f.get_array(Dx, 0)[0]
    e = f.get_array(Ex, 0)[0]
    b = f.get_array(By, 0)[0]
    return {"D": float(np.abs(d - d_exp).max() / d_exp), "E": float(np.abs(e * eps - d_exp).max() / d_exp),
            "B": float(np.abs(b - b_exp).max() / b_exp)}


def conductive_pml_limit(lib, sigma=1e-9, a=10, freqs=None):
    """Relative change of the DFT fields of a 2D TM point source in a cell with uPML on
    every face when a vanishing conductivity sigma is added to Ez and Hx, Hy.  The
    conductive branches of the curl update (the fcnd auxiliary where uPML acts on f,
    src/step_generic.cpp:158-248) must reduce to the non-conductive ones as sigma -> 0, so
    the fields may move by O(sigma) only."""
    if freqs is None:
        freqs = np.linspace(0.9, 1.1, 5)
    out = []
    for sg in (0.0, sigma):
        s = structure(lib, dim=2, size=(5.0, 4.0), a=a, pml=[(d, side, 1.0) for d in (X, Y) for side in (High, Low)])
        if sg:
            for c in (Ez, Hx, Hy):
                s.set_conductivity(c, lambda x, y, z: sg)
        f = fields(s)
        f.use_real_fields()
        f.add_volume_source(Ez, gaussian_src_time(1.0, 0.5), (2.1, 1.7, 0), (2.1, 1.7, 0))
        mon = f.add_dft_fields([Ez, Hx, Hy], (0.3, 0.3, 0), (4.7, 3.7, 0), freqs, decimation_factor=1)
        f.step_n(int(round(30.0 / s.dt)))
        out.append([mon.dft(q) for q in range(mon.num_chunks())])
    return max(float(np.abs(x - y).max() / np.abs(y).max()) for x, y in zip(out[1], out[0]))


def pml_locality(lib, a=10, steps=10):
    """Relative change of Dx along a current plane that crosses a y uPML region when a z uPML
    is added 25 cells away (beyond the 2 cells per step the update can reach).  The reference steps each uPML
    region as a chunk of its own (src/structure.cpp:128-137), whose sigma arrays cover only
    the directions whose uPML it meets (structure_chunk::use_pml, :645-654), so a point of the
    y uPML takes the y branch alone wherever a z uPML is: the current it receives is damped by
    the y profile, and nothing can depend on the far z uPML."""
    out = []
    for zpml in (False, True):
        terms = [(Y, side, 0.8) for side in (High, Low)]
        if zpml:
            terms += [(Z, side, 0.5) for side in (High, Low)]
        s = structure(lib, dim=3, size=(1.2, 2.4, 6.0), a=a, pml=terms)
        f = fields(s)
        f.use_real_fields()
        f.add_volume_source(Ex, gaussian_src_time(0.8, 0.5), (0.0, 0.0, 3.0), (1.2, 2.4, 3.0))
        f.step_n(steps)
        out.append(f.get_array(Dx, 0)[0])
    return float(np.abs(out[1] - out[0]).max() / np.abs(out[0]).max())


def _dyadic_green(dim, x, freq, eps, mu, x0, c0, f0):
    """The six fields of a point current from the scalar Green's function (module docstring)."""
    from scipy.special import hankel1
    d = np.array(x, float) - np.array(x0, float)
    if dim == 2:
        d[2] = 0.0
    r = np.linalg.norm(d)
    rh = d / r
    w = 2 * math.pi * freq
    k = w * math.sqrt(eps * mu)
    if dim == 3:
        G = np.exp(1j * k * r) / (4 * math.pi * r)
        G1 = G * (1j * k - 1 / r)
        G2 = G * ((1j * k - 1 / r) ** 2 + 1 / r ** 2)
    else:
        G = 0.25j * hankel1(0, k * r)
        G1 = -0.25j * k * hankel1(1, k * r)
        G2 = -0.25j * k * k * (hankel1(0, k * r) - hankel1(1, k * r) / (k * r))
    p = np.zeros(3)
    p[{0: 0, 1: 1, 4: 2}[c0 % 5]] = 1.0
    proj = p.copy()
    if dim == 2:
        proj[2] = 0.0
    own = G * p + (G2 * rh * (rh @ p) + G1 / r * (proj - rh * (rh @ p))) / k ** 2
    cross = G1 * np.cross(rh, p)
    if c0 < 5:
        EH = np.concatenate([1j * w * mu * own, cross])
    else:
        EH = np.concatenate([-cross, 1j * w * eps * own])
    return EH * f0


def near2far_green(lib):
    """Largest error of green2d / green3d relative to the largest field, over both
    dimensions, the six current components and a few geometries and media."""
    from meep_amd.capi import green
    cases = [((0.9, -0.4, 0.55), (0.1, 0.2, -0.3), 0.4, 2.0, 1.3),
             ((3.1, 2.2, -1.7), (0.0, 0.0, 0.0), 1.1, 1.0, 1.0),
             ((-0.2, 0.35, 0.1), (0.15, -0.05, 0.3), 0.25, 12.0, 1.0)]
    err = 0.0
    for dim in (2, 3):
        for c0 in (Ex, Ey, Ez, Hx, Hy, Hz):
            for x, x0, fr, eps, mu in cases:
                got = green(lib, dim, x, fr, eps, mu, x0, c0, 0.3 + 0.7j)
                want = _dyadic_green(dim, x, fr, eps, mu, x0, c0, 0.3 + 0.7j)
                err = max(err, float(np.abs(got - want).max() / np.abs(want).max()))
    return err


def _n2f_box(cen, L, dim):
    """The closed box of half-width L around cen as near2far / flux surfaces (vmin, vmax,
    weight = the outward sign)."""
    out = []
    for d in range(dim):
        for sgn in (1.0, -1.0):
            lo = [c - L for c in cen]
            hi = [c + L for c in cen]
            lo[d] = hi[d] = cen[d] + sgn * L
            out.append((tuple(lo) + (0.0,) * (3 - dim), tuple(hi) + (0.0,) * (3 - dim), sgn, d))
    return out


def near2far_equivalence(lib, dim=2, c0=Ez, a=12, sz=6.0, L=0.8, freq=0.5, single=False, far_R=12.0):
    """Errors of the far-field transform of a closed box around a point current (module
    docstring): 'fields', relative to the directly accumulated fields at six points outside
    the box, and 'flux', the transform's flux out of a box of half-width far_R relative to the
    DFT flux out of the near box."""
    size = (sz,) * dim
    s = structure(lib, dim=dim, size=size, a=a, eps=lambda x, y, z: 2.0,
                  pml=[(d, side, 1.0) for d in range(dim) for side in (High, Low)], single=single)
    f = fields(s)
    cen = [sz / 2] * dim
    f.add_volume_source(c0, gaussian_src_time(freq, 0.4), tuple(cen) + (0.0,) * (3 - dim),
                        tuple(cen) + (0.0,) * (3 - dim))
    box = _n2f_box(cen, L, dim)
    n2f = f.add_dft_near2far([(lo, hi, w) for lo, hi, w, _ in box], [freq], decimation_factor=1)
    near = [(f.add_dft_flux(d, lo, hi, [freq], decimation_factor=1), w) for lo, hi, w, d in box]
    pts = []
    for i in range(6):
        r = L + 0.3 + 0.1 * i
        v = np.array([1.0, 0.5, 0.3][:dim])
        p = np.array(cen) + r * v / np.linalg.norm(v)
        pts.append(tuple(p) + (0.0,) * (3 - dim))
    comps = (Ex, Ey, Ez, Hx, Hy, Hz)
    mons = [f.add_dft_fields(list(comps), p, p, [freq], decimation_factor=1) for p in pts]
    f.step_n(int(round(40.0 / s.dt)))
    direct = np.array([[complex(m.get_dft_array(c, 0)) for c in comps] for m in mons])
    far = np.array([n2f.farfield(p)[0] for p in pts])
    fields_err = float(np.linalg.norm(far - direct) / np.linalg.norm(direct))
    flux_near = sum(w * fl.flux()[0] for fl, w in near)
    flux_far = 0.0
    for lo, hi, w, d in _n2f_box(cen, far_R, dim):
        flux_far += w * n2f.flux(d, lo, hi, 4.0)[0]
    return {"fields": fields_err, "flux": abs(flux_far - flux_near) / abs(flux_near)}


def near2far_periodic(lib, nperiods=3, ky=0.27, a=10):
    """Largest difference, relative to the field, between the transform of a surface spanning
    a Bloch-periodic cell with nperiods copies and the phased sum of single-period transforms
    at the shifted points (module docstring)."""
    sx, sy = 4.0, 2.0
    s = structure(lib, dim=2, size=(sx, sy), a=a, pml=[(X, side, 1.0) for side in (High, Low)])
    f = fields(s)
    f.use_bloch((0.0, ky))
    f.add_volume_source(Ez, gaussian_src_time(0.6, 0.4), (1.5, 0.7, 0), (1.5, 0.7, 0))
    surf = [((2.6, 0.0, 0), (2.6, sy, 0), 1.0)]
    many = f.add_dft_near2far(surf, [0.6], decimation_factor=1, nperiods=nperiods)
    one = f.add_dft_near2far(surf, [0.6], decimation_factor=1, nperiods=1)
    f.step_n(int(round(20.0 / s.dt)))
    err = 0.0
    for x in ((5.0, 0.4, 0.0), (7.5, 1.9, 0.0)):
        got = many.farfield(x)[0]
        want = sum(np.exp(2j * math.pi * ky * sy * n) * one.farfield((x[0], x[1] - n * sy, 0.0))[0]
                   for n in range(-nperiods, nperiods + 1))
        err = max(err, float(np.abs(got - want).max() / np.abs(want).max()))
    return err, many.periodic_d, many.periodic_n


def _reload_run(lib, single, before_steps=None, nsteps=150):
    s = structure(lib, dim=2, size=(5.0, 4.0), a=10, pml=[(X, side, 1.0) for side in (High, Low)],
                  single=single)
    f = fields(s)
    f.add_volume_source(Ez, gaussian_src_time(0.7, 0.4), (1.6, 2.05, 0), (1.6, 2.05, 0))
    fl = f.add_dft_flux(X, (3.3, 0.2, 0), (3.3, 3.8, 0), np.linspace(0.5, 0.9, 5), decimation_factor=1)
    n2f = f.add_dft_near2far([((2.9, 0.5, 0), (2.9, 3.5, 0), 1.0)], np.linspace(0.5, 0.9, 5), decimation_factor=1)
    if before_steps:
        before_steps(fl, n2f)
    f.step_n(nsteps)
    return f, fl, n2f


def dft_data_reload(lib, single=False):
    """Errors of the flux-data reload pattern and of scale_dfts (module docstring)."""
    x = (8.0, 1.0, 0.0)
    f1, fl1, n1 = _reload_run(lib, single)
    data, ndata = fl1.get_data(), n1.get_data()
    flux1, far1 = np.asarray(fl1.flux()), n1.farfield(x)

    def load(fl, n2f):
        fl.load_minus_data(data)
        n2f.load_minus_data(ndata)
    f2, fl2, n2 = _reload_run(lib, single, load)
    out = {"minus_flux": float(np.abs(fl2.flux()).max() / np.abs(flux1).max()),
           "minus_far": float(np.abs(n2.farfield(x)).max() / np.abs(far1).max())}
    sc = 2.0 + 1.0j
    fl1.scale_dfts(sc)
    n1.scale_dfts(sc)
    out["scale_flux"] = float(np.abs(np.asarray(fl1.flux()) - abs(sc) ** 2 * flux1).max() / np.abs(flux1).max() / 5)
    out["scale_far"] = float(np.abs(n1.farfield(x) - sc * far1).max() / np.abs(far1).max())
    return out


def dft_energy_plane_wave(lib, a=20, single=False, L=1.5, spectra=False):
    """Relative departures of a travelling wave's DFT energies from equipartition and from L
    times its DFT flux (module docstring), over five frequencies."""
    s = structure(lib, dim=1, size=(0, 0, 8.0), a=a, pml=[(Z, side, 1.0) for side in (High, Low)], single=single)
    f = fields(s)
    freqs = np.linspace(0.3, 0.5, 5)
    f.add_volume_source(Ex, gaussian_src_time(0.4, 0.3), (0, 0, 2.0), (0, 0, 2.0))
    fl = f.add_dft_flux(Z, (0, 0, 3.5), (0, 0, 3.5), freqs, decimation_factor=1)
    en = f.add_dft_energy([((0, 0, 3.5), (0, 0, 3.5 + L))], freqs, decimation_factor=1)
    f.step_n(int(round(60.0 / s.dt)))
    flux, e, m, t = np.asarray(fl.flux()), en.electric(), en.magnetic(), en.total()
    if spectra:
        return e, m, t
    return {"equipartition": float(np.abs(e - m).max() / np.abs(t).max()),
            "energy_flux": float(np.abs(t - L * flux).max() / np.abs(t).max()),
            "total": float(np.abs(t - e - m).max() / np.abs(t).max())}


def dft_force_stress(lib, a=20, single=False, ky=0.15, spectra=False):
    """Relative departures of dft_force from the stress of travelling waves (module
    docstring)."""
    s = structure(lib, dim=1, size=(0, 0, 8.0), a=a, pml=[(Z, side, 1.0) for side in (High, Low)], single=single)
    f = fields(s)
    freqs = np.linspace(0.3, 0.5, 5)
    f.add_volume_source(Ex, gaussian_src_time(0.4, 0.3), (0, 0, 2.0), (0, 0, 2.0))
    fl = f.add_dft_flux(Z, (0, 0, 3.5), (0, 0, 3.5), freqs, decimation_factor=1)
    fz = f.add_dft_force([((0, 0, 3.5), (0, 0, 3.5), 104)], freqs, decimation_factor=1)
    f.step_n(int(round(60.0 / s.dt)))
    flux1, force1 = np.asarray(fl.flux()), fz.force()
    s = structure(lib, dim=2, size=(8.0, 1.0), a=a, pml=[(X, side, 1.0) for side in (High, Low)], single=single)
    f = fields(s)
    f.use_bloch_dir(Y, ky)
    freqs = np.linspace(0.4, 0.6, 5)
    f.add_volume_source(Ez, gaussian_src_time(0.5, 0.3), (2.0, 0, 0), (2.0, 1.0, 0),
                        amp_func=lambda x, y, z: np.exp(2j * np.pi * ky * y))
    fl = f.add_dft_flux(X, (3.5, 0, 0), (3.5, 1.0, 0), freqs, decimation_factor=1)
    fy = f.add_dft_force([((3.5, 0, 0), (3.5, 1.0, 0), 101)], freqs, decimation_factor=1)
    fx = f.add_dft_force([((3.5, 0, 0), (3.5, 1.0, 0), 100)], freqs, decimation_factor=1)
    f.step_n(int(round(60.0 / s.dt)))
    flux2, force_y, force_x = np.asarray(fl.flux()), fy.force(), fx.force()
    if spectra:
        return force1, force_y, force_x
    sin = ky / freqs
    return {"diag_1d": float(np.abs(force1 + flux1).max() / np.abs(flux1).max()),
            "offdiag": float(np.abs(force_y + sin * flux2).max() / np.abs(flux2).max()),
            "diag_oblique": float(np.abs(force_x + np.sqrt(1 - sin ** 2) * flux2).max() / np.abs(flux2).max())}
