"""Properties the mathematics fixes, checked on the oracle where the reference's own golden
values do not reach: its DFT accumulation, flux and field-energy code.

* Energy conservation: in a closed metallic cavity without loss, fields::field_energy is
  constant once the source is off (to the leapfrog scheme's O((omega dt)^2) oscillation, far
  below any dropped term or wrong sign, which changes it at O(1)).
* Lossless 1D transport: a pulse crosses two flux planes unchanged, so their flux spectra agree
  (a wrong weight, sign or E/H time offset in the flux breaks this).
* Source spectrum: a current sheet in 1D vacuum radiates a flux proportional to |J(omega)|^2, so
  the ratio of the flux spectrum to |J(omega)|^2 (J's discrete Fourier transform, computed
  here from the Gaussian's closed form) is flat across the band.
"""

import math

import numpy as np
import pytest

from meep_amd.capi import (Ex, Ez, Z, fields, gaussian_src_time, linspace, pml, structure, volone,
                           voltwo)


def _cavity_energy_spread(lib, a):
    s = structure(lib, eps=lambda x, y, z: 2.0 if x > 1.0 else 1.0, **voltwo(2.0, 2.0, a))
    f = fields(s)
    f.use_real_fields()
    f.add_volume_source(Ez, gaussian_src_time(0.3, 0.2), (0.73, 0.91, 0), (0.73, 0.91, 0))
    while f.time() < f.last_source_time() + 1.0:
        f.step()
    e = []
    for _ in range(24):
        f.step_n(int(round(0.25 / s.dt)))
        e.append(f.field_energy())
    e = np.array(e)
    return (e.max() - e.min()) / e.mean(), e.mean()


def test_cavity_energy_is_conserved(oracle_lib):
    # the reference's quadrature of the energy (loop_in_chunks weights on each component's own
    # grid) is first-order accurate, so the energy of the sealed cavity wanders by O(dx) and
    # halves when the resolution doubles; a dropped magnetic or electric part would make it
    # swing by O(1) as energy moves between E and H
    v10, e10 = _cavity_energy_spread(oracle_lib, 10)
    v20, e20 = _cavity_energy_spread(oracle_lib, 20)
    assert v10 < 0.05 and v20 < 0.03, (v10, v20)
    assert 0.35 < v20 / v10 < 0.65, (v10, v20)
    assert abs(e20 - e10) / e20 < 0.1


def _line(lib, zsrc, planes, nfreq=20, L=24.0, a=20):
    s = structure(lib, pml=pml(2.0), **volone(L, a))
    f = fields(s)
    f.use_real_fields()
    f.add_volume_source(Ex, gaussian_src_time(1.0, 0.5), (0, 0, zsrc), (0, 0, zsrc))
    fl = [f.add_dft_flux(Z, (0, 0, zp), (0, 0, zp), linspace(0.6, 1.4, nfreq), decimation_factor=1)
          for zp in planes]
    return s, f, fl


def test_flux_is_the_same_through_two_planes_in_1d(oracle_lib):
    s, f, (f1, f2) = _line(oracle_lib, 5.0, (10.0, 16.0))
    f.step_n(int(70.0 / s.dt))
    a, b = f1.flux(), f2.flux()
    assert a.min() > 0
    assert np.abs(a - b).max() <= 1e-4 * a.max(), np.abs(a - b).max() / a.max()


def test_flux_spectrum_follows_the_source_spectrum(oracle_lib):
    s, f, (fl,) = _line(oracle_lib, 5.0, (10.0,))
    dt = s.dt
    n = int(70.0 / dt)
    f.step_n(n)
    flux = fl.flux()
    # the Gaussian current of python/source.py semantics: width 1/df, peak at 5 widths, the
    # current being the discrete time derivative of the dipole exp(-(t-t0)^2/2w^2) e^{-iwt}/(-iw)
    w, t0, fc = 2.0, 10.0, 1.0
    def dipole(t):
        return np.exp(-((t - t0) ** 2) / (2 * w * w)) * np.exp(-2j * math.pi * fc * t) / (-2j * math.pi * fc)
    t = (np.arange(n) + 0.5) * dt
    cur = (dipole(t + dt) - dipole(t)) / dt
    freqs = linspace(0.6, 1.4, 20)
    J = np.array([np.sum(cur.real * np.exp(2j * math.pi * fr * t)) * dt for fr in freqs])
    # over the band where the pulse has its energy (|J|^2 above 1e-2 of its peak; outside it
    # the spectrum is 1e-6 of the peak and truncation of the pulse dominates)
    band = np.abs(J) ** 2 > 1e-2 * np.max(np.abs(J) ** 2)
    assert band.sum() >= 8
    ratio = flux[band] / np.abs(J[band]) ** 2
    assert np.all(ratio > 0)
    assert (ratio.max() - ratio.min()) / ratio.mean() < 2e-2, ratio
