"""Each closed-form check of tests/closed_forms.py detects the mistake it is there to catch:
the oracle built with one planted mistake (-DMUT_<name>, oracle/Makefile `mutants`) fails
the check, while the unmodified oracle passes it (test_oracle_closed_forms.py).  The mistakes
are the plausible ones a restatement of the reference could make: a dropped DFT scale, a
flipped DFT phase sign, H transformed at the E time, a current injected without dt, flux
weights dropped, a Drude pole with a restoring term, pole damping off by 2 pi, a magnetic
wall at the wrong plane, a conductive update with dt for dt / 2, an electric conductivity
not scaled by chi1inv, a current in a conductive medium not scaled by condinv, the conductive
uPML branch without its (kap - sig) factor, a uPML branch taken over the whole chunk instead
of its region, and a DFT array that keeps the integration weights or the stored weight, sums
the planes around an off-grid monitor unweighted, or places a periodic piece without its
shift; a time-domain flux that adds its second pair, reads fields off the centred grid, or
keeps only the flux after the E update; a uPML term whose sigma ignores its profile's integral or
whose kappa uses the wrong one; an amplitude array interpolated without its half-cell offset, a
source point moved to its periodic owner with the phase instead of its conjugate."""

import ctypes
import os
import subprocess

import pytest

import test_oracle_closed_forms as T
from meep_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECKS = {
    "DFT_SCALE": T.check_line_source,
    "DFT_SIGN": T.check_line_source,
    "DFT_HTIME": T.check_line_source,
    "SRC_DT": T.check_line_source,
    "FLUX_WEIGHT": T.check_plane_wave,
    "DRUDE": lambda lib: T.check_pole(lib, True),
    "GAMMA": lambda lib: T.check_pole(lib, False),
    "MAGNETIC": T.check_cavity,
    "COND_HALF": T.check_conductive_decay,
    "COND_CHI": T.check_conductive_decay,
    "COND_SRC": T.check_conductive_line,
    "COND_PML": T.check_conductive_pml_limit,
    "PML_CHUNKWIDE": T.check_pml_locality,
    "DFTARR_WEIGHT": T.check_dft_array,
    "DFTARR_SIGN": T.check_dft_array,
    "DFTARR_INTERP": T.check_dft_array,
    "DFTARR_SHIFT": T.check_dft_array,
    "FLUXT_SIGN": T.check_flux_time_domain,
    "FLUXT_CENTRE": T.check_flux_time_domain,
    "FLUXT_HALF": T.check_flux_time_domain,
    "PML_INTEGRAL": T.check_pml_reflection,
    "PML_STRETCH": T.check_pml_stretch,
    "AMPARR_OFFSET": T.check_source_array,
    "FIXBND_PHASE": T.check_srcdata,
    "N2F_SIGN": T.check_n2f_equivalence,
    "N2F_PHASE": T.check_n2f_periodic,
    "HPOLE_SUB": lambda lib: T.check_pole(lib, False, magnetic=True),
    "SCALE_DFTS": T.check_dft_reload,
    "DFTENERGY_WEIGHT": T.check_dft_energy,
    "FORCE_DIAG": T.check_dft_force,
    "DFTARR_SQRT": T.check_force_dft_array,
}


@pytest.fixture(scope="module")
def mutants():
    subprocess.check_call(["make", "-s", "-j8", "-C", os.path.join(ROOT, "oracle"), "mutants"])
    return {m: capi.bind(ctypes.CDLL(os.path.join(ROOT, "oracle", "_build", f"libmeep_oracle_mut_{m}.so")))
            for m in CHECKS}


@pytest.mark.parametrize("mutant", sorted(CHECKS))
def test_planted_mistake_is_detected(mutants, mutant):
    with pytest.raises(AssertionError):
        CHECKS[mutant](mutants[mutant])
