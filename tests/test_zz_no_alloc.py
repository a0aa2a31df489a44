"""mp_fields_step_n_noalloc, the entry point that allocates no device memory, against
mp_fields_step_n on identical fields: the fields, auxiliary and polarisation arrays equal
bit for bit (the same kernels run), and the DFT monitors within 1e-12 (FP64) / 1e-6 (FP32)
of their maximum (they accumulate every step instead of in deferred batches of eight, with
the same per-step arithmetic).  Two calls in a row share the library's staging array: the
3D uPML + Lorentz-Drude scenario with a plane source and monitors, then the smaller 2D
waveguide.  Fields never stepped are refused before any step."""

import numpy as np
import pytest

from meep_amd import capi
from meep_amd.capi import dft_fields
from scenarios import box_3d_pml_lorentz, state, waveguide_2d

pytestmark = pytest.mark.gpu


def _monitors(ms):
    out = []
    for m in ms:
        if isinstance(m, dft_fields):
            out += [m.dft(i) for i in range(m.num_chunks())]
        else:
            out += [m.dft(w, i) for w in (0, 1) for i in range(m.num_chunks(w))]
    return out


@pytest.mark.parametrize("single", [False, True], ids=["f64", "f32"])
def test_noalloc_matches_step_n(gpu_lib, single):
    for builder, n in ((box_3d_pml_lorentz, 50), (waveguide_2d, 30)):
        print(f"[noalloc] {builder.__name__} single={single} steps={n}", flush=True)
        (sa, fa, ma), (sb, fb, mb) = builder(gpu_lib, single=single), builder(gpu_lib, single=single)
        fa.step_n(3)
        fb.step_n(3)
        fa.step_n(n)
        fb.step_n_noalloc(n)
        assert fa.t == fb.t == n + 3
        a, b = state(fa, sa, 2), state(fb, sb, 2)
        assert set(a) == set(b)
        for key in sorted(a):
            assert np.array_equal(a[key], b[key]), f"{builder.__name__} {key}"
        tol = 1e-6 if single else 1e-12
        for x, y in zip(_monitors(ma), _monitors(mb)):
            assert x.shape == y.shape
            assert np.abs(x - y).max() <= tol * max(np.abs(x).max(), 1e-30)


def test_noalloc_refuses_unstepped_fields(gpu_lib):
    s, f, _ = waveguide_2d(gpu_lib)
    with pytest.raises(capi.MeepError, match="step_n"):
        f.step_n_noalloc(5)
    assert f.t == 0
