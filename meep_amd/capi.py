"""ctypes mirror of the C-ABI in include/meep_amd.h.

The classes below follow the reference's C++ API names (``structure``, ``fields``,
``dft_flux``; src/meep.hpp) so scripts read like the reference's own tests
(tests/known_results.cpp).  Each class takes the ``ctypes.CDLL`` it drives; the package
entry point :func:`meep_amd.load_library` returns the MI355X engine library.
"""

import ctypes as C

import numpy as np

# reference enums (src/meep/vec.hpp:31-79, src/meep.hpp:1594)
Ex, Ey, Ez, Hx, Hy, Hz = 0, 1, 4, 5, 6, 9
E_stuff, H_stuff = 0, 1
Dx, Dy, Dz, Bx, By, Bz = 10, 11, 14, 15, 16, 19
X, Y, Z = 0, 1, 2
High, Low = 0, 1
Periodic, Metallic, Magnetic, NoBoundary = 0, 1, 2, 3
ARR_F, ARR_F_U, ARR_F_W, ARR_P, ARR_P_PREV, ARR_F_COND = 0, 1, 2, 3, 4, 5
K_STEP_BH, K_STEP_DE, K_HALO, K_DFT, K_SRC, K_BH_INTERIOR, K_DE_INTERIOR, K_FARFIELD = 0, 1, 2, 3, 4, 5, 6, 7
DFT_REALNUM, DFT_F64, DFT_KAHAN32 = 0, 1, 2

component_name = {Ex: "ex", Ey: "ey", Ez: "ez", Hx: "hx", Hy: "hy", Hz: "hz",
                  Dx: "dx", Dy: "dy", Dz: "dz", Bx: "bx", By: "by", Bz: "bz"}

SCALAR_FUNC = C.CFUNCTYPE(C.c_double, C.POINTER(C.c_double), C.c_void_p)
COMPLEX_FUNC = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.c_void_p,
                           C.POINTER(C.c_double), C.POINTER(C.c_double))


class SrcTime(C.Structure):
    """mp_src_time: gaussian_src_time / continuous_src_time (src/sources.cpp:67-145)."""
    _fields_ = [("kind", C.c_int), ("freq", C.c_double), ("width", C.c_double),
                ("start_time", C.c_double), ("end_time", C.c_double),
                ("slowness", C.c_double), ("is_integrated", C.c_int)]


def gaussian_src_time(freq, fwidth, start_time=0.0, cutoff=5.0, is_integrated=False):
    """Python GaussianSource semantics (python/source.py:326-336)."""
    width = 1.0 / fwidth
    return SrcTime(0, freq, width, start_time, start_time + 2 * width * cutoff, 3.0,
                   int(is_integrated))


def continuous_src_time(freq, width=0.0, start_time=0.0, end_time=1e20, slowness=3.0,
                        is_integrated=False):
    return SrcTime(1, freq, width, start_time, end_time, slowness, int(is_integrated))


_SIGNATURES = {
    "mp_last_error": (C.c_char_p, []),
    "mp_engine_name": (C.c_char_p, []),
    "mp_structure_new": (C.c_void_p, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                      SCALAR_FUNC, C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                      C.POINTER(C.c_int), C.POINTER(C.c_double),
                                      C.POINTER(C.c_double), C.c_double, C.c_int]),
    "mp_structure_new_pml": (C.c_void_p, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, SCALAR_FUNC,
                                          C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                          C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                          C.POINTER(SCALAR_FUNC), C.POINTER(C.c_void_p), C.POINTER(C.c_double),
                                          C.POINTER(C.c_double), C.c_double, C.c_int]),
    "mp_structure_delete": (None, [C.c_void_p]),
    "mp_structure_ntot": (C.c_size_t, [C.c_void_p]),
    "mp_structure_num": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "mp_structure_dt": (C.c_double, [C.c_void_p]),
    "mp_structure_set_chi1inv": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "mp_structure_get_chi1inv": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "mp_structure_add_susceptibility": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_int,
                                                  SCALAR_FUNC, C.c_void_p]),
    "mp_structure_add_susceptibility_ft": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int,
                                                     SCALAR_FUNC, C.c_void_p]),
    "mp_structure_set_sigma": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "mp_structure_set_conductivity": (C.c_int, [C.c_void_p, C.c_int, SCALAR_FUNC, C.c_void_p]),
    "mp_fields_new": (C.c_void_p, [C.c_void_p]),
    "mp_fields_delete": (None, [C.c_void_p]),
    "mp_fields_use_real_fields": (C.c_int, [C.c_void_p]),
    "mp_fields_use_bloch": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "mp_fields_use_bloch_dir": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double]),
    "mp_fields_set_boundary": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "mp_fields_add_point_source_legacy": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double,
                                                    C.c_double, C.c_double,
                                                    C.POINTER(C.c_double), C.c_double,
                                                    C.c_double, C.c_int]),
    "mp_fields_add_volume_source": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(SrcTime),
                                              C.POINTER(C.c_double), C.POINTER(C.c_double),
                                              C.c_double, C.c_double, COMPLEX_FUNC,
                                              C.c_void_p]),
    "mp_fields_step": (C.c_int, [C.c_void_p]),
    "mp_fields_step_n": (C.c_int, [C.c_void_p, C.c_long]),
    "mp_fields_step_n_noalloc": (C.c_int, [C.c_void_p, C.c_long]),
    "mp_fields_time": (C.c_double, [C.c_void_p]),
    "mp_fields_round_time": (C.c_double, [C.c_void_p]),
    "mp_fields_t": (C.c_long, [C.c_void_p]),
    "mp_fields_dt": (C.c_double, [C.c_void_p]),
    "mp_fields_last_source_time": (C.c_double, [C.c_void_p]),
    "mp_fields_get_field": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "mp_fields_get_array": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.POINTER(C.c_double)]),
    "mp_fields_set_array": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.POINTER(C.c_double)]),
    "mp_fields_have_component": (C.c_int, [C.c_void_p, C.c_int]),
    "mp_fields_get_times": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "mp_fields_set_dft_accumulator": (C.c_int, [C.c_void_p, C.c_int]),
    "mp_fields_set_kernel_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "mp_fields_get_kernel_timing": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double),
                                              C.POINTER(C.c_long)]),
    "mp_fields_reset_kernel_timing": (C.c_int, [C.c_void_p]),
    "mp_fields_add_dft_flux": (C.c_void_p, [C.c_void_p, C.c_int, C.POINTER(C.c_double),
                                            C.POINTER(C.c_double), C.POINTER(C.c_double),
                                            C.c_int, C.c_int, C.c_int, C.c_int]),
    "mp_dft_flux_delete": (None, [C.c_void_p]),
    "mp_dft_flux_flux": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "mp_dft_flux_num_chunks": (C.c_int, [C.c_void_p, C.c_int]),
    "mp_dft_flux_chunk_points": (C.c_long, [C.c_void_p, C.c_int, C.c_int]),
    "mp_dft_flux_decimation": (C.c_int, [C.c_void_p]),
    "mp_dft_flux_get_dft": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "mp_fields_field_energy": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "mp_adjoint_new": (C.c_void_p, [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_int),
                                    C.POINTER(C.c_int)]),
    "mp_adjoint_delete": (None, [C.c_void_p]),
    "mp_batch_new": (C.c_void_p, [C.POINTER(C.c_void_p), C.c_int]),
    "mp_batch_delete": (None, [C.c_void_p]),
    "mp_batch_step_n": (C.c_int, [C.c_void_p, C.c_long]),
    "mp_batch_set_options": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "mp_batch_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "mp_adjoint_objective": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "mp_adjoint_gradient": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "mp_fields_add_dft_fields": (C.c_void_p, [C.c_void_p, C.POINTER(C.c_int), C.c_int,
                                              C.POINTER(C.c_double), C.POINTER(C.c_double),
                                              C.POINTER(C.c_double), C.c_int, C.c_int, C.c_int]),
    "mp_dft_fields_delete": (None, [C.c_void_p]),
    "mp_dft_fields_num_chunks": (C.c_int, [C.c_void_p]),
    "mp_dft_fields_chunk_points": (C.c_long, [C.c_void_p, C.c_int]),
    "mp_dft_fields_chunk_component": (C.c_int, [C.c_void_p, C.c_int]),
    "mp_dft_fields_get_dft": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "mp_fields_add_volume_source_array": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(SrcTime), C.POINTER(C.c_double),
                                                    C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_size_t,
                                                    C.c_size_t, C.c_size_t, C.c_double, C.c_double]),
    "mp_fields_add_srcdata": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(SrcTime), C.c_long, C.POINTER(C.c_long),
                                        C.POINTER(C.c_double), C.c_int]),
    "mp_fields_require_source_components": (C.c_int, [C.c_void_p]),
    "mp_fields_add_flux_vol": (C.c_void_p, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "mp_flux_vol_delete": (None, [C.c_void_p]),
    "mp_flux_vol_flux": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "mp_fields_flux_in_box": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                        C.POINTER(C.c_double)]),
    "mp_fields_flux_in_box_wrongH": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double),
                                               C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "mp_fields_energy_in_box": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                          C.POINTER(C.c_double)]),
    "mp_dft_flux_get_dft_array": (C.c_long, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int),
                                             C.POINTER(C.c_size_t), C.POINTER(C.c_double)]),
    "mp_dft_fields_get_dft_array": (C.c_long, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int),
                                               C.POINTER(C.c_size_t), C.POINTER(C.c_double)]),
    "mp_fields_add_dft_near2far": (C.c_void_p, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                                C.POINTER(C.c_double), C.POINTER(C.c_double),
                                                C.POINTER(C.c_double), C.c_int, C.c_int, C.c_int]),
    "mp_near2far_delete": (None, [C.c_void_p]),
    "mp_near2far_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)]),
    "mp_near2far_farfield": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "mp_near2far_get_farfields_array": (C.c_long, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                                   C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_size_t),
                                                   C.POINTER(C.c_double)]),
    "mp_near2far_flux": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                   C.c_double, C.POINTER(C.c_double)]),
    "mp_near2far_get_dft_array": (C.c_long, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int),
                                             C.POINTER(C.c_size_t), C.POINTER(C.c_double)]),
    "mp_dft_flux_set_dft": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "mp_dft_flux_scale_dfts": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "mp_near2far_num_chunks": (C.c_int, [C.c_void_p]),
    "mp_near2far_chunk_points": (C.c_long, [C.c_void_p, C.c_int]),
    "mp_near2far_get_dft": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "mp_near2far_set_dft": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "mp_near2far_scale_dfts": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "mp_fields_add_dft_energy": (C.c_void_p, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                              C.POINTER(C.c_double), C.c_int, C.c_int]),
    "mp_dft_energy_delete": (None, [C.c_void_p]),
    "mp_dft_energy_get": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "mp_fields_add_dft_force": (C.c_void_p, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                             C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                             C.c_int, C.c_int]),
    "mp_dft_force_delete": (None, [C.c_void_p]),
    "mp_dft_force_force": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "mp_dft_force_scale_dfts": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "mp_dft_energy_scale_dfts": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "mp_dft_force_get_dft_array": (C.c_long, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int),
                                              C.POINTER(C.c_size_t), C.POINTER(C.c_double)]),
    "mp_near2far_green": (C.c_int, [C.c_int, C.POINTER(C.c_double), C.c_double, C.c_double, C.c_double,
                                    C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]),
}

EXPORTED_SYMBOLS = tuple(_SIGNATURES)


def bind(lib):
    """Attach argtypes/restype to every entry point of ``lib``; return ``lib``."""
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


class MeepError(RuntimeError):
    """Raised where the reference calls meep::abort (src/mympi.cpp:238-256)."""


def _check(lib, rc):
    if rc is not None and rc < 0:
        raise MeepError(lib.mp_last_error().decode())
    return rc


def _dvec(vals):
    arr = (C.c_double * 3)(*(list(vals) + [0.0] * (3 - len(vals))))
    return arr


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class structure:
    """meep::structure (src/meep.hpp:815-930) with eps given as a function of (x, y, z)."""

    def __init__(self, lib, dim, size, a, eps=None, pml=(), courant=0.5, single=False):
        self.lib = lib
        self._keep = []
        size = list(size) + [0.0] * (3 - len(size))
        if dim == 1:
            size = [0.0, 0.0, size[0]] if size[2] == 0 else size
        eps_cb = SCALAR_FUNC(0) if eps is None else SCALAR_FUNC(
            lambda r, _u: float(eps(r[0], r[1], r[2])))
        self._keep.append(eps_cb)
        # a term is (direction, side, thickness[, R[, mean_stretch[, profile[, integral,
        # integral_u]]]]): boundary_region's uPML parameters (src/meep.hpp:665-672); profile
        # is a function of u in [0, 1] (None: u^2), the integrals default to 1/3 and 1/4
        n = len(pml)
        dirs = (C.c_int * max(n, 1))(*[p[0] for p in pml])
        sides = (C.c_int * max(n, 1))(*[p[1] for p in pml])
        thick = (C.c_double * max(n, 1))(*[p[2] for p in pml])
        R = (C.c_double * max(n, 1))(*[(p[3] if len(p) > 3 else 1e-15) for p in pml])
        stretch = (C.c_double * max(n, 1))(*[(p[4] if len(p) > 4 else 1.0) for p in pml])
        profs = (SCALAR_FUNC * max(n, 1))()
        for i, p in enumerate(pml):
            if len(p) > 5 and p[5] is not None:
                cb = SCALAR_FUNC(lambda r, _u, fn=p[5]: float(fn(r[0])))
                self._keep.append(cb)
                profs[i] = cb
        integ = (C.c_double * max(n, 1))(*[(p[6] if len(p) > 6 else 0.0) for p in pml])
        integ_u = (C.c_double * max(n, 1))(*[(p[7] if len(p) > 7 else 0.0) for p in pml])
        self.ptr = lib.mp_structure_new_pml(dim, size[0], size[1], size[2], a, eps_cb, None, n,
                                            dirs, sides, thick, R, stretch, profs, None, integ,
                                            integ_u, courant, int(single))
        if not self.ptr:
            raise MeepError(lib.mp_last_error().decode())
        self.dim = dim
        self.single = single
        self.a = a
        self.pml = [tuple(p) for p in pml]

    @property
    def ntot(self):
        return self.lib.mp_structure_ntot(self.ptr)

    @property
    def num(self):
        n = (C.c_int * 3)()
        self.lib.mp_structure_num(self.ptr, n)
        return tuple(n)

    @property
    def shape(self):
        """Array shape of a field in the reference layout (slowest direction first)."""
        n = self.num
        if self.dim == 1:
            return (n[2] + 1,)
        if self.dim == 2:
            return (n[0] + 1, n[1] + 1)
        return (n[0] + 1, n[1] + 1, n[2] + 1)

    @property
    def dt(self):
        return self.lib.mp_structure_dt(self.ptr)

    def set_chi1inv(self, c, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        assert arr.size == self.ntot
        _check(self.lib, self.lib.mp_structure_set_chi1inv(self.ptr, c, _dptr(arr)))

    def get_chi1inv(self, c):
        out = np.empty(self.ntot)
        _check(self.lib, self.lib.mp_structure_get_chi1inv(self.ptr, c, _dptr(out)))
        return out.reshape(self.shape)

    def add_susceptibility(self, sigma, omega_0, gamma, drude=False, ft=E_stuff):
        """add_susceptibility(sigma, ft, lorentzian_susceptibility(omega_0, gamma, drude)),
        ft E_stuff or H_stuff (a magnetic pole).  sigma None leaves sigma zero everywhere
        (fill it with set_sigma)."""
        cb = SCALAR_FUNC(0) if sigma is None else SCALAR_FUNC(
            lambda r, _u: float(sigma(r[0], r[1], r[2])))
        self._keep.append(cb)
        return _check(self.lib, self.lib.mp_structure_add_susceptibility_ft(
            self.ptr, ft, omega_0, gamma, int(drude), cb, None))

    def set_conductivity(self, c, cond):
        """structure::set_conductivity(c, double C(const vec &)) (src/meep.hpp:855-856): cond is
        a function of (x, y, z); for an E or H component it is multiplied by chi1inv."""
        cb = SCALAR_FUNC(lambda r, _u: float(cond(r[0], r[1], r[2])))
        self._keep.append(cb)
        _check(self.lib, self.lib.mp_structure_set_conductivity(self.ptr, c, cb, None))

    def set_sigma(self, isus, c, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        _check(self.lib, self.lib.mp_structure_set_sigma(self.ptr, isus, c, _dptr(arr)))

    def __del__(self):
        if getattr(self, "ptr", None):
            self.lib.mp_structure_delete(self.ptr)
            self.ptr = None


def _get_dft_array(fn, ptr, lib, single, c, num_freq):
    """The array of fields::get_dft_array as the reference's Python layer returns it
    (python/meep.i:450-490): shape dims, or a 0-d array holding the single value (rank 0) or
    zero (a null array); complex64 for single-precision fields."""
    rank = C.c_int(0)
    dims = (C.c_size_t * 3)()
    n = _check(lib, fn(ptr, c, num_freq, C.byref(rank), dims, None))
    out = np.empty(2 * max(n, 1))
    n = _check(lib, fn(ptr, c, num_freq, C.byref(rank), dims, _dptr(out)))
    dtype = np.complex64 if single else np.complex128
    if n == 0:
        return np.zeros((), dtype=dtype)
    vals = (out[0:2 * n:2] + 1j * out[1:2 * n:2]).astype(dtype)
    if rank.value == 0:
        return vals.reshape(())
    return vals.reshape(tuple(dims[r] for r in range(rank.value)))


class dft_flux:
    """meep::dft_flux (src/meep.hpp:1209-1255)."""

    def __init__(self, fields, ptr, nfreq):
        self.f = fields
        self.lib = fields.lib
        self.ptr = ptr
        self.nfreq = nfreq

    def flux(self):
        out = np.empty(self.nfreq)
        _check(self.lib, self.lib.mp_dft_flux_flux(self.ptr, _dptr(out)))
        return out

    def num_chunks(self, which):
        return self.lib.mp_dft_flux_num_chunks(self.ptr, which)

    @property
    def decimation(self):
        return self.lib.mp_dft_flux_decimation(self.ptr)

    def dft(self, which, i=0):
        n = self.lib.mp_dft_flux_chunk_points(self.ptr, which, i)
        out = np.empty(2 * n * self.nfreq)
        _check(self.lib, self.lib.mp_dft_flux_get_dft(self.ptr, which, i, _dptr(out)))
        return (out[0::2] + 1j * out[1::2]).reshape(n, self.nfreq)

    def set_dft(self, which, i, values):
        """Replace chunk i's accumulators (shape (points, nfreq))."""
        v = np.ascontiguousarray(values, dtype=np.complex128).reshape(-1)
        raw = np.empty(2 * v.size)
        raw[0::2], raw[1::2] = v.real, v.imag
        _check(self.lib, self.lib.mp_dft_flux_set_dft(self.ptr, which, i, _dptr(raw)))

    def scale_dfts(self, scale):
        """dft_flux::scale_dfts (src/dft.cpp:597-600)."""
        scale = complex(scale)
        _check(self.lib, self.lib.mp_dft_flux_scale_dfts(self.ptr, scale.real, scale.imag))

    def get_data(self):
        """The accumulators of the E and H lists (the Python layer's get_flux_data,
        python/simulation.py:3628-3636), per chunk."""
        return ([self.dft(0, i) for i in range(self.num_chunks(0))],
                [self.dft(1, i) for i in range(self.num_chunks(1))])

    def load_data(self, data):
        """load_flux_data (python/simulation.py:3638-3647): replace the accumulators."""
        E, H = data
        if len(E) != self.num_chunks(0) or len(H) != self.num_chunks(1):
            raise MeepError("meep: flux data from a monitor with a different chunk layout")
        for which, lst in ((0, E), (1, H)):
            for i, v in enumerate(lst):
                self.set_dft(which, i, v)

    def load_minus_data(self, data):
        """load_minus_flux_data (python/simulation.py:3651-3659): load, then scale by -1."""
        self.load_data(data)
        self.scale_dfts(-1.0)

    def get_dft_array(self, c, num_freq):
        """fields::get_dft_array(dft_flux, c, num_freq, ...) (src/dft.cpp:1259-1268)."""
        return _get_dft_array(self.lib.mp_dft_flux_get_dft_array, self.ptr, self.lib,
                              self.f.s.single, c, num_freq)


class dft_fields:
    """meep::dft_fields (src/meep.hpp:1420-1441): accumulators per chunk and component."""

    def __init__(self, fields, ptr, nfreq):
        self.f = fields
        self.lib = fields.lib
        self.ptr = ptr
        self.nfreq = nfreq

    def num_chunks(self):
        return self.lib.mp_dft_fields_num_chunks(self.ptr)

    def component(self, i):
        c = self.lib.mp_dft_fields_chunk_component(self.ptr, i)
        if c < 0:
            raise MeepError(self.lib.mp_last_error().decode())
        return c

    def points(self, i):
        """Points of chunk i (interpolation planes of the monitor volume included)."""
        n = self.lib.mp_dft_fields_chunk_points(self.ptr, i)
        if n < 0:
            raise MeepError(self.lib.mp_last_error().decode())
        return n

    def dft(self, i):
        n = self.lib.mp_dft_fields_chunk_points(self.ptr, i)
        out = np.empty(2 * n * self.nfreq)
        _check(self.lib, self.lib.mp_dft_fields_get_dft(self.ptr, i, _dptr(out)))
        return (out[0::2] + 1j * out[1::2]).reshape(n, self.nfreq)

    def get_dft_array(self, c, num_freq):
        """fields::get_dft_array(dft_fields, c, num_freq, ...) (src/dft.cpp:1281-1290)."""
        return _get_dft_array(self.lib.mp_dft_fields_get_dft_array, self.ptr, self.lib,
                              self.f.s.single, c, num_freq)


class flux_vol:
    """meep::flux_vol (src/meep.hpp:2322-2352): the flux of E x H through a volume, taken by
    every step; it stays with its fields."""

    def __init__(self, fields, ptr):
        self.f = fields
        self.lib = fields.lib
        self.ptr = ptr

    def flux(self):
        out = C.c_double()
        _check(self.lib, self.lib.mp_flux_vol_flux(self.ptr, C.byref(out)))
        return out.value

    def __del__(self):
        if getattr(self, "ptr", None):
            self.lib.mp_flux_vol_delete(self.ptr)
            self.ptr = None


class dft_energy:
    """meep::dft_energy (src/meep.hpp:1258-1290): energy-density spectra in volumes."""

    def __init__(self, fields, ptr, nfreq):
        self.f = fields
        self.lib = fields.lib
        self.ptr = ptr
        self.nfreq = nfreq

    def _get(self, which):
        out = np.empty(self.nfreq)
        _check(self.lib, self.lib.mp_dft_energy_get(self.ptr, which, _dptr(out)))
        return out

    def total(self):
        """dft_energy::total (src/dft.cpp:713-723)."""
        return self._get(0)

    def electric(self):
        """dft_energy::electric (src/dft.cpp:681-695)."""
        return self._get(1)

    def magnetic(self):
        """dft_energy::magnetic (src/dft.cpp:697-711)."""
        return self._get(2)

    def scale_dfts(self, scale):
        """dft_energy::scale_dfts (src/dft.cpp:785-790)."""
        scale = complex(scale)
        _check(self.lib, self.lib.mp_dft_energy_scale_dfts(self.ptr, scale.real, scale.imag))


class dft_force:
    """meep::dft_force (src/meep.hpp:1292-1320): the Maxwell-stress force spectrum."""

    def __init__(self, fields, ptr, nfreq):
        self.f = fields
        self.lib = fields.lib
        self.ptr = ptr
        self.nfreq = nfreq

    def force(self):
        """dft_force::force (src/stress.cpp:101-114)."""
        out = np.empty(self.nfreq)
        _check(self.lib, self.lib.mp_dft_force_force(self.ptr, _dptr(out)))
        return out

    def scale_dfts(self, scale):
        """dft_force::scale_dfts (src/stress.cpp:144-148)."""
        scale = complex(scale)
        _check(self.lib, self.lib.mp_dft_force_scale_dfts(self.ptr, scale.real, scale.imag))

    def get_dft_array(self, c, num_freq):
        """fields::get_dft_array(dft_force, c, num_freq, ...) (src/dft.cpp:1270-1280)."""
        return _get_dft_array(self.lib.mp_dft_force_get_dft_array, self.ptr, self.lib,
                              self.f.s.single, c, num_freq)


class near2far:
    """meep::dft_near2far (src/meep.hpp:1455-1520): the tangential fields on closed or open
    surfaces, transformed to the fields anywhere in the surrounding homogeneous medium."""

    def __init__(self, fields, ptr, nfreq):
        self.f = fields
        self.lib = fields.lib
        self.ptr = ptr
        self.nfreq = nfreq
        em = (C.c_double * 2)()
        per = (C.c_int * 4)()
        npts = C.c_long()
        self.lib.mp_near2far_info(ptr, em, per, C.byref(npts))
        self.npoints = npts.value
        self.eps, self.mu = em[0], em[1]
        self.periodic_d = (per[0], per[2])
        self.periodic_n = (per[1], per[3])

    def farfield(self, x):
        """dft_near2far::farfield(x) (src/near2far.cpp:399-408): array (nfreq, 6) of
        Ex, Ey, Ez, Hx, Hy, Hz."""
        out = np.empty(12 * self.nfreq)
        _check(self.lib, self.lib.mp_near2far_farfield(self.ptr, _dvec(x), _dptr(out)))
        return (out[0::2] + 1j * out[1::2]).reshape(self.nfreq, 6)

    def get_farfields_array(self, vmin, vmax, resolution):
        """dft_near2far::get_farfields_array (src/near2far.cpp:411-483): array
        (6, 2, N, nfreq) and the grid dims with singletons dropped."""
        rank = C.c_int(0)
        dims = (C.c_size_t * 3)()
        n = self.lib.mp_near2far_get_farfields_array(self.ptr, _dvec(vmin), _dvec(vmax), resolution,
                                                     C.byref(rank), dims, None)
        n = _check(self.lib, n)
        out = np.empty(12 * n * self.nfreq)
        _check(self.lib, self.lib.mp_near2far_get_farfields_array(self.ptr, _dvec(vmin), _dvec(vmax),
                                                                  resolution, C.byref(rank), dims,
                                                                  _dptr(out)))
        return out.reshape(6, 2, n, self.nfreq), tuple(dims[r] for r in range(rank.value))

    def get_farfields(self, vmin, vmax, resolution):
        """The Python layer's get_farfields (python/simulation.py): a dict of complex arrays
        of shape dims (+ nfreq when nfreq > 1) per component."""
        EH, dims = self.get_farfields_array(vmin, vmax, resolution)
        shape = dims + ((self.nfreq,) if self.nfreq > 1 else ())
        names = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
        return {names[k]: (EH[k, 0] + 1j * EH[k, 1]).reshape(shape) for k in range(6)}

    def flux(self, d, vmin, vmax, resolution):
        """dft_near2far::flux(direction, where, resolution) (src/near2far.cpp:517-565)."""
        out = np.empty(self.nfreq)
        _check(self.lib, self.lib.mp_near2far_flux(self.ptr, d, _dvec(vmin), _dvec(vmax), resolution,
                                                   _dptr(out)))
        return out

    def num_chunks(self):
        return self.lib.mp_near2far_num_chunks(self.ptr)

    def dft(self, i):
        n = self.lib.mp_near2far_chunk_points(self.ptr, i)
        out = np.empty(2 * n * self.nfreq)
        _check(self.lib, self.lib.mp_near2far_get_dft(self.ptr, i, _dptr(out)))
        return (out[0::2] + 1j * out[1::2]).reshape(n, self.nfreq)

    def scale_dfts(self, scale):
        """dft_near2far::scale_dfts (src/near2far.cpp:124-126)."""
        scale = complex(scale)
        _check(self.lib, self.lib.mp_near2far_scale_dfts(self.ptr, scale.real, scale.imag))

    def get_data(self):
        """get_near2far_data (python/simulation.py:3355-3362), per chunk."""
        return [self.dft(i) for i in range(self.num_chunks())]

    def load_data(self, data):
        """load_near2far_data (python/simulation.py:3364-3371)."""
        if len(data) != self.num_chunks():
            raise MeepError("meep: near2far data from a monitor with a different chunk layout")
        for i, vals in enumerate(data):
            v = np.ascontiguousarray(vals, dtype=np.complex128).reshape(-1)
            raw = np.empty(2 * v.size)
            raw[0::2], raw[1::2] = v.real, v.imag
            _check(self.lib, self.lib.mp_near2far_set_dft(self.ptr, i, _dptr(raw)))

    def load_minus_data(self, data):
        """load_minus_near2far_data (python/simulation.py:3373-3380)."""
        self.load_data(data)
        self.scale_dfts(-1.0)

    def get_dft_array(self, c, num_freq):
        """fields::get_dft_array(dft_near2far, c, num_freq, ...) (src/dft.cpp:1282-1290)."""
        return _get_dft_array(self.lib.mp_near2far_get_dft_array, self.ptr, self.lib,
                              self.f.s.single, c, num_freq)


def green(lib, dim, x, freq, eps, mu, x0, c0, f0=1.0):
    """green2d / green3d (src/near2far.cpp:133-267): the six fields at x of a point current
    c0 of amplitude f0 at x0 in the medium eps, mu."""
    out = np.empty(12)
    f = _dvec([complex(f0).real, complex(f0).imag])
    _check(lib, lib.mp_near2far_green(dim, _dvec(x), freq, eps, mu, _dvec(x0), c0, f, _dptr(out)))
    return out[0::2] + 1j * out[1::2]


class adjoint:
    """Reverse-mode gradient of J = sum |a|^2 over a dft_fields monitor's accumulators with
    respect to chi1inv of electric component c on the design box [vmin, vmax]; create it
    before the steps to differentiate (see mp_adjoint_new)."""

    def __init__(self, fields, monitor, c, vmin, vmax, checkpoint_interval=0):
        self.f = fields
        self.lib = fields.lib
        lo, n = (C.c_int * 3)(), (C.c_int * 3)()
        self.ptr = self.lib.mp_adjoint_new(fields.ptr, monitor.ptr, c, _dvec(vmin), _dvec(vmax),
                                           int(checkpoint_interval), lo, n)
        if not self.ptr:
            raise MeepError(self.lib.mp_last_error().decode())
        self.box_lo = tuple(lo)
        self.box_n = tuple(n)
        fields._adjoints.append(self)

    @property
    def shape(self):
        """Design box shape in the reference layout (absent directions dropped)."""
        dim = self.f.s.dim
        dirs = {1: (2,), 2: (0, 1), 3: (0, 1, 2)}[dim]
        return tuple(self.box_n[d] for d in dirs)

    def objective(self):
        J = C.c_double()
        _check(self.lib, self.lib.mp_adjoint_objective(self.ptr, C.byref(J)))
        return J.value

    def gradient(self):
        out = np.empty(int(np.prod(self.box_n)))
        _check(self.lib, self.lib.mp_adjoint_gradient(self.ptr, _dptr(out)))
        return out.reshape(self.shape)

    def close(self):
        if self.ptr:
            self.lib.mp_adjoint_delete(self.ptr)
            self.ptr = None


class batch:
    """Many fields objects of one precision stepped together (see mp_batch_new)."""

    def __init__(self, field_objects, use_graph=True, autotune=True, block=0):
        self.fields = list(field_objects)
        self.lib = self.fields[0].lib
        arr = (C.c_void_p * len(self.fields))(*[f.ptr for f in self.fields])
        self.ptr = self.lib.mp_batch_new(arr, len(self.fields))
        if not self.ptr:
            raise MeepError(self.lib.mp_last_error().decode())
        _check(self.lib, self.lib.mp_batch_set_options(self.ptr, int(use_graph), int(autotune),
                                                       int(block)))

    def step_n(self, n):
        _check(self.lib, self.lib.mp_batch_step_n(self.ptr, int(n)))

    def stats(self):
        b, ms = C.c_int(), C.c_double()
        _check(self.lib, self.lib.mp_batch_stats(self.ptr, C.byref(b), C.byref(ms)))
        return b.value, ms.value

    def __del__(self):
        if getattr(self, "ptr", None):
            self.lib.mp_batch_delete(self.ptr)
            self.ptr = None


class fields:
    """meep::fields (src/meep.hpp:1716-2320)."""

    def __init__(self, s):
        self.lib = s.lib
        self.s = s
        self._keep = []
        self.ptr = self.lib.mp_fields_new(s.ptr)
        if not self.ptr:
            raise MeepError(self.lib.mp_last_error().decode())
        self._fluxes = []
        self._adjoints = []

    def use_real_fields(self):
        _check(self.lib, self.lib.mp_fields_use_real_fields(self.ptr))

    def use_bloch(self, k):
        kk = (C.c_double * 3)(*(list(k) + [0.0] * (3 - len(k))))
        _check(self.lib, self.lib.mp_fields_use_bloch(self.ptr, kk))

    def use_bloch_dir(self, d, k):
        k = complex(k)
        _check(self.lib, self.lib.mp_fields_use_bloch_dir(self.ptr, d, k.real, k.imag))

    def set_boundary(self, side, d, cond):
        _check(self.lib, self.lib.mp_fields_set_boundary(self.ptr, side, d, cond))

    def add_point_source(self, c, freq, width, peaktime, cutoff, p, amp=1.0, is_continuous=0):
        """Deprecated C++ add_point_source (src/sources.cpp:193-213)."""
        amp = complex(amp)
        _check(self.lib, self.lib.mp_fields_add_point_source_legacy(
            self.ptr, c, freq, width, peaktime, cutoff, _dvec(p), amp.real, amp.imag,
            int(is_continuous)))

    def add_volume_source(self, c, src, vmin, vmax, amp=1.0, amp_func=None):
        amp = complex(amp)
        if amp_func is None:
            cb = COMPLEX_FUNC(0)
        else:
            def _cb(r, _u, re, im):
                v = complex(amp_func(r[0], r[1], r[2]))
                re[0] = v.real
                im[0] = v.imag
            cb = COMPLEX_FUNC(_cb)
        self._keep.append(cb)
        _check(self.lib, self.lib.mp_fields_add_volume_source(
            self.ptr, c, C.byref(src), _dvec(vmin), _dvec(vmax), amp.real, amp.imag, cb, None))

    def add_volume_source_array(self, c, src, vmin, vmax, arr, amp=1.0):
        """add_volume_source(c, src, where, complex<double> *arr, dim1, dim2, dim3, amp)
        (src/sources.cpp:394-417): arr, up to three dimensions, is interpolated over the
        volume."""
        amp = complex(amp)
        a = np.asarray(arr, dtype=np.complex128)
        shape = list(a.shape) + [1] * (3 - a.ndim)
        flat = np.ascontiguousarray(np.stack([a.real.ravel(), a.imag.ravel()], axis=1).ravel())
        _check(self.lib, self.lib.mp_fields_add_volume_source_array(
            self.ptr, c, C.byref(src), _dvec(vmin), _dvec(vmax), _dptr(flat), shape[0], shape[1], shape[2],
            amp.real, amp.imag))

    def add_srcdata(self, c, src, index, amps, needs_boundary_fix=False):
        """fields::add_srcdata (src/sources.cpp:314-334): a source on the given points of c's
        array (reference-layout indices) with the given complex amplitudes."""
        idx = np.ascontiguousarray(index, dtype=np.int64)
        a = np.asarray(amps, dtype=np.complex128).ravel()
        flat = np.ascontiguousarray(np.stack([a.real, a.imag], axis=1).ravel())
        _check(self.lib, self.lib.mp_fields_add_srcdata(
            self.ptr, c, C.byref(src), len(idx), idx.ctypes.data_as(C.POINTER(C.c_long)), _dptr(flat),
            int(needs_boundary_fix)))

    def require_source_components(self):
        """fields::require_source_components (src/fields.cpp:508-530), with fix_boundary_sources."""
        _check(self.lib, self.lib.mp_fields_require_source_components(self.ptr))

    def add_point_source_src(self, c, src, p, amp=1.0):
        self.add_volume_source(c, src, p, p, amp)

    def step(self):
        _check(self.lib, self.lib.mp_fields_step(self.ptr))

    def step_n(self, n):
        _check(self.lib, self.lib.mp_fields_step_n(self.ptr, int(n)))

    def step_n_noalloc(self, n):
        """n steps with no device allocation (mp_fields_step_n_noalloc): the fields must have
        been stepped by step_n since their last change."""
        _check(self.lib, self.lib.mp_fields_step_n_noalloc(self.ptr, int(n)))

    def time(self):
        return self.lib.mp_fields_time(self.ptr)

    def round_time(self):
        return self.lib.mp_fields_round_time(self.ptr)

    @property
    def t(self):
        return self.lib.mp_fields_t(self.ptr)

    @property
    def dt(self):
        return self.lib.mp_fields_dt(self.ptr)

    def last_source_time(self):
        return self.lib.mp_fields_last_source_time(self.ptr)

    def field_energy(self):
        """fields::field_energy() (src/energy_and_flux.cpp:48)."""
        e = C.c_double()
        _check(self.lib, self.lib.mp_fields_field_energy(self.ptr, C.byref(e)))
        return e.value

    def _energy_in_box(self, which, vmin, vmax):
        e = C.c_double()
        _check(self.lib, self.lib.mp_fields_energy_in_box(self.ptr, which, _dvec(vmin), _dvec(vmax), C.byref(e)))
        return e.value

    def field_energy_in_box(self, vmin, vmax):
        """fields::field_energy_in_box(const volume &) (src/energy_and_flux.cpp:54-59)."""
        return self._energy_in_box(0, vmin, vmax)

    def electric_energy_in_box(self, vmin, vmax):
        """fields::electric_energy_in_box (src/energy_and_flux.cpp:84-88)."""
        return self._energy_in_box(1, vmin, vmax)

    def magnetic_energy_in_box(self, vmin, vmax):
        """fields::magnetic_energy_in_box (src/energy_and_flux.cpp:90-95)."""
        return self._energy_in_box(2, vmin, vmax)

    def add_flux_vol(self, d, vmin, vmax):
        """fields::add_flux_vol(direction, volume) (src/energy_and_flux.cpp:230-235)."""
        ptr = self.lib.mp_fields_add_flux_vol(self.ptr, d, _dvec(vmin), _dvec(vmax))
        if not ptr:
            raise MeepError(self.lib.mp_last_error().decode())
        return flux_vol(self, ptr)

    def add_flux_plane(self, vmin, vmax):
        """fields::add_flux_plane(p1, p2) (src/energy_and_flux.cpp:238-245): the direction is
        the volume's normal."""
        return self.add_flux_vol(-1, vmin, vmax)

    def flux_in_box(self, d, vmin, vmax):
        """fields::flux_in_box(direction, volume) (src/energy_and_flux.cpp:223-228)."""
        out = C.c_double()
        _check(self.lib, self.lib.mp_fields_flux_in_box(self.ptr, d, _dvec(vmin), _dvec(vmax), C.byref(out)))
        return out.value

    def flux_in_box_wrongH(self, d, vmin, vmax):
        """fields::flux_in_box_wrongH(direction, volume) (src/energy_and_flux.cpp:186-221)."""
        out = C.c_double()
        _check(self.lib, self.lib.mp_fields_flux_in_box_wrongH(self.ptr, d, _dvec(vmin), _dvec(vmax),
                                                                C.byref(out)))
        return out.value

    def get_field(self, c, p):
        re, im = C.c_double(), C.c_double()
        _check(self.lib, self.lib.mp_fields_get_field(self.ptr, c, _dvec(p), C.byref(re),
                                                      C.byref(im)))
        return complex(re.value, im.value)

    def have_component(self, c):
        return bool(self.lib.mp_fields_have_component(self.ptr, c))

    def get_array(self, c, cmp=0, kind=ARR_F, ipol=0):
        out = np.empty(self.s.ntot)
        rc = _check(self.lib, self.lib.mp_fields_get_array(self.ptr, c, cmp, kind, ipol,
                                                           _dptr(out)))
        return out.reshape(self.s.shape), rc == 0

    def set_array(self, c, cmp, kind, arr, ipol=0):
        arr = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        _check(self.lib, self.lib.mp_fields_set_array(self.ptr, c, cmp, kind, ipol, _dptr(arr)))

    def times(self):
        t = (C.c_double * 7)()
        self.lib.mp_fields_get_times(self.ptr, t)
        return list(t)

    def set_kernel_timing(self, on=True):
        _check(self.lib, self.lib.mp_fields_set_kernel_timing(self.ptr, int(on)))

    def kernel_timing(self, which):
        """(total milliseconds, launches) of one kernel class since the last reset."""
        ms, n = C.c_double(), C.c_long()
        _check(self.lib, self.lib.mp_fields_get_kernel_timing(self.ptr, which, C.byref(ms),
                                                              C.byref(n)))
        return ms.value, n.value

    def reset_kernel_timing(self):
        _check(self.lib, self.lib.mp_fields_reset_kernel_timing(self.ptr))

    def set_dft_accumulator(self, kind):
        _check(self.lib, self.lib.mp_fields_set_dft_accumulator(self.ptr, kind))

    def add_dft_flux(self, d, vmin, vmax, freqs, decimation_factor=0, use_symmetry=True,
                     centered_grid=True):
        """fields::add_dft_flux(direction, volume, freq, use_symmetry, centered_grid,
        decimation_factor) (src/meep.hpp:2051-2059)."""
        freqs = np.ascontiguousarray(freqs, dtype=np.float64)
        ptr = self.lib.mp_fields_add_dft_flux(self.ptr, d, _dvec(vmin), _dvec(vmax),
                                              _dptr(freqs), len(freqs), int(use_symmetry),
                                              int(centered_grid), decimation_factor)
        if not ptr:
            raise MeepError(self.lib.mp_last_error().decode())
        fl = dft_flux(self, ptr, len(freqs))
        self._fluxes.append(fl)
        return fl

    def add_dft_fields(self, components, vmin, vmax, freqs, decimation_factor=0,
                       use_centered_grid=True):
        """fields::add_dft_fields(components, num_components, volume, freq,
        use_centered_grid, decimation_factor) (src/meep.hpp:2085-2093)."""
        freqs = np.ascontiguousarray(freqs, dtype=np.float64)
        comps = (C.c_int * len(components))(*components)
        ptr = self.lib.mp_fields_add_dft_fields(self.ptr, comps, len(components), _dvec(vmin),
                                                _dvec(vmax), _dptr(freqs), len(freqs),
                                                int(use_centered_grid), decimation_factor)
        if not ptr:
            raise MeepError(self.lib.mp_last_error().decode())
        df = dft_fields(self, ptr, len(freqs))
        self._fluxes.append(df)
        return df

    def add_dft_energy(self, volumes, freqs, decimation_factor=0):
        """fields::add_dft_energy(volume_list, freq, Nfreq, decimation_factor)
        (src/dft.cpp:725-750); volumes: (vmin, vmax) pairs."""
        freqs = np.ascontiguousarray(freqs, dtype=np.float64)
        vmins = np.ascontiguousarray([[float(v) for v in (list(lo) + [0.0] * 3)[:3]] for lo, _ in volumes])
        vmaxs = np.ascontiguousarray([[float(v) for v in (list(hi) + [0.0] * 3)[:3]] for _, hi in volumes])
        ptr = self.lib.mp_fields_add_dft_energy(self.ptr, len(volumes), _dptr(vmins), _dptr(vmaxs), _dptr(freqs),
                                                len(freqs), decimation_factor)
        if not ptr:
            raise MeepError(self.lib.mp_last_error().decode())
        en = dft_energy(self, ptr, len(freqs))
        self._fluxes.append(en)
        return en

    def add_dft_force(self, volumes, freqs, decimation_factor=0):
        """fields::add_dft_force(volume_list, freq, Nfreq, decimation_factor)
        (src/stress.cpp:153-198); volumes: (vmin, vmax, force component[, weight]) with the
        component an E component or Sx = 100, Sy = 101, Sz = 104."""
        freqs = np.ascontiguousarray(freqs, dtype=np.float64)
        n = len(volumes)
        comps = (C.c_int * n)(*[int(v[2]) for v in volumes])
        vmins = np.ascontiguousarray([[float(x) for x in (list(v[0]) + [0.0] * 3)[:3]] for v in volumes])
        vmaxs = np.ascontiguousarray([[float(x) for x in (list(v[1]) + [0.0] * 3)[:3]] for v in volumes])
        w = np.ascontiguousarray([[complex(v[3] if len(v) > 3 else 1.0).real, complex(v[3] if len(v) > 3 else 1.0).imag]
                                  for v in volumes])
        ptr = self.lib.mp_fields_add_dft_force(self.ptr, n, comps, _dptr(vmins), _dptr(vmaxs), _dptr(w), _dptr(freqs),
                                               len(freqs), decimation_factor)
        if not ptr:
            raise MeepError(self.lib.mp_last_error().decode())
        fo = dft_force(self, ptr, len(freqs))
        self._fluxes.append(fo)
        return fo

    def add_dft_near2far(self, surfaces, freqs, decimation_factor=0, nperiods=1):
        """fields::add_dft_near2far(volume_list, freq, Nfreq, decimation_factor, Nperiods)
        (src/near2far.cpp:569-654).  surfaces: (vmin, vmax, weight) or (vmin, vmax, weight,
        normal direction) tuples, the direction defaulting to the volume's normal."""
        freqs = np.ascontiguousarray(freqs, dtype=np.float64)
        n = len(surfaces)
        dirs = (C.c_int * n)(*[int(sf[3]) if len(sf) > 3 else -1 for sf in surfaces])
        vmins = np.ascontiguousarray([[float(v) for v in (list(sf[0]) + [0.0] * 3)[:3]] for sf in surfaces])
        vmaxs = np.ascontiguousarray([[float(v) for v in (list(sf[1]) + [0.0] * 3)[:3]] for sf in surfaces])
        weights = np.ascontiguousarray([[complex(sf[2]).real, complex(sf[2]).imag] for sf in surfaces])
        ptr = self.lib.mp_fields_add_dft_near2far(self.ptr, n, dirs, _dptr(vmins), _dptr(vmaxs), _dptr(weights),
                                                  _dptr(freqs), len(freqs), decimation_factor, nperiods)
        if not ptr:
            raise MeepError(self.lib.mp_last_error().decode())
        nf = near2far(self, ptr, len(freqs))
        self._fluxes.append(nf)
        return nf

    def __del__(self):
        if getattr(self, "ptr", None):
            for adj in self._adjoints:
                adj.close()
            for fl in self._fluxes:
                if fl.ptr:
                    if isinstance(fl, near2far):
                        self.lib.mp_near2far_delete(fl.ptr)
                    elif isinstance(fl, dft_energy):
                        self.lib.mp_dft_energy_delete(fl.ptr)
                    elif isinstance(fl, dft_force):
                        self.lib.mp_dft_force_delete(fl.ptr)
                    elif isinstance(fl, dft_fields):
                        self.lib.mp_dft_fields_delete(fl.ptr)
                    else:
                        self.lib.mp_dft_flux_delete(fl.ptr)
                    fl.ptr = None
            self.lib.mp_fields_delete(self.ptr)
            self.ptr = None


def linspace(fmin, fmax, n):
    """meep::linspace (src/dft.cpp:30-40)."""
    if n <= 1:
        return np.array([(fmin + fmax) * 0.5])
    df = (fmax - fmin) / (n - 1)
    return np.array([fmin + i * df for i in range(n)])


def volone(zsize, a):
    return dict(dim=1, size=(0.0, 0.0, zsize), a=a)


def voltwo(xsize, ysize, a):
    return dict(dim=2, size=(xsize, ysize, 0.0), a=a)


def vol3d(xsize, ysize, zsize, a):
    return dict(dim=3, size=(xsize, ysize, zsize), a=a)


def pml(thickness, direction=None, side=None, R=1e-15, mean_stretch=1.0):
    """boundary_region pml(...) terms (src/structure.cpp:252-273)."""
    dirs = [X, Y, Z] if direction is None else [direction]
    sides = [Low, High] if side is None else [side]
    return [(d, s, thickness, R, mean_stretch) for d in dirs for s in sides]
