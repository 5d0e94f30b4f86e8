"""Python binding of the MI355X continuum-LDOS library (libcqpi_amd.so).

``AmdQPI`` mirrors the reference's GPU plug-in interface, class GPUQPI
(include/gpuqpi.h:18-32) as implemented by HipQPI (src/qpihipkernel.h:64-88):
same constructor arguments, ``wannierldos(scat, g0)``, ``retrieveResult``,
``printinfo``; errors raise ``CqpiError`` instead of exiting the process
(src/qpihipkernel.cc:372-380).  The whole-chain extension (``set_model``,
``set_impurity``, ``run``) moves the eigensolve, resolvent, k->R transform and
T-matrix (src/qpi.cc:454-486, :557-578, :833-854) onto the device too.

The compute path is the HIP library only; there is no CPU fallback.  Loading
fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import synth

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(os.path.dirname(_HERE), "libcqpi_amd.so")
_LIB = None

PREC_FP64, PREC_FP32, PREC_MIXED = 0, 1, 2
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "mixed": PREC_MIXED}
# cqpi_set_surface modes (green= of the reference driver)
GREENS = {"normal": 0, "surface": 1, "bulk": 2}

# cqpi_flagentry: {i, j, o1, o2 (uint32), factor (double)}, 24 bytes
FLAG_DTYPE = np.dtype([("i", "<u4"), ("j", "<u4"), ("o1", "<u4"), ("o2", "<u4"), ("factor", "<f8")])
STAGES = ("eigensolve", "resolvent+transform_ky", "transform_kx", "tmatrix+bulk", "contraction")

_sz = C.c_size_t
_dp = C.POINTER(C.c_double)
_vp = C.c_void_p

# every symbol include/cqpi_amd.h declares, with its ctypes signature
SIGNATURES = {
    "cqpi_last_error": (C.c_char_p, []),
    "cqpi_version": (C.c_char_p, []),
    "cqpi_device_count": (C.c_int, []),
    "cqpi_create": (C.c_int, [_sz, _sz, _sz, _sz, _sz, _sz, _sz, C.c_int, _dp, C.POINTER(_vp)]),
    "cqpi_wannierldos": (C.c_int, [_vp, _dp, C.POINTER(_vp), _sz]),
    "cqpi_wannierjosephson": (C.c_int, [_vp, _dp, C.POINTER(_vp), _sz, C.c_double, C.c_double]),
    "cqpi_wannierldos_device": (C.c_int, [_vp, _vp, _vp, _sz, _vp]),
    "cqpi_set_josephson": (C.c_int, [_vp, C.c_int, C.c_double, C.c_double]),
    "cqpi_set_surface": (C.c_int, [_vp, C.c_int, C.c_double, _vp, _sz]),
    "cqpi_set_surface_steps": (C.c_int, [_vp, C.c_int]),
    "cqpi_set_pair_symmetry": (C.c_int, [_vp, C.c_int]),
    "cqpi_get_pair_symmetry": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "cqpi_orbital_tile": (C.c_int, [C.c_char_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.c_double, _sz, _sz, _dp]),
    "cqpi_orbital_range": (C.c_double, [C.c_double, C.c_double]),
    "cqpi_xsf_tile": (C.c_int, [C.c_char_p, C.c_double, _sz, _sz, C.c_double, _dp]),
    "cqpi_bands_histogram": (C.c_int, [_vp, _sz, _sz, C.c_double, C.c_double, _sz, _dp]),
    "cqpi_create_list": (C.c_int, [_sz, _sz, _sz, _sz, _sz, _sz, _vp, C.POINTER(_sz), C.POINTER(_vp)]),
    "cqpi_create_spectral": (C.c_int, [_sz, _sz, _sz, _sz, C.c_int, _dp, _sz, _dp, _sz, C.POINTER(_vp)]),
    "cqpi_spf": (C.c_int, [_vp, C.POINTER(_vp), _sz]),
    "cqpi_uspf": (C.c_int, [_vp, C.POINTER(_vp), _sz]),
    "cqpi_set_spectral": (C.c_int, [_vp, C.c_int]),
    "cqpi_dos": (C.c_int, [_vp, _dp, _sz, C.c_double, _dp]),
    "cqpi_wannierldoslist": (C.c_int, [_vp, _dp, C.POINTER(_vp), _sz]),
    "cqpi_make_threshold_list": (C.c_int, [_dp, _sz, _sz, _sz, C.c_double, _sz, C.c_int, _vp, _sz,
                                           C.POINTER(_sz), C.POINTER(_sz)]),
    "cqpi_make_threshold_list_spinarr": (C.c_int, [_dp, _sz, _sz, _sz, C.c_double, C.POINTER(C.c_int), _sz, _vp,
                                                   _sz, C.POINTER(_sz), C.POINTER(_sz)]),
    "cqpi_retrieve": (C.c_int, [_vp, _dp]),
    "cqpi_printinfo": (C.c_int, [_vp, C.c_char_p, _sz]),
    "cqpi_destroy": (None, [_vp]),
    "cqpi_set_precision": (C.c_int, [_vp, C.c_int]),
    "cqpi_set_model": (C.c_int, [_vp, _vp, _sz, C.c_double, C.c_int]),
    "cqpi_set_impurity": (C.c_int, [_vp, _dp]),
    "cqpi_compute_eigen": (C.c_int, [_vp]),
    "cqpi_run": (C.c_int, [_vp, _dp, _sz, C.c_double, _dp]),
    "cqpi_retrieve_layers": (C.c_int, [_vp, _sz, _sz, _dp]),
    "cqpi_device_ldos": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_sz)]),
    "cqpi_window_half": (C.c_int, [_vp, C.POINTER(_sz)]),
    "cqpi_retrieve_g0_window": (C.c_int, [_vp, _sz, _dp]),
    "cqpi_retrieve_tmatrix": (C.c_int, [_vp, _sz, _dp]),
    "cqpi_stage_times": (C.c_int, [_vp, _dp, C.POINTER(C.c_int64)]),
    "cqpi_set_timing": (C.c_int, [_vp, C.c_int]),
    "cqpi_read_hr": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(_vp), C.POINTER(_sz), C.POINTER(_sz),
                               C.c_char_p, _sz]),
    "cqpi_free": (None, [_vp]),
    "cqpi_set_keep_windows": (C.c_int, [_vp, C.c_int]),
    "cqpi_gradient": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "cqpi_gradient_device": (C.c_int, [_vp, _vp, _dp, _dp, _dp]),
    "cqpi_set_tuning": (C.c_int, [_vp, C.c_int, C.c_int]),
    "cqpi_get_tuning": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "cqpi_set_tuning_ext": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int]),
    "cqpi_get_tuning_ext": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]),
    "cqpi_autotune": (C.c_int, [_vp, _dp, _sz, C.c_double, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "cqpi_batch_create": (C.c_int, [_vp, _sz, C.c_int, C.POINTER(_vp)]),
    "cqpi_batch_autotune": (C.c_int, [_vp, C.c_int]),
    "cqpi_batch_run": (C.c_int, [_vp]),
    "cqpi_batch_size": (C.c_int, [_vp, C.POINTER(_sz)]),
    "cqpi_batch_context": (C.c_int, [_vp, _sz, C.POINTER(_vp)]),
    "cqpi_batch_info": (C.c_int, [_vp, _dp, C.POINTER(_sz), C.POINTER(_sz)]),
    "cqpi_batch_destroy": (None, [_vp]),
}


class _Instance(C.Structure):
    """cqpi_instance of include/cqpi_amd.h."""
    _fields_ = [("kpoints", _sz), ("n", _sz), ("window", _sz), ("oversamp", _sz), ("bands", _sz),
                ("maxband", _sz), ("spin", C.c_int), ("scmodel", C.c_int), ("hops", _vp), ("nhops", _sz),
                ("efermi", C.c_double), ("V", _dp), ("wf", _dp), ("omegas", _dp), ("E", _sz),
                ("eta", C.c_double)]


class _DeviceArray:
    """Zero-copy view of library device memory (float64) for torch.as_tensor."""

    def __init__(self, ptr, shape):
        self.__cuda_array_interface__ = {"shape": tuple(shape), "typestr": "<f8", "data": (int(ptr), False),
                                         "version": 3, "strides": None}


class CqpiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("cqpi error %d: %s" % (code, msg))
        self.code = code


def lib():
    """The loaded libcqpi_amd.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError("libcqpi_amd.so not built (run __graft_entry__.build() or "
                               "make -C calcqpi-release_amd): %s" % LIBPATH)
        # torch carries its own HIP and HSA runtime libraries.  Loaded after
        # this library has initialised the GPU, that runtime finds no device
        # ("No HIP GPUs are available"); loaded first, both work side by side.
        # ldos_tensor() and cqpi.autograd hand device memory to torch, so torch
        # is loaded here, before the library, when it is installed.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _check(rc):
    if rc:
        raise CqpiError(rc, lib().cqpi_last_error().decode(errors="replace"))


def _d(a):
    return a.ctypes.data_as(_dp)


def device_count() -> int:
    return int(lib().cqpi_device_count())


def version() -> str:
    return lib().cqpi_version().decode()


def read_hr(path, scmodel=False):
    """Reads a Wannier90 _hr.dat file with the library's reader; returns
    (hops as synth.HOP_DTYPE array, bands, name)."""
    L = lib()
    hp = _vp()
    nh = _sz()
    nb = _sz()
    name = C.create_string_buffer(1024)
    _check(L.cqpi_read_hr(os.fsencode(path), int(scmodel), C.byref(hp), C.byref(nh), C.byref(nb), name, 1024))
    n = nh.value
    try:
        raw = C.string_at(hp.value, n * synth.HOP_DTYPE.itemsize) if n else b""
    finally:
        L.cqpi_free(hp)
    return np.frombuffer(raw, dtype=synth.HOP_DTYPE).copy(), nb.value, name.value.decode(errors="replace")


class AmdQPI:
    """GPUQPI drop-in plus the whole-chain extension (one HIP context)."""

    def __init__(self, wanniern, kpoints, n, window, oversamp, bands, maxband, spin, wf):
        self.W = int(wanniern)
        self.K = int(kpoints)
        self.n = int(n)
        self.window = int(window)
        self.oversamp = int(oversamp)
        self.bands = int(bands)
        self.maxband = int(maxband) if maxband else int(bands)
        self.spin = bool(spin)
        T = (2 * self.window + 1) * self.oversamp
        wf = np.ascontiguousarray(wf, dtype=np.float64)
        if wf.shape != (self.maxband, T, T):
            raise ValueError("wf must have shape (maxband, T, T) = %s, got %s" % ((self.maxband, T, T), wf.shape))
        self._wf = wf
        self._ctx = _vp()
        _check(lib().cqpi_create(self.W, self.K, self.n, self.window, self.oversamp, self.bands,
                                 self.maxband, int(self.spin), _d(wf), C.byref(self._ctx)))
        self.nlayers = 0
        self._nhops = 0

    # -- GPUQPI ------------------------------------------------------------
    def _dropin_args(self, scat, g0):
        b = self.bands
        scat = np.ascontiguousarray(scat, dtype=np.complex128)
        g0 = np.ascontiguousarray(g0, dtype=np.complex128)
        if scat.shape != (b, b) or g0.shape != (self.K, self.K, b, b):
            raise ValueError("shape mismatch")
        base = g0.ctypes.data
        ptrs = (base + np.arange(self.K * self.K, dtype=np.uint64) * np.uint64(b * b * 16)).astype(np.uint64)
        return scat, g0, ptrs

    def wannierldos(self, scat, g0):
        """scat: T (b x b complex); g0: G0(R) on the k-grid, shape (K, K, b, b)
        complex, g0[ix, iy] = G0(R = (ix, iy) mod K)."""
        scat, g0, ptrs = self._dropin_args(scat, g0)
        _check(lib().cqpi_wannierldos(self._ctx, _d(scat.view(np.float64)),
                                      ptrs.ctypes.data_as(C.POINTER(_vp)), self.bands))
        self.nlayers = 1

    def wannierldos_device(self, d_T, d_g0win, E, d_ldos):
        """cqpi_wannierldos_device: device pointers (ints) of E matrices T,
        E G0(R) windows (layout of g0_window) and the E output maps; nothing
        is allocated, the kernels run on the context stream and the call
        returns without waiting."""
        _check(lib().cqpi_wannierldos_device(self._ctx, int(d_T), int(d_g0win), int(E), int(d_ldos)))

    @classmethod
    def from_list(cls, wanniern, kpoints, n, window, oversamp, bands, flags, counts):
        """A threshold-list context (cqpi_create_list): flags FLAG_DTYPE entries,
        counts the (2w+1)^4 list lengths in IDX4CC order."""
        self = cls.__new__(cls)
        self.W, self.K, self.n = int(wanniern), int(kpoints), int(n)
        self.window, self.oversamp, self.bands = int(window), int(oversamp), int(bands)
        self.maxband, self.spin = self.bands, False
        self._flags = np.ascontiguousarray(flags, FLAG_DTYPE)
        self._counts = np.ascontiguousarray(counts, np.uintp)
        if self._counts.size != (2 * self.window + 1) ** 4 or int(self._counts.sum()) != self._flags.size:
            raise ValueError("counts must hold (2w+1)^4 list lengths summing to len(flags)")
        self._wf = None
        self._ctx = _vp()
        _check(lib().cqpi_create_list(self.W, self.K, self.n, self.window, self.oversamp, self.bands,
                                      self._flags.ctypes.data, self._counts.ctypes.data_as(C.POINTER(_sz)),
                                      C.byref(self._ctx)))
        self.nlayers = 0
        self._nhops = 0
        return self

    @classmethod
    def spectral(cls, kpoints, n, bands, maxbands=0, spin=False, pos=None, prearr=None):
        """A spectral-function context (cqpi_create_spectral): maps n x n; pos
        (maxbands x 2) and prearr for the unfolded spectral function."""
        self = cls.__new__(cls)
        self.K, self.n, self.W = int(kpoints), int(n), int(n)
        self.window, self.oversamp, self.bands = 0, 1, int(bands)
        self.maxband, self.spin = int(maxbands) if maxbands else int(bands), bool(spin)
        self._pos = np.ascontiguousarray(np.zeros((0, 2)) if pos is None else np.asarray(pos)[:, :2], np.float64)
        self._pre = np.ascontiguousarray(np.zeros(0) if prearr is None else prearr, np.float64)
        self._wf = None
        self._ctx = _vp()
        _check(lib().cqpi_create_spectral(self.K, self.n, self.bands, self.maxband, int(self.spin),
                                          _d(self._pos) if len(self._pos) else None, len(self._pos),
                                          _d(self._pre) if len(self._pre) else None, len(self._pre),
                                          C.byref(self._ctx)))
        self.nlayers = 0
        self._nhops = 0
        return self

    def _g0k_table(self, g0k):
        b = self.bands
        g0k = np.ascontiguousarray(g0k, dtype=np.complex128)
        if g0k.shape != (self.K, self.K, b, b):
            raise ValueError("g0 must have shape (K, K, b, b)")
        ptrs = (g0k.ctypes.data + np.arange(self.K * self.K, dtype=np.uint64) * np.uint64(b * b * 16)).astype(np.uint64)
        return g0k, ptrs

    def spf(self, g0k):
        """GPUQPI::spf: g0k = G0(k) on the grid, shape (K, K, b, b)."""
        g0k, ptrs = self._g0k_table(g0k)
        _check(lib().cqpi_spf(self._ctx, ptrs.ctypes.data_as(C.POINTER(_vp)), self.bands))
        self.nlayers = 1

    def uspf(self, g0k):
        """GPUQPI::uspf: g0k = G0(k) on the grid, shape (K, K, b, b)."""
        g0k, ptrs = self._g0k_table(g0k)
        _check(lib().cqpi_uspf(self._ctx, ptrs.ctypes.data_as(C.POINTER(_vp)), self.bands))
        self.nlayers = 1

    def set_spectral(self, mode):
        """cqpi_run computes SPF ("spf"), USPF ("uspf") maps, or none (None)."""
        m = {None: 0, "spf": 1, "uspf": 2}[mode]
        _check(lib().cqpi_set_spectral(self._ctx, m))

    def dos(self, omegas, eta):
        """Density of states -sum Im G0(R=0)_oo per energy (cqpi_dos)."""
        om = np.ascontiguousarray(omegas, np.float64)
        out = np.zeros(len(om))
        _check(lib().cqpi_dos(self._ctx, _d(om), len(om), float(eta), _d(out)))
        return out

    def wannierldoslist(self, scat, g0):
        """GPUQPI::wannierldoslist for a context made by from_list."""
        scat, g0, ptrs = self._dropin_args(scat, g0)
        _check(lib().cqpi_wannierldoslist(self._ctx, _d(scat.view(np.float64)),
                                          ptrs.ctypes.data_as(C.POINTER(_vp)), self.bands))
        self.nlayers = 1

    def wannierjosephson(self, scat, g0, tip):
        """As wannierldos, for the Josephson map of a superconducting model
        (particle-hole block, columns o + maxband) with the complex tip factor
        (src/qpihipkernel.cc:600-622)."""
        scat, g0, ptrs = self._dropin_args(scat, g0)
        tip = complex(tip)
        _check(lib().cqpi_wannierjosephson(self._ctx, _d(scat.view(np.float64)),
                                           ptrs.ctypes.data_as(C.POINTER(_vp)), self.bands, tip.real, tip.imag))
        self.nlayers = 1

    def retrieveResult(self, ldos=None):
        out = np.empty(self.W * self.W) if ldos is None else ldos
        _check(lib().cqpi_retrieve(self._ctx, _d(out)))
        return out

    def printinfo(self) -> str:
        buf = C.create_string_buffer(2048)
        _check(lib().cqpi_printinfo(self._ctx, buf, 2048))
        return buf.value.decode()

    # -- whole chain --------------------------------------------------------
    def set_josephson(self, on=True, deltat=0.001, etat=float("nan")):
        """cqpi_run computes the Josephson map (tip gap deltat, tip broadening
        etat; NaN: the run's eta) instead of the LDOS map."""
        _check(lib().cqpi_set_josephson(self._ctx, int(bool(on)), float(deltat), float(etat)))

    def bands_histogram(self, pix, oversamp, le, ue, layers):
        """Band-structure histogram (bsfile=): pix x pix x layers, [x, y, layer]."""
        out = np.zeros(layers * pix * pix)
        _check(lib().cqpi_bands_histogram(self._ctx, int(pix), int(oversamp), float(le), float(ue), int(layers),
                                          out.ctypes.data_as(_dp)))
        return out.reshape(layers, pix, pix).transpose(2, 1, 0)

    def set_surface(self, mode="surface", epserr=1e-5, shops=None):
        """cqpi_run uses the surface ("surface") or bulk ("bulk") Green's
        function of the layered model, or G0(k) of the eigen data ("normal");
        shops: in-plane hoppings of the surface layer (default: the model's)."""
        mode = GREENS.get(mode, mode)
        sh = np.ascontiguousarray(shops if shops is not None else np.zeros(0, synth.HOP_DTYPE), dtype=synth.HOP_DTYPE)
        _check(lib().cqpi_set_surface(self._ctx, int(mode), float(epserr), sh.ctypes.data if len(sh) else None,
                                      len(sh)))

    def set_surface_steps(self, maxit):
        """Step limit of the surface Green's function's decimation (default 500)."""
        _check(lib().cqpi_set_surface_steps(self._ctx, int(maxit)))

    def set_pair_symmetry(self, enable):
        """Lets runs transform only the orbital pairs r <= c where the model
        has H(k)^T = H(-k) (default on)."""
        _check(lib().cqpi_set_pair_symmetry(self._ctx, int(bool(enable))))

    def pair_symmetry(self):
        """(model, used): the model has H(k)^T = H(-k); the last run used it."""
        m, u = C.c_int(0), C.c_int(0)
        _check(lib().cqpi_get_pair_symmetry(self._ctx, C.byref(m), C.byref(u)))
        return bool(m.value), bool(u.value)

    def set_precision(self, mode):
        mode = PRECISIONS.get(mode, mode)
        _check(lib().cqpi_set_precision(self._ctx, int(mode)))

    def set_model(self, hops, efermi=0.0, scmodel=False):
        hops = np.ascontiguousarray(hops, dtype=synth.HOP_DTYPE)
        self._nhops = len(hops)
        _check(lib().cqpi_set_model(self._ctx, hops.ctypes.data if len(hops) else None, len(hops),
                                    float(efermi), int(scmodel)))

    def set_impurity(self, V):
        V = np.ascontiguousarray(V, dtype=np.complex128)
        if V.shape != (self.bands, self.bands):
            raise ValueError("V must be b x b")
        _check(lib().cqpi_set_impurity(self._ctx, _d(V.view(np.float64))))

    def compute_eigen(self):
        _check(lib().cqpi_compute_eigen(self._ctx))

    def set_timing(self, on=True):
        _check(lib().cqpi_set_timing(self._ctx, int(on)))

    def run(self, omegas, eta, host=True):
        om = np.ascontiguousarray(omegas, dtype=np.float64)
        E = len(om)
        out = np.empty((E, self.W * self.W)) if host else None
        _check(lib().cqpi_run(self._ctx, _d(om), E, float(eta), _d(out) if host else None))
        self.nlayers = E
        return out

    def retrieve_layers(self, first, count):
        out = np.empty((count, self.W * self.W))
        _check(lib().cqpi_retrieve_layers(self._ctx, first, count, _d(out)))
        return out

    def window_half(self):
        h = _sz()
        _check(lib().cqpi_window_half(self._ctx, C.byref(h)))
        return h.value

    def g0_window(self, e):
        h = self.window_half()
        nw = 2 * h + 1
        out = np.empty((nw, nw, self.bands, self.bands), np.complex128)
        _check(lib().cqpi_retrieve_g0_window(self._ctx, e, _d(out.view(np.float64))))
        return out

    def tmatrix(self, e):
        out = np.empty((self.bands, self.bands), np.complex128)
        _check(lib().cqpi_retrieve_tmatrix(self._ctx, e, _d(out.view(np.float64))))
        return out

    def stage_times(self):
        ms = np.zeros(5)
        n = (C.c_int64 * 5)()
        _check(lib().cqpi_stage_times(self._ctx, _d(ms), n))
        return {STAGES[i]: (float(ms[i]), int(n[i])) for i in range(5)}

    def set_keep_windows(self, on=True):
        _check(lib().cqpi_set_keep_windows(self._ctx, int(on)))

    def gradient(self, rhobar):
        """Adjoint of the last run (kept windows): rhobar = dL/drho, shape
        (E, W*W).  Returns (g_hops (nhops, 2) = dL/d(Re t, Im t),
        g_mu, g_V (b x b complex, dL = Re sum conj(g_V) dV))."""
        rb = np.ascontiguousarray(rhobar, dtype=np.float64)
        if rb.size != self.nlayers * self.W * self.W:
            raise ValueError("rhobar must hold E*W*W values of the last run")
        nh = self._nhops
        gh = np.zeros((max(nh, 1), 2))
        gmu = np.zeros(1)
        gV = np.zeros((self.bands, self.bands), np.complex128)
        _check(lib().cqpi_gradient(self._ctx, _d(rb), _d(gh), _d(gmu), _d(gV.view(np.float64))))
        return gh[:nh], float(gmu[0]), gV

    def gradient_device(self, rhobar_ptr):
        """As gradient() with dL/drho already in device memory (address of
        E*W*W float64 values); the caller synchronises its writes first."""
        nh = self._nhops
        gh = np.zeros((max(nh, 1), 2))
        gmu = np.zeros(1)
        gV = np.zeros((self.bands, self.bands), np.complex128)
        _check(lib().cqpi_gradient_device(self._ctx, C.c_void_p(int(rhobar_ptr)), _d(gh), _d(gmu),
                                          _d(gV.view(np.float64))))
        return gh[:nh], float(gmu[0]), gV

    def set_tuning(self, tpw=0, eb=0):
        _check(lib().cqpi_set_tuning(self._ctx, int(tpw), int(eb)))

    def tuning(self):
        t, e = C.c_int(), C.c_int()
        _check(lib().cqpi_get_tuning(self._ctx, C.byref(t), C.byref(e)))
        return t.value, e.value

    def set_tuning_ext(self, tpw=0, eb=0, js=0, streams=0):
        _check(lib().cqpi_set_tuning_ext(self._ctx, int(tpw), int(eb), int(js), int(streams)))

    def tuning_ext(self):
        """(tpw, eb, js, streams) of the context."""
        v = [C.c_int() for _ in range(4)]
        _check(lib().cqpi_get_tuning_ext(self._ctx, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def autotune(self, omegas, eta, reps=3):
        """Times candidate launch tunings for this shape on the device and keeps
        the fastest; returns (tpw, eb)."""
        om = np.ascontiguousarray(omegas, dtype=np.float64)
        t, e = C.c_int(), C.c_int()
        _check(lib().cqpi_autotune(self._ctx, _d(om), len(om), float(eta), int(reps), C.byref(t), C.byref(e)))
        self.nlayers = len(om)
        return t.value, e.value

    def ldos_tensor(self):
        """The maps of the last run as a torch CUDA tensor (E, W*W), a
        zero-copy view of the library buffer (valid until the next run)."""
        import torch

        ptr, cnt = self.device_ldos()
        return torch.as_tensor(_DeviceArray(ptr, (self.nlayers, self.W * self.W)), device="cuda")

    def device_ldos(self):
        p = _vp()
        cnt = _sz()
        _check(lib().cqpi_device_ldos(self._ctx, C.byref(p), C.byref(cnt)))
        return p.value, cnt.value

    def close(self):
        if getattr(self, "_ctx", None) and self._ctx.value:
            lib().cqpi_destroy(self._ctx)
            self._ctx = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def orbital_tile(name, radius=0.5, height=5.0, xpos=0.0, ypos=0.0, theta=0.0, prefact=1.0, window=2, oversamp=4):
    """An analytic orbital sampled on the Wannier tile grid (cqpi_orbital_tile):
    T x T, T = (2 window + 1) oversamp; theta in radians."""
    T = (2 * window + 1) * oversamp
    out = np.zeros((T, T))
    _check(lib().cqpi_orbital_tile(name.encode(), float(radius), float(height), float(xpos), float(ypos),
                                   float(theta), float(prefact), int(window), int(oversamp),
                                   out.ctypes.data_as(_dp)))
    return out


def xsf_tile(path, height=5.0, window=2, oversamp=4, prefact=1.0):
    """Wannier tile of an .xsf orbital file (cqpi_xsf_tile): T x T."""
    T = (2 * window + 1) * oversamp
    out = np.zeros((T, T))
    _check(lib().cqpi_xsf_tile(os.fsencode(path), float(height), int(window), int(oversamp), float(prefact),
                               out.ctypes.data_as(_dp)))
    return out


def make_threshold_list(tiles, window, oversamp, threshold, maxband=0, spin=False):
    """cqpi_make_threshold_list (src/qpi.cc:139-211): (flags FLAG_DTYPE,
    counts per window pair in IDX4CC order)."""
    tiles = np.ascontiguousarray(tiles, np.float64)
    norb = tiles.shape[0]
    counts = np.zeros((2 * window + 1) ** 4, np.uintp)
    total = _sz(0)
    _check(lib().cqpi_make_threshold_list(_d(tiles), norb, window, oversamp, float(threshold), maxband, int(spin),
                                          None, 0, counts.ctypes.data_as(C.POINTER(_sz)), C.byref(total)))
    flags = np.zeros(max(total.value, 1), FLAG_DTYPE)
    _check(lib().cqpi_make_threshold_list(_d(tiles), norb, window, oversamp, float(threshold), maxband, int(spin),
                                          flags.ctypes.data, total.value, counts.ctypes.data_as(C.POINTER(_sz)),
                                          C.byref(total)))
    return flags[:total.value], counts


def make_threshold_list_spinarr(tiles, window, oversamp, threshold, spinarr):
    """cqpi_make_threshold_list_spinarr (src/qpi.cc:213-253): the list over
    orbital pairs with equal nonzero spin labels; (flags FLAG_DTYPE, counts
    per window pair in IDX4CC order)."""
    tiles = np.ascontiguousarray(tiles, np.float64)
    norb = tiles.shape[0]
    sp = np.ascontiguousarray(spinarr, np.int32)
    spp = sp.ctypes.data_as(C.POINTER(C.c_int))
    counts = np.zeros((2 * window + 1) ** 4, np.uintp)
    total = _sz(0)
    _check(lib().cqpi_make_threshold_list_spinarr(_d(tiles), norb, window, oversamp, float(threshold), spp, len(sp),
                                                  None, 0, counts.ctypes.data_as(C.POINTER(_sz)), C.byref(total)))
    flags = np.zeros(max(total.value, 1), FLAG_DTYPE)
    _check(lib().cqpi_make_threshold_list_spinarr(_d(tiles), norb, window, oversamp, float(threshold), spp, len(sp),
                                                  flags.ctypes.data, total.value,
                                                  counts.ctypes.data_as(C.POINTER(_sz)), C.byref(total)))
    return flags[:total.value], counts


def engine_for(wl: "synth.Workload", precision="fp64") -> AmdQPI:
    """A context for a synth.Workload with model and impurity uploaded."""
    q = AmdQPI(wl.W, wl.K, wl.n, wl.window, wl.oversamp, wl.bands, wl.maxband, wl.spin, wl.tiles)
    q.set_model(wl.hops, wl.efermi, wl.scmodel)
    q.set_impurity(wl.V)
    q.set_precision(precision)
    return q


class _Borrowed(AmdQPI):
    """An instance context owned by a Batch (not destroyed by this object)."""

    def __init__(self, ctx, wl, nlayers):
        self._ctx = ctx
        self.W, self.K, self.n = wl.W, wl.K, wl.n
        self.window, self.oversamp, self.bands = wl.window, wl.oversamp, wl.bands
        self.maxband, self.spin = wl.maxband, wl.spin
        self.nlayers = nlayers
        self._nhops = len(wl.hops)

    def close(self):
        self._ctx = _vp()


class Batch:
    """Many independent instances (synth.Workload each) run as one captured
    HIP graph: cqpi_batch_* of include/cqpi_amd.h."""

    def __init__(self, workloads, precision="fp64"):
        self.workloads = list(workloads)
        self._keep = []
        arr = (_Instance * len(self.workloads))()
        for i, wl in enumerate(self.workloads):
            hops = np.ascontiguousarray(wl.hops, dtype=synth.HOP_DTYPE)
            V = np.ascontiguousarray(wl.V, dtype=np.complex128)
            wf = np.ascontiguousarray(wl.tiles, dtype=np.float64)
            om = np.ascontiguousarray(wl.energies, dtype=np.float64)
            self._keep += [hops, V, wf, om]
            arr[i] = _Instance(wl.K, wl.n, wl.window, wl.oversamp, wl.bands, wl.maxband, int(wl.spin),
                               int(wl.scmodel), hops.ctypes.data if len(hops) else None, len(hops),
                               float(wl.efermi), _d(V.view(np.float64)), _d(wf), _d(om), len(om), float(wl.eta))
        self._b = _vp()
        _check(lib().cqpi_batch_create(arr, len(self.workloads), PRECISIONS.get(precision, precision),
                                       C.byref(self._b)))
        self._keep = []

    def autotune(self, reps=3):
        _check(lib().cqpi_batch_autotune(self._b, int(reps)))

    def run(self):
        """Launches the batch graph once (captured on the first call); returns
        the device milliseconds of the launch."""
        _check(lib().cqpi_batch_run(self._b))
        return self.info()["last_ms"]

    def info(self):
        ms = np.zeros(1)
        nodes, mem = _sz(), _sz()
        _check(lib().cqpi_batch_info(self._b, _d(ms), C.byref(nodes), C.byref(mem)))
        return {"last_ms": float(ms[0]), "graph_nodes": nodes.value, "gpumem": mem.value}

    def instance(self, i) -> AmdQPI:
        """Context of instance i (valid while the batch lives)."""
        c = _vp()
        _check(lib().cqpi_batch_context(self._b, i, C.byref(c)))
        return _Borrowed(c, self.workloads[i], len(self.workloads[i].energies))

    def maps(self, i):
        wl = self.workloads[i]
        return self.instance(i).retrieve_layers(0, len(wl.energies))

    def close(self):
        if getattr(self, "_b", None) and self._b.value:
            lib().cqpi_batch_destroy(self._b)
            self._b = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
