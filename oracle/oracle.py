"""ctypes wrapper of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
Build with ``make -C oracle``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

HOP_DTYPE = np.dtype(
    [("rx", "<f8"), ("ry", "<f8"), ("rz", "<f8"), ("o1", "<i4"), ("o2", "<i4"),
     ("tre", "<f8"), ("tim", "<f8"), ("anti", "<i4"), ("reserved", "<i4")],
    align=True,
)

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        sz = C.c_size_t
        L.or_read_hr.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_void_p), C.POINTER(sz),
                                 C.POINTER(sz), C.c_char_p, sz]
        L.or_free.argtypes = [C.c_void_p]
        L.or_setmatrix.argtypes = [C.c_void_p, sz, sz, C.c_double, sz, C.c_double, C.c_double, _dp]
        L.or_eigh.argtypes = [sz, _dp, _dp, _dp]
        L.or_calchamiltonian.argtypes = [C.c_void_p, sz, sz, C.c_double, sz, sz, _dp, _dp]
        L.or_setgreensfunction.argtypes = [sz, sz, _dp, _dp, C.c_double, C.c_double, _dp]
        L.or_realspace.argtypes = [sz, sz, _dp]
        L.or_tmatrix.argtypes = [sz, _dp, _dp, _dp]
        L.or_tmatrix.restype = C.c_int
        L.or_wannierldos.argtypes = [sz, sz, sz, sz, sz, sz, C.c_int, _dp, _dp, _dp, _ip, _dp]
        L.or_wannierldos_cells.argtypes = [sz, sz, sz, sz, sz, sz, C.c_int, _dp, _dp, _dp, _ip, sz, sz, sz, sz,
                                           _dp]
        L.or_wannierjosephson.argtypes = [sz, sz, sz, sz, sz, sz, C.c_int, _dp, _dp, _dp, _ip,
                                          C.c_double, C.c_double, _dp]
        L.or_mkthresholdlist.argtypes = [_dp, sz, sz, sz, C.c_double, sz, C.c_int, C.c_void_p, sz,
                                         C.POINTER(sz)]
        L.or_mkthresholdlist.restype = sz
        L.or_mkthresholdlist_spinarr.argtypes = [_dp, sz, sz, sz, C.c_double, _ip, sz, C.c_void_p, sz,
                                                 C.POINTER(sz)]
        L.or_mkthresholdlist_spinarr.restype = sz
        L.or_wannierldoslist.argtypes = [sz, sz, sz, sz, sz, _dp, _dp, C.c_void_p, C.POINTER(sz), _dp]
        L.or_spf.argtypes = [sz, sz, sz, sz, _dp, _dp]
        L.or_surfacegreen.argtypes = [C.c_void_p, sz, C.c_void_p, sz, sz, C.c_double, sz, sz, C.c_double,
                                      C.c_double, C.c_double, C.c_int, C.c_int, _dp]
        L.or_surfacegreen.restype = C.c_int
        L.or_uspf.argtypes = [sz, sz, sz, sz, C.c_int, _dp, _dp, _dp, _dp]
        L.or_cldos_run.argtypes = [C.c_void_p, sz, sz, C.c_double, C.c_int, sz, sz, sz, sz, sz,
                                   C.c_int, C.c_double, _dp, sz, _dp, _dp, _ip, _dp]
        L.or_cldos_run.restype = C.c_int
        L.or_getbias.argtypes = [C.c_double, C.c_double, sz, sz]
        L.or_getbias.restype = C.c_double
        L.or_idl_write.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, sz, sz, sz, _dp, _dp]
        L.or_idl_read_header.argtypes = [C.c_char_p, C.POINTER(sz), _dp]
        L.or_idl_read.argtypes = [C.c_char_p, _dp, sz]
        L.or_num_threads.restype = C.c_int
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(_dp)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def read_hr(path, scmodel=False):
    L = lib()
    hp = C.c_void_p()
    nh = C.c_size_t()
    nb = C.c_size_t()
    name = C.create_string_buffer(512)
    err = L.or_read_hr(path.encode(), int(scmodel), C.byref(hp), C.byref(nh), C.byref(nb), name, 512)
    if err:
        raise ValueError("or_read_hr failed with code %d" % err)
    n = nh.value
    buf = (C.c_char * (n * HOP_DTYPE.itemsize)).from_address(hp.value)
    hops = np.frombuffer(bytes(buf), dtype=HOP_DTYPE).copy()
    L.or_free(hp)
    return hops, nb.value, name.value.decode()


def setmatrix(hops, b, efermi, kx, ky, scmodel=False):
    hops = _c(hops, HOP_DTYPE)
    H = np.zeros((b, b), np.complex128)
    scmax = b >> 1 if scmodel else b
    lib().or_setmatrix(hops.ctypes.data, len(hops), b, efermi, scmax, kx, ky, _d(H.view(np.float64)))
    return H


def eigh(H):
    H = _c(H, np.complex128)
    b = H.shape[0]
    ev = np.zeros(b)
    V = np.zeros((b, b), np.complex128)
    lib().or_eigh(b, _d(H.view(np.float64)), _d(ev), _d(V.view(np.float64)))
    return ev, V


def calchamiltonian(hops, b, efermi, K, scmodel=False):
    hops = _c(hops, HOP_DTYPE)
    ev = np.zeros((K, K, b))
    V = np.zeros((K, K, b, b), np.complex128)
    scmax = b >> 1 if scmodel else b
    lib().or_calchamiltonian(hops.ctypes.data, len(hops), b, efermi, scmax, K, _d(ev),
                             _d(V.view(np.float64)))
    return ev, V


def greens_k(ev, V, omega, eta):
    K, _, b = ev.shape
    g = np.zeros((K, K, b, b), np.complex128)
    lib().or_setgreensfunction(K, b, _d(_c(ev, np.float64)), _d(_c(V, np.complex128).view(np.float64)),
                               omega, eta, _d(g.view(np.float64)))
    return g


def realspace(g0k):
    g = _c(g0k, np.complex128).copy()
    K, _, b, _ = g.shape
    lib().or_realspace(K, b, _d(g.view(np.float64)))
    return g


def tmatrix(V, g00):
    V = _c(V, np.complex128)
    b = V.shape[0]
    T = np.zeros((b, b), np.complex128)
    err = lib().or_tmatrix(b, _d(V.view(np.float64)), _d(_c(g00, np.complex128).view(np.float64)),
                           _d(T.view(np.float64)))
    if err:
        raise ZeroDivisionError("singular 1 - V G0(0)")
    return T


def wannierldos(g0R, T, tiles, n, window, oversamp, maxband=0, spin=False, wfzero=None):
    K, _, b, _ = g0R.shape
    W = n * oversamp
    out = np.zeros(W * W)
    tiles = _c(tiles, np.float64)
    if not maxband:
        maxband = b
    wz = None
    if wfzero is not None:
        wz = _c(wfzero, np.int32).ctypes.data_as(_ip)
    lib().or_wannierldos(K, n, window, oversamp, b, maxband, int(spin),
                         _d(_c(T, np.complex128).view(np.float64)),
                         _d(_c(g0R, np.complex128).view(np.float64)), _d(tiles), wz, _d(out))
    return out


def wannierldos_cells(g0R, T, tiles, n, window, oversamp, cells, maxband=0, spin=False):
    """or_wannierldos on the cells i in [i0, i1), j in [j0, j1) only
    (cells = (i0, i1, j0, j1)); the other pixels of the W*W map are 0."""
    K, _, b, _ = g0R.shape
    W = n * oversamp
    out = np.zeros(W * W)
    if not maxband:
        maxband = b
    i0, i1, j0, j1 = cells
    lib().or_wannierldos_cells(K, n, window, oversamp, b, maxband, int(spin),
                               _d(_c(T, np.complex128).view(np.float64)),
                               _d(_c(g0R, np.complex128).view(np.float64)), _d(_c(tiles, np.float64)), None,
                               i0, i1, j0, j1, _d(out))
    return out


def wannierjosephson(g0R, T, tiles, n, window, oversamp, tip, maxband=0, spin=False, wfzero=None):
    """Josephson map (src/qpi.cc:953-1033): tiles for maxband orbitals (0:
    bands/2), particle-hole column o1 + maxband, complex tip factor."""
    K, _, b, _ = g0R.shape
    W = n * oversamp
    out = np.zeros(W * W)
    tiles = _c(tiles, np.float64)
    if not maxband:
        maxband = b // 2
    wz = None
    if wfzero is not None:
        wz = _c(wfzero, np.int32).ctypes.data_as(_ip)
    tip = complex(tip)
    lib().or_wannierjosephson(K, n, window, oversamp, b, maxband, int(spin),
                              _d(_c(T, np.complex128).view(np.float64)),
                              _d(_c(g0R, np.complex128).view(np.float64)), _d(tiles), wz,
                              tip.real, tip.imag, _d(out))
    return out


# {i, j, o1, o2, factor}: the layout of the reference's qpigpuflaglist
# (src/qpihipkernel.h:47-50) and of cqpi_flagentry
FLAG_DTYPE = np.dtype([("i", "<u4"), ("j", "<u4"), ("o1", "<u4"), ("o2", "<u4"), ("factor", "<f8")])


def mkthresholdlist(tiles, window, oversamp, threshold, maxband=0, spin=False):
    """src/qpi.cc:139-211: (entries FLAG_DTYPE, counts per window pair in
    IDX4CC order)."""
    tiles = _c(tiles, np.float64)
    norb = tiles.shape[0]
    wins = 2 * window + 1
    counts = np.zeros(wins ** 4, np.uintp)
    total = lib().or_mkthresholdlist(_d(tiles), norb, window, oversamp, float(threshold), maxband, int(spin),
                                     None, 0, counts.ctypes.data_as(C.POINTER(C.c_size_t)))
    out = np.zeros(max(total, 1), FLAG_DTYPE)
    lib().or_mkthresholdlist(_d(tiles), norb, window, oversamp, float(threshold), maxband, int(spin),
                             out.ctypes.data, total, counts.ctypes.data_as(C.POINTER(C.c_size_t)))
    return out[:total], counts


def mkthresholdlist_spinarr(tiles, window, oversamp, threshold, spinarr):
    """src/qpi.cc:213-253 (spin label per orbital): (entries FLAG_DTYPE,
    counts per window pair in IDX4CC order)."""
    tiles = _c(tiles, np.float64)
    norb = tiles.shape[0]
    sp = np.ascontiguousarray(spinarr, np.int32)
    wins = 2 * window + 1
    counts = np.zeros(wins ** 4, np.uintp)
    total = lib().or_mkthresholdlist_spinarr(_d(tiles), norb, window, oversamp, float(threshold),
                                             sp.ctypes.data_as(_ip), len(sp), None, 0,
                                             counts.ctypes.data_as(C.POINTER(C.c_size_t)))
    out = np.zeros(max(total, 1), FLAG_DTYPE)
    lib().or_mkthresholdlist_spinarr(_d(tiles), norb, window, oversamp, float(threshold), sp.ctypes.data_as(_ip),
                                     len(sp), out.ctypes.data, total, counts.ctypes.data_as(C.POINTER(C.c_size_t)))
    return out[:total], counts


def wannierldoslist(g0R, T, flags, counts, n, window, oversamp):
    """src/qpi.cc:1088-1135: the map from a threshold list."""
    K, _, b, _ = g0R.shape
    W = n * oversamp
    out = np.zeros(W * W)
    flags = np.ascontiguousarray(flags, FLAG_DTYPE)
    counts = np.ascontiguousarray(counts, np.uintp)
    lib().or_wannierldoslist(K, n, window, oversamp, b, _d(_c(T, np.complex128).view(np.float64)),
                             _d(_c(g0R, np.complex128).view(np.float64)), flags.ctypes.data,
                             counts.ctypes.data_as(C.POINTER(C.c_size_t)), _d(out))
    return out


def surfacegreen(hops, b, efermi, K, omega, eta, epserr=1e-5, surface=True, shops=None, scmodel=False,
                 maxit=500):
    """src/qpi.cc:488-554: G0(k) (K, K, b, b) of the surface (or bulk) of a
    layered model (in-plane hops rz == 0, coupling to the next layer rz < 0),
    and the largest iteration count (-1 if maxit was reached)."""
    hops = _c(hops, HOP_DTYPE)
    sh = None if shops is None else _c(shops, HOP_DTYPE)
    g = np.zeros((K, K, b, b), np.complex128)
    scmax = b >> 1 if scmodel else b
    it = lib().or_surfacegreen(hops.ctypes.data, len(hops), None if sh is None else sh.ctypes.data,
                               0 if sh is None else len(sh), b, efermi, scmax, K, omega, eta, epserr,
                               int(surface), maxit, _d(g.view(np.float64)))
    return g, it


_ORB_ANG = {
    # real harmonic times r^l, then divided by r^l: wannierfunctions::swave ...
    # fxx2wave, src/wannierfunctions.cc:21-124 (polynomial in x, y, z over r^l)
    "px": (1, lambda x, y, z: np.sqrt(3 / 4 / np.pi) * x),
    "py": (1, lambda x, y, z: np.sqrt(3 / 4 / np.pi) * y),
    "pz": (1, lambda x, y, z: np.sqrt(3 / 4 / np.pi) * z),
    "dx2": (2, lambda x, y, z: np.sqrt(15 / 4 / np.pi) * (x * x - y * y) / 2),
    "dxy": (2, lambda x, y, z: np.sqrt(15 / 4 / np.pi) * x * y),
    "dxz": (2, lambda x, y, z: np.sqrt(15 / 4 / np.pi) * x * z),
    "dyz": (2, lambda x, y, z: np.sqrt(15 / 4 / np.pi) * y * z),
    "dr2": (2, lambda x, y, z: np.sqrt(5 / np.pi) * (2 * z * z - x * x - y * y) / 4),
    "fy3x2": (3, lambda x, y, z: np.sqrt(35 / 2 / np.pi) * y * (3 * x * x - y * y) / 4),
    "fxyz": (3, lambda x, y, z: np.sqrt(105 / 4 / np.pi) * x * y * z),
    "fyz2": (3, lambda x, y, z: np.sqrt(21 / 2 / np.pi) * y * (4 * z * z - x * x - y * y) / 4),
    "fz3": (3, lambda x, y, z: np.sqrt(7 / np.pi) * z * (2 * z * z - 3 * x * x - 3 * y * y) / 4),
    "fxz2": (3, lambda x, y, z: np.sqrt(21 / 2 / np.pi) * x * (4 * z * z - x * x - y * y) / 4),
    "fzx2": (3, lambda x, y, z: np.sqrt(105 / np.pi) * z * (x * x - y * y) / 4),
    "fxx2": (3, lambda x, y, z: np.sqrt(35 / 2 / np.pi) * x * (x * x - 3 * y * y) / 4),
}


def orbital_value(name, radius, x, y, z, prefact=1.0, symmetry=4.0):
    """wannierfunctions::<name>wave at (x, y, z) in the orbital's frame
    (src/wannierfunctions.cc:21-124)."""
    x, y, z = np.broadcast_arrays(*(np.asarray(a, np.float64) for a in (x, y, z)))
    r = np.sqrt(x * x + y * y + z * z)
    radial = prefact * np.exp(-r / radius)
    if name == "z":
        return np.zeros_like(r)
    if name == "s":
        return radial / np.sqrt(4 * np.pi)
    if name == "n":
        c = np.cos(symmetry * np.arctan2(x, y) / 2)
        return radial * c * c * r / radius
    l, f = _ORB_ANG[name]
    with np.errstate(invalid="ignore", divide="ignore"):
        v = radial * f(x, y, z) / r ** l
    return np.where(r == 0.0, 0.0, v)


def orbital_tile(name, radius, height, xpos, ypos, theta, prefact, window, oversamp):
    """wannierfunctions::precalculate (src/wannierfunctions.cc:436-442) of the
    orbital wannierfunctions(name, radius, height, xpos, ypos, theta):
    getwave(x, y) with z = 0, shifted by (xpos, ypos, height), rotated by theta
    about z unless s (src/wannierfunctions.cc:126-222, .h:34-40)."""
    T = (2 * window + 1) * oversamp
    g = np.arange(T) / oversamp - window - 0.5
    x, y = np.meshgrid(g - xpos, g - ypos, indexing="ij")
    z = np.full_like(x, -height)
    if theta != 0.0 and name != "s":
        x, y = x * np.cos(theta) + y * np.sin(theta), -x * np.sin(theta) + y * np.cos(theta)
    return orbital_value(name, radius, x, y, z, prefact)


def xsf_read(path):
    """xcfloader::readfile (src/loadxcf.cc:33-185): CONVVEC rows, atom z
    coordinates, data-grid origin, spanning vectors (rows) and values [i, j, k]."""
    lines = [ln for ln in open(path).read().splitlines() if "#" not in ln]
    conv, atoms, grid = np.eye(3), [], None
    i = 0
    while i < len(lines):
        ln = lines[i]
        if "CONVVEC" in ln:
            conv = np.array([[float(t) for t in lines[i + r].split()[:3]] for r in (1, 2, 3)])
            i += 3
        elif "PRIMCOORD" in ln:
            n = int(lines[i + 1].split()[0])
            atoms = [float(lines[i + 2 + a].split()[3]) for a in range(n)]
            i += 1 + n
        elif "BEGIN_BLOCK_DATAGRID_3D" in ln:
            nx, ny, nz = (int(t) for t in lines[i + 3].split()[:3])
            origin = np.array([float(t) for t in lines[i + 4].split()[:3]])
            span = np.array([[float(t) for t in lines[i + 5 + a].split()[:3]] for a in range(3)])
            toks = " ".join(lines[i + 8:]).split()
            vals = np.array([float(t) for t in toks[:nx * ny * nz]])
            grid = (origin, span, vals.reshape(nz, ny, nx).transpose(2, 1, 0))
            break
        i += 1
    return conv, np.array(atoms), grid


def xsf_tile(path, height, window, oversamp, prefact=1.0):
    """wannierfunctions(filename, height, window, oversamp, prefact)
    (src/wannierfunctions.cc:444-461) through xcfloader::getvalue, operator()
    and getzheight (src/loadxcf.cc:220-236, 330-349)."""
    conv, atoms, (origin, span, cube) = xsf_read(path)
    zh = atoms.max() + height if height > 0 else atoms.min() + height if height < 0 else np.nan
    T = (2 * window + 1) * oversamp
    g = np.arange(T) / oversamp - window
    xu, yu = np.meshgrid(g, g, indexing="ij")
    p = np.stack([xu * conv[0, 0] + yu * conv[1, 0] - origin[0], xu * conv[0, 1] + yu * conv[1, 1] - origin[1],
                  np.full_like(xu, zh - origin[2])], -1)
    f = np.linalg.solve(span.T, p.reshape(-1, 3).T).T.reshape(T, T, 3)
    n = np.array(cube.shape)
    with np.errstate(invalid="ignore"):
        idx = np.sign(f * n) * np.floor(np.abs(f * n) + 0.5)  # C round()
        ok = np.all((idx >= 0) & (idx < n), -1)
    out = np.zeros((T, T))
    ii = np.where(ok[..., None], idx, 0).astype(int)
    out[ok] = cube[ii[..., 0], ii[..., 1], ii[..., 2]][ok]
    return prefact * out


def bands_histogram(hops, b, efermi, pix, oversamp, le, ue, layers, scmodel=False):
    """tightbind::calcbst3dhisthr (src/tightbinding.cc:1158-1193, ux = uy = 1,
    all bands): pix x pix x layers, [x, y, layer]."""
    ospix = pix * oversamp
    g = np.arange(ospix) / (ospix - 1) - 0.5
    kx, ky = np.meshgrid(g, g, indexing="ij")
    H = np.zeros((ospix, ospix, b, b), complex)
    scmax = b >> 1 if scmodel else b
    for o in range(b):
        H[..., o, o] = -efermi if o < scmax else efermi
    for h in np.asarray(hops):
        ph = 2 * np.pi * (h["rx"] * kx + h["ry"] * ky) * (-1 if h["anti"] else 1)
        H[..., h["o1"], h["o2"]] += (h["tre"] + 1j * h["tim"]) * np.exp(1j * ph)
    ev = np.linalg.eigvalsh(H)
    out = np.zeros((pix, pix, layers))
    binsize = (ue - le) / (layers - 1)
    w = (ue - le) / layers * 2.0
    for io in range(ospix):
        for jo in range(ospix):
            for e in ev[io, jo]:
                if not (le - 2 * binsize < e < ue + 2 * binsize):
                    continue
                v = (e - le) / (ue - le) * (layers - 1)
                lay = int(np.sign(v) * np.floor(abs(v) + 0.5))  # C round(): halves away from zero
                for nl in range(lay - 1, lay + 2):
                    if 0 <= nl < layers:
                        x = (ue - le) * nl / (layers - 1) + le - e
                        out[io // oversamp, jo // oversamp, nl] += w / 2 / (x * x + w * w / 4) / np.pi
    return out


def spf(g0k, n, maxband=0):
    """Spectral function map (src/qpihipkernel.cc:317-329): n x n, x + y*n."""
    K, _, b, _ = g0k.shape
    out = np.zeros(n * n)
    lib().or_spf(K, n, b, maxband, _d(_c(g0k, np.complex128).view(np.float64)), _d(out))
    return out


def uspf(g0k, n, pos, prearr, maxband=0, spin=False):
    """Unfolded spectral function map (src/qpi.cc:731-775): pos maxband x 2."""
    K, _, b, _ = g0k.shape
    mb = maxband or b
    out = np.zeros(n * n)
    lib().or_uspf(K, n, b, mb, int(spin), _d(_c(np.asarray(pos)[:mb, :2], np.float64)),
                  _d(_c(np.asarray(prearr)[:mb], np.float64)), _d(_c(g0k, np.complex128).view(np.float64)),
                  _d(out))
    return out


def dos(g0R, maxband=0):
    """tmatrix::calcdos (src/qpi.cc:856-862): -sum_{i<maxband} Im G0(R=0)_ii."""
    b = g0R.shape[2]
    mb = maxband or b
    return -float(np.sum(np.imag(np.diag(g0R[0, 0])[:mb])))


def cldos_run(wl, energies=None):
    """Whole FP64 path for a synth.Workload; returns E x W*W (x + y*W)."""
    energies = wl.energies if energies is None else np.asarray(energies, np.float64)
    hops = _c(wl.hops, HOP_DTYPE)
    E = len(energies)
    out = np.zeros((E, wl.W * wl.W))
    err = lib().or_cldos_run(hops.ctypes.data, len(hops), wl.bands, wl.efermi, int(wl.scmodel),
                             wl.K, wl.n, wl.window, wl.oversamp, wl.maxband, int(wl.spin), wl.eta,
                             _d(_c(energies, np.float64)), E,
                             _d(_c(wl.V, np.complex128).view(np.float64)),
                             _d(_c(wl.tiles, np.float64)), None, _d(out))
    if err:
        raise ZeroDivisionError("singular 1 - V G0(0)")
    return out


def getbias(lb, hb, layers, i):
    return lib().or_getbias(lb, hb, layers, i)


def idl_write(path, data_xyz, name="", date="", hdr=(1, 1, 0, 0, 0, 0, 0)):
    sx, sy, sz = data_xyz.shape
    flat = np.ascontiguousarray(np.transpose(data_xyz, (2, 1, 0)), dtype=np.float64)
    h = np.array(hdr, np.float64)
    err = lib().or_idl_write(path.encode(), name.encode(), date.encode(), sx, sy, sz, _d(h), _d(flat))
    if err:
        raise IOError("or_idl_write failed")


def idl_read(path):
    L = lib()
    dims = (C.c_size_t * 3)()
    h = np.zeros(7)
    if L.or_idl_read_header(path.encode(), dims, _d(h)):
        raise IOError("bad idl header")
    sx, sy, sz = dims
    flat = np.zeros(sx * sy * sz)
    if L.or_idl_read(path.encode(), _d(flat), flat.size):
        raise IOError("bad idl data")
    return flat.reshape(sz, sy, sx).transpose(2, 1, 0), h


def num_threads():
    return lib().or_num_threads()
