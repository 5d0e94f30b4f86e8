"""Analytic orbitals of the driver's orbitals= mode: the oracle's restatement
pinned by what the mathematics fixes (orthonormal real harmonics, Unsoeld's
sums, rotations that map harmonics onto each other, the tile grid's centre
and shift), then the library's cqpi_orbital_tile against it (host code, no GPU)."""
import numpy as np
import pytest

import cqpi
import oracle

NAMES = ["s", "px", "py", "pz", "dx2", "dxy", "dxz", "dyz", "dr2",
         "fy3x2", "fxyz", "fyz2", "fz3", "fxz2", "fzx2", "fxx2"]
SHELLS = {0: ["s"], 1: ["px", "py", "pz"], 2: ["dx2", "dxy", "dxz", "dyz", "dr2"],
          3: ["fy3x2", "fxyz", "fyz2", "fz3", "fxz2", "fzx2", "fxx2"]}


def _sphere():
    """Gauss-Legendre in cos(theta) x uniform phi: exact for the products of
    two harmonics of degree <= 3 on the unit sphere."""
    mu, wmu = np.polynomial.legendre.leggauss(8)
    phi = np.arange(16) * 2 * np.pi / 16
    M, P = np.meshgrid(mu, phi, indexing="ij")
    s = np.sqrt(1 - M * M)
    w = np.outer(wmu, np.full(16, 2 * np.pi / 16))
    return s * np.cos(P), s * np.sin(P), M, w


def test_orbital_angular_parts_are_orthonormal():
    """At r = 1 and radius -> infinity (radial factor 1) every orbital is a
    real spherical harmonic: the 16 are orthonormal over the sphere."""
    x, y, z, w = _sphere()
    Y = {n: oracle.orbital_value(n, 1e300, x, y, z) for n in NAMES}
    G = np.array([[np.sum(w * Y[a] * Y[b]) for b in NAMES] for a in NAMES])
    assert np.allclose(G, np.eye(len(NAMES)), atol=1e-13)


def test_orbital_unsoeld_sums():
    """sum over a shell of Y^2 = (2l+1)/(4 pi) at every point, with the same
    radial factor: (2l+1) times the s orbital squared."""
    rng = np.random.default_rng(3)
    x, y, z = rng.normal(size=(3, 50))
    s2 = oracle.orbital_value("s", 0.7, x, y, z) ** 2
    for l, names in SHELLS.items():
        tot = sum(oracle.orbital_value(n, 0.7, x, y, z) ** 2 for n in names)
        assert np.allclose(tot, (2 * l + 1) * s2, rtol=1e-13)


def test_orbital_tile_grid_and_rotation():
    w, os_, h = 2, 4, 0.3
    c = os_ * w + os_ // 2  # x = c/os - w - 1/2 = 0
    t = oracle.orbital_tile("s", 0.5, h, 0.0, 0.0, 0.0, 2.0, w, os_)
    assert np.unravel_index(np.argmax(t), t.shape) == (c, c)
    assert abs(t[c, c] - 2.0 * np.exp(-h / 0.5) / np.sqrt(4 * np.pi)) < 1e-15
    # a quarter turn maps px onto py; an eighth turn maps dxy onto -dx2
    assert np.allclose(oracle.orbital_tile("px", 0.5, h, 0, 0, np.pi / 2, 1, w, os_),
                       oracle.orbital_tile("py", 0.5, h, 0, 0, 0.0, 1, w, os_), atol=1e-15)
    assert np.allclose(oracle.orbital_tile("dxy", 0.5, h, 0, 0, np.pi / 4, 1, w, os_),
                       -oracle.orbital_tile("dx2", 0.5, h, 0, 0, 0.0, 1, w, os_), atol=1e-15)
    # a shift by one sample moves the tile by one index
    a = oracle.orbital_tile("dxz", 0.5, h, 1.0 / os_, 0, 0.0, 1, w, os_)
    b = oracle.orbital_tile("dxz", 0.5, h, 0.0, 0, 0.0, 1, w, os_)
    assert np.allclose(a[1:], b[:-1], atol=1e-15)


@pytest.mark.parametrize("name", NAMES + ["n", "z"])
def test_library_orbital_tile_matches_oracle(name):
    for theta, xp, yp, h in [(0.0, 0.0, 0.0, 5.0), (0.3, 0.25, -0.5, 0.7), (np.pi / 3, 0.0, 0.1, -0.4)]:
        got = cqpi.orbital_tile(name, 0.8, h, xp, yp, theta, 1.5, 2, 4)
        ref = oracle.orbital_tile(name, 0.8, h, xp, yp, theta, 1.5, 2, 4)
        assert np.max(np.abs(got - ref)) <= 1e-14 * max(np.max(np.abs(ref)), 1e-300)


def test_library_orbital_rejects_unknown_names():
    with pytest.raises(cqpi.CqpiError):
        cqpi.orbital_tile("dz2")
    assert abs(lib_range := cqpi.lib().cqpi_orbital_range(0.5, 0.01)) > 0
    assert abs(lib_range - 0.5 * np.sqrt(np.log(100.0))) < 1e-15


def _write_xsf(path, cube, origin, span, conv=np.eye(3), atoms_z=(1.0, 2.0)):
    nx, ny, nz = cube.shape
    with open(path, "w") as f:
        f.write("# test orbital\nCRYSTAL\nPRIMVEC\n")
        for r in np.eye(3):
            f.write("%g %g %g\n" % tuple(r))
        f.write("CONVVEC\n")
        for r in conv:
            f.write("%.17g %.17g %.17g\n" % tuple(r))
        f.write("PRIMCOORD\n%d 1\n" % len(atoms_z))
        for z in atoms_z:
            f.write("X 0.0 0.0 %.17g\n" % z)
        f.write("BEGIN_BLOCK_DATAGRID_3D\n3D_field\nBEGIN_DATAGRID_3D_UNKNOWN\n%d %d %d\n" % (nx, ny, nz))
        f.write("%.17g %.17g %.17g\n" % tuple(origin))
        for r in span:
            f.write("%.17g %.17g %.17g\n" % tuple(r))
        vals = cube.transpose(2, 1, 0).ravel()  # x fastest
        for q in range(0, len(vals), 6):
            f.write(" ".join("%.17g" % v for v in vals[q:q + 6]) + "\n")
        f.write("END_DATAGRID_3D\nEND_BLOCK_DATAGRID_3D\n")


def test_xsf_tile_grid_pins(tmp_path):
    """A grid whose nodes sit on the tile samples (spacing 1/os, origin at
    (-w, -w)): the tile is the node plane at zh = top atom + height; CONVVEC
    rows swapped transpose it; a negative height counts from the lowest atom."""
    w, os_, nz, dz = 1, 2, 8, 0.5
    T = (2 * w + 1) * os_
    i, j, k = np.meshgrid(np.arange(T), np.arange(T), np.arange(nz), indexing="ij")
    cube = i + 100.0 * j + 1e4 * k
    span = np.diag([T / os_, T / os_, nz * dz])
    p = str(tmp_path / "a.xsf")
    _write_xsf(p, cube, (-w, -w, 0.0), span)
    ti, tj = np.meshgrid(np.arange(T), np.arange(T), indexing="ij")
    assert np.array_equal(oracle.xsf_tile(p, 0.5, w, os_, 2.0), 2.0 * (ti + 100.0 * tj + 5e4))   # zh 2.5: k 5
    assert np.array_equal(oracle.xsf_tile(p, -0.5, w, os_), ti + 100.0 * tj + 1e4)              # zh 0.5: k 1
    q = str(tmp_path / "b.xsf")
    _write_xsf(q, cube, (-w, -w, 0.0), span, conv=np.array([[0.0, 1, 0], [1, 0, 0], [0, 0, 1]]))
    assert np.array_equal(oracle.xsf_tile(q, 0.5, w, os_), (ti + 100.0 * tj + 5e4).T)
    r = str(tmp_path / "c.xsf")
    _write_xsf(r, cube, (-w + 1.0, -w, 0.0), span)   # the first 1/os * os = 2 rows fall outside
    t = oracle.xsf_tile(r, 0.5, w, os_)
    assert np.all(t[:2] == 0) and np.array_equal(t[2:], ti[:-2] + 100.0 * tj[2:] + 5e4)


def test_library_xsf_tile_matches_oracle(tmp_path):
    rng = np.random.default_rng(5)
    cube = rng.normal(size=(17, 13, 9))
    span = np.array([[3.1, 0.2, 0.0], [-0.4, 2.7, 0.1], [0.05, 0.0, 4.3]])
    conv = np.array([[1.0, 0.0, 0.0], [0.5, np.sqrt(3) / 2, 0.0], [0.0, 0.0, 6.0]])   # hexagonal
    p = str(tmp_path / "h.xsf")
    _write_xsf(p, cube, (-1.3, -1.1, 0.2), span, conv=conv, atoms_z=(0.7, 1.9, 1.2))
    for h in (1.0, -0.3):
        got = cqpi.xsf_tile(p, h, 2, 3, 1.5)
        ref = oracle.xsf_tile(p, h, 2, 3, 1.5)
        assert np.count_nonzero(ref) > 10 and np.array_equal(got, ref)
    with pytest.raises(cqpi.CqpiError):
        cqpi.xsf_tile(str(tmp_path / "missing.xsf"))
