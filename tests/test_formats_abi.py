"""CPU tests of the host side: the C ABI library loads and exports every
symbol include/cqpi_amd.h declares; the library's _hr.dat reader agrees with
the oracle's restatement; idl files round-trip; error paths report codes."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import cqpi
import oracle
from cqpi import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "cqpi_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cqpi_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = cqpi.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(cqpi.SIGNATURES), set(names) ^ set(cqpi.SIGNATURES)


def test_hopping_struct_layout():
    assert synth.HOP_DTYPE.itemsize == 56
    assert synth.HOP_DTYPE == oracle.HOP_DTYPE


def test_no_device_is_reported_not_faked():
    if cqpi.device_count() > 0:
        pytest.skip("device present")
    T = np.zeros((1, 3, 3))
    with pytest.raises(cqpi.CqpiError) as ei:
        cqpi.AmdQPI(4, 4, 4, 1, 1, 1, 1, False, T)
    assert ei.value.code == 6


@pytest.mark.parametrize("cfg", [1, 2, 3])
def test_hr_reader_matches_oracle(tmp_path, cfg):
    wl = {1: synth.square_1orb, 2: lambda: synth.square_3orb(K=8, n=4),
          3: lambda: synth.hex_5d_spin(K=8, n=4)}[cfg]()
    p = str(tmp_path / "m_hr.dat")
    synth.write_hr(p, wl)
    h1, b1, n1 = cqpi.read_hr(p)
    h2, b2, n2 = oracle.read_hr(p)
    assert b1 == b2 == wl.bands and n1 == n2 == wl.name
    assert len(h1) == len(h2) == len(wl.hops)
    for f in ("rx", "ry", "rz", "o1", "o2", "tre", "tim", "anti"):
        assert np.array_equal(h1[f], h2[f]), f
    # amplitudes survive the text format to the last digit written (%.16e)
    assert np.allclose(h1["tre"], wl.hops["tre"], rtol=1e-15, atol=1e-300)


def test_hr_degeneracy_and_grouping(tmp_path):
    """Degeneracies spread over several lines divide the amplitudes of their R;
    R changes are detected line by line."""
    p = tmp_path / "deg_hr.dat"
    lines = ["test model", "1", "3", "  2", " 4    1", "0 0 0 1 1 0.8 0.0", "1 0 0 1 1 -1.0 0.0",
             "1 0 0 1 1 -0.5 0.2", "0 1 0 1 1 2.0 0.0", ""]
    p.write_text("\n".join(lines))
    for reader in (cqpi.read_hr, oracle.read_hr):
        h, b, name = reader(str(p))
        assert b == 1 and name == "test model"
        assert np.allclose(h["tre"], [0.4, -0.25, -0.125, 2.0])
        assert np.allclose(h["tim"], [0.0, 0.0, 0.05, 0.0])


def test_hr_scmodel_anti_flag(tmp_path):
    p = tmp_path / "sc_hr.dat"
    lines = ["sc", "4", "1", "1"] + ["0 0 0 %d %d 1.0 0.0" % (u, v) for u in range(1, 5) for v in range(1, 5)]
    p.write_text("\n".join(lines) + "\n")
    for reader in (cqpi.read_hr, oracle.read_hr):
        h, _, _ = reader(str(p), scmodel=True)
        anti = {(int(a), int(b)): int(f) for a, b, f in zip(h["o1"], h["o2"], h["anti"])}
        assert anti[(2, 3)] == 1 and anti[(0, 3)] == 0 and anti[(1, 1)] == 0


@pytest.mark.parametrize("bad", ["badorb", "short", "missing"])
def test_hr_errors(tmp_path, bad):
    p = tmp_path / "bad_hr.dat"
    if bad == "badorb":
        p.write_text("m\n1\n1\n1\n0 0 0 2 1 1.0 0.0\n")
    elif bad == "short":
        p.write_text("m\n1\n1\n1\n0 0 0 1 1 1.0\n")
    else:
        p = tmp_path / "does_not_exist"
    with pytest.raises(cqpi.CqpiError) as ei:
        cqpi.read_hr(str(p))
    assert ei.value.code in (1, 5)


def test_idl_roundtrip_between_writers(tmp_path):
    rng = np.random.default_rng(0)
    d = rng.normal(size=(5, 4, 3)).astype(np.float32).astype(np.float64)
    p1, p2 = str(tmp_path / "a.idl"), str(tmp_path / "b.idl")
    # file order: scan range (2), offset (2), bias range (2), setpoint -- as the
    # cLDOS output map idl(W,W,L,2.0,2.0,lb,hb,0.0,-1.0,-1.0) writes it
    hdr = (2.0, 2.0, -1.0, -1.0, -0.1, 0.1, 0.0)
    synth.write_idl(p1, d, name="map", date="today", hdr=hdr)
    oracle.idl_write(p2, d, name="map", date="today", hdr=hdr)
    assert open(p1, "rb").read() == open(p2, "rb").read()
    got, h = synth.read_idl(p1)
    assert np.array_equal(got, d) and h["name"] == "map" and h["bias"] == [-0.1, 0.1]
    got2, h2 = oracle.idl_read(p2)
    assert np.array_equal(got2, d) and list(h2) == list(hdr)


def test_idl_bytes_follow_the_file_format(tmp_path):
    """Known-answer bytes written out by hand from the format of idl::write
    (src/idl.cc:78-95): name, date, the three sizes, scan range x and y,
    offset x and y, junction bias low and high, setpoint, one per line as
    an ostream prints doubles (6 significant digits), then float32 values with
    x fastest, then y, then the layer.  The header values are those of the
    driver's map idl(n, n, L, 2.0, 2.0, lb, hb, 0.0, -1.0, -1.0)
    (src/calcqpi.cc:655; constructor src/idl.cc:662-674), so the file order
    puts the offset (-1, -1) before the bias (lb, hb).  Both writers must
    produce these bytes and both readers must read them back."""
    import struct
    sx, sy, sz = 3, 2, 2
    head = b"map\nThu Dec  3 10:00:00 2026\n3\n2\n2\n2\n2\n-1\n-1\n-0.125\n0.125\n0\n"
    vals = [100.0 * k + 10.0 * j + i for k in range(sz) for j in range(sy) for i in range(sx)]
    expect = head + struct.pack("<12f", *vals)
    d = np.zeros((sx, sy, sz))
    for i in range(sx):
        for j in range(sy):
            for k in range(sz):
                d[i, j, k] = 100.0 * k + 10.0 * j + i
    hdr = (2.0, 2.0, -1.0, -1.0, -0.125, 0.125, 0.0)
    p1, p2 = str(tmp_path / "a.idl"), str(tmp_path / "b.idl")
    synth.write_idl(p1, d, name="map", date="Thu Dec  3 10:00:00 2026", hdr=hdr)
    oracle.idl_write(p2, d, name="map", date="Thu Dec  3 10:00:00 2026", hdr=hdr)
    assert open(p1, "rb").read() == expect
    assert open(p2, "rb").read() == expect
    p3 = tmp_path / "c.idl"
    p3.write_bytes(expect)
    got, h = synth.read_idl(str(p3))
    assert np.array_equal(got, d)
    assert (h["scan_range"], h["offset"], h["bias"], h["setpoint"]) == ([2.0, 2.0], [-1.0, -1.0], [-0.125, 0.125], 0.0)
    got2, h2 = oracle.idl_read(str(p3))
    assert np.array_equal(got2, d) and list(h2) == list(hdr)


def test_write_inputs(tmp_path):
    wl = synth.square_3orb(K=16, n=8, E=5)
    p = synth.write_inputs(str(tmp_path), wl)
    txt = open(p).read()
    assert "scattering=(0.5,0.5,0.5);" in txt and "window=2;" in txt
    tiles, _ = synth.read_idl(str(tmp_path / "orbitals.idl"))
    assert np.array_equal(np.transpose(tiles, (2, 0, 1)), wl.tiles)


def test_flag_entry_layout():
    assert cqpi.FLAG_DTYPE.itemsize == 24 and cqpi.FLAG_DTYPE == oracle.FLAG_DTYPE


@pytest.mark.parametrize("threshold,spin", [(-1.0, False), (0.05, False), (0.02, True)])
def test_threshold_list_builder_matches_oracle(threshold, spin):
    """cqpi_make_threshold_list (host code of the library) builds the same
    entries, in the same order, as the oracle's mkthresholdlist."""
    wl = synth.square_norb(5 if spin else 3, K=8, n=4, E=1, seed=71, w=2, os_=4)
    tiles = wl.tiles[:4] if spin else wl.tiles
    got, gc = cqpi.make_threshold_list(tiles, 2, 4, threshold, spin=spin)
    ref, rc = oracle.mkthresholdlist(tiles, 2, 4, threshold, spin=spin)
    assert np.array_equal(gc, rc)
    assert np.array_equal(got, ref)
    assert len(got) > 0


@pytest.mark.parametrize("spinarr", [[1, 1, 2, 2], [2, 0, 2, 1], [1, 2], [3, 3, 3, 3, 3, 9]])
def test_threshold_list_spinarr_builder_matches_oracle(spinarr):
    """cqpi_make_threshold_list_spinarr builds the same entries, in the same
    order, as the oracle's restatement of mkthresholdlist(..., spinarr): labels
    resized to the orbital count, zero labels without entries."""
    tiles = synth.square_norb(4, K=8, n=4, E=1, seed=73, w=1, os_=3).tiles
    got, gc = cqpi.make_threshold_list_spinarr(tiles, 1, 3, 0.01, spinarr)
    ref, rc = oracle.mkthresholdlist_spinarr(tiles, 1, 3, 0.01, spinarr)
    assert np.array_equal(gc, rc)
    assert np.array_equal(got, ref)
    assert len(got) > 0
    # the selection rule: only equal nonzero labels pair up
    lab = (list(spinarr) + [0] * 4)[:4]
    assert all(lab[e["o1"]] != 0 and lab[e["o1"]] == lab[e["o2"]] for e in got)


def test_threshold_list_spinarr_pins():
    """Pins of the spin-label rule against the other list rule: one common
    label is the plain list over every orbital (same entries, same order);
    two halves labelled 1 and 2 hold the spin list's entries (its order
    interleaves the halves); all labels 0 give no entry."""
    tiles = synth.square_norb(4, K=8, n=4, E=1, seed=74, w=1, os_=2).tiles
    one, oc = oracle.mkthresholdlist_spinarr(tiles, 1, 2, 0.02, [1, 1, 1, 1])
    plain, pc = oracle.mkthresholdlist(tiles, 1, 2, 0.02)
    assert np.array_equal(one, plain) and np.array_equal(oc, pc)
    halves, hc = oracle.mkthresholdlist_spinarr(tiles, 1, 2, 0.02, [1, 1, 2, 2])
    spin, sc = oracle.mkthresholdlist(tiles, 1, 2, 0.02, spin=True)
    assert np.array_equal(hc, sc)
    key = lambda a: sorted((int(e["i"]), int(e["j"]), int(e["o1"]), int(e["o2"]), float(e["factor"])) for e in a)
    assert key(halves) == key(spin)
    none, nc = oracle.mkthresholdlist_spinarr(tiles, 1, 2, 0.02, [0, 0, 0, 0])
    assert len(none) == 0 and not nc.any()


def test_threshold_list_builder_rejects_short_buffer():
    wl = synth.square_norb(3, K=8, n=4, E=1, seed=71, w=1, os_=2)
    counts = np.zeros(81, np.uintp)
    total = C.c_size_t(0)
    flags = np.zeros(3, cqpi.FLAG_DTYPE)
    rc = cqpi.lib().cqpi_make_threshold_list(
        wl.tiles.ctypes.data_as(C.POINTER(C.c_double)), 3, 1, 2, 0.0, 0, 0, flags.ctypes.data, 3,
        counts.ctypes.data_as(C.POINTER(C.c_size_t)), C.byref(total))
    assert rc == 1 and total.value > 3


def test_plugin_shim_compiles_against_the_interface():
    """The GPUQPI shim header (the file a calcQPI maintainer includes) compiles
    with the interface's stand-in headers and -Wall -Wextra -Werror."""
    import shutil
    import subprocess
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("no g++")
    shim = os.path.join(ROOT, "tests", "shim")
    r = subprocess.run([cxx, "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-Wno-unused-parameter",
                        "-I", shim, "-I", os.path.join(ROOT, "include"), "-I",
                        os.path.join(ROOT, "calcqpi-release_amd", "host"), os.path.join(shim, "shim_check.cc")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
