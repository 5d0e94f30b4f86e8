"""The GPUQPI plug-in shim (calcqpi-release_amd/host/gpuqpi_amd_shim.h), the
header a calcQPI maintainer includes, compiled into a C++ driver
(tests/shim/shim_check.cc, with test-only stand-ins for the calcQPI headers
it includes) and run as the reference's tmatrix calls the plug-in: the
wannierldos constructor and call with GSL matrices of leading dimension > b,
g0 overwritten after the call returns, the three retrieveResult overloads,
the threshold-list constructor with wannierldoslist, and the spectral
constructor with spf.  Maps against the FP64 oracle at 1e-10."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import refmath as rm
from cqpi import synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "shim", "shim_check")


def test_shim_drives_the_library_like_the_reference(gpu, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("tests/shim/shim_check not built (__graft_entry__.build())")
    wl = synth.square_3orb(K=32, n=12, E=1)
    b, K = wl.bands, wl.K
    ev, V = oracle.calchamiltonian(wl.hops, b, wl.efermi, K)
    g0k = oracle.greens_k(ev, V, wl.energies[0], wl.eta)
    gR = oracle.realspace(g0k)
    T = oracle.tmatrix(wl.V, gR[0, 0])
    threshold = 0.01
    flags, counts = oracle.mkthresholdlist(wl.tiles, wl.window, wl.oversamp, threshold)
    d = str(tmp_path)
    with open(os.path.join(d, "params.txt"), "w") as f:
        f.write("%d %d %d %d %d %d %d %d %r\n" % (wl.W, K, wl.n, wl.window, wl.oversamp, b, wl.maxband,
                                                  int(wl.spin), threshold))
    np.ascontiguousarray(wl.tiles, np.float64).tofile(os.path.join(d, "tiles.bin"))
    np.ascontiguousarray(gR, np.complex128).view(np.float64).tofile(os.path.join(d, "g0.bin"))
    np.ascontiguousarray(T, np.complex128).view(np.float64).tofile(os.path.join(d, "scat.bin"))
    np.ascontiguousarray(g0k, np.complex128).view(np.float64).tofile(os.path.join(d, "g0k.bin"))
    counts.astype(np.float64).tofile(os.path.join(d, "counts.bin"))
    fl = np.stack([flags["i"], flags["j"], flags["o1"], flags["o2"], flags["factor"]], axis=1).astype(np.float64)
    fl.tofile(os.path.join(d, "flags.bin"))
    print("case: shim driver, 3 orbitals K=32 n=12 (ldos, 3 retrieveResult overloads, list, spf)", flush=True)
    r = subprocess.run([EXE, d], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0 and "shim_check done" in r.stdout
    assert "HIP GPU" in r.stdout   # printinfo went through the shim

    ref = oracle.wannierldos(gR, T, wl.tiles, wl.n, wl.window, wl.oversamp)
    outs = [np.fromfile(os.path.join(d, f)) for f in ("out_ptr.bin", "out_idl.bin", "out_vec.bin")]
    for o in outs:
        assert rm.rel_errors(o, ref)[0] < 1e-10
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])

    lref = oracle.wannierldoslist(gR, T, flags, counts, wl.n, wl.window, wl.oversamp)
    lst = np.fromfile(os.path.join(d, "out_list.bin"))
    assert rm.rel_errors(lst, lref)[0] < 1e-10

    sref = oracle.spf(g0k, K)
    spf = np.fromfile(os.path.join(d, "out_spf.bin"))
    assert rm.rel_errors(spf, sref)[0] < 1e-10
