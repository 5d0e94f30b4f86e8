"""GPU parity at the geometries of the measured workloads, against the FP64
oracle (restated reference algorithm, src/qpi.cc:454-950 with the spin branch
:924-940), through the whole device chain (cqpi_run) and through the GPUQPI
drop-in (cqpi_wannierldos, host G0(R) and T as src/qpi.cc:876-883 hands them
over).

* hexagonal 5 d orbitals x spin (10 bands, SOC, full Hermitian V) at the
  workload's own window w = 3 and oversampling os = 6 (a 16-row MFMA tile
  holds 12 live pixel rows), on small k-grids and maps, and at the
  workload's k-grid K = 1024 with a small map;
* the square 3-orbital workload at full size (K = 512, 128 x 128 cells,
  os = 4, w = 2) on two of its energies;
* the square 1-orbital workload at full size (K = 256, 64 x 64 cells, os = 8,
  w = 1) on two of its energies.

FP64 agrees to 1e-10, FP32 and mixed to 1e-4, per energy, in the metric
max|rho - rho_oracle| / max(|rho_oracle|, 1e-3 max|rho_oracle|)."""
import numpy as np
import pytest

import cqpi
import oracle
import refmath as rm
from cqpi import synth

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-10, "fp32": 1e-4, "mixed": 1e-4}


def _check(got, ref, prec, label):
    for e in range(ref.shape[0]):
        mx, rms = rm.rel_errors(got[e], ref[e])
        print("%s %s energy %d: max_rel %.3e rms_rel %.3e" % (label, prec, e, mx, rms), flush=True)
        assert mx < TOL[prec], (label, prec, e, mx, rms)


def _hex(K, n, E=2):
    wl = synth.hex_5d_spin(K=K, n=n, E=E)
    assert (wl.window, wl.oversamp, wl.bands, wl.spin) == (3, 6, 10, True)
    # two energies inside the workload's range, away from the range ends
    wl.energies = np.linspace(-0.13, 0.07, E)
    return wl


_REF = {}


def _oracle(key, wl):
    if key not in _REF:
        _REF[key] = oracle.cldos_run(wl)
    return _REF[key]


@pytest.mark.parametrize("prec", ["fp64", "mixed", "fp32"])
@pytest.mark.parametrize("K,n", [(16, 6), (32, 8)])
def test_hex_spin_w3_os6_chain(gpu, K, n, prec):
    wl = _hex(K, n)
    print("case: hex 10 bands spin w=3 os=6 K=%d n=%d E=2 prec=%s (cqpi_run)" % (K, n, prec), flush=True)
    with cqpi.engine_for(wl, prec) as q:
        got = q.run(wl.energies, wl.eta)
    _check(got, _oracle(("hex", K, n), wl), prec, "hex K=%d n=%d" % (K, n))


@pytest.mark.parametrize("prec", ["fp64", "mixed", "fp32"])
def test_hex_spin_w3_os6_dropin(gpu, prec):
    K, n = 32, 8
    wl = _hex(K, n)
    print("case: hex 10 bands spin w=3 os=6 K=%d n=%d prec=%s (drop-in)" % (K, n, prec), flush=True)
    ev, V = oracle.calchamiltonian(wl.hops, wl.bands, wl.efermi, wl.K)
    q = cqpi.AmdQPI(wl.W, wl.K, wl.n, wl.window, wl.oversamp, wl.bands, wl.maxband, wl.spin, wl.tiles)
    try:
        q.set_precision(prec)
        for e, om in enumerate(wl.energies):
            gR = oracle.realspace(oracle.greens_k(ev, V, om, wl.eta))
            T = oracle.tmatrix(wl.V, gR[0, 0])
            ref = oracle.wannierldos(gR, T, wl.tiles, wl.n, wl.window, wl.oversamp, spin=True)
            q.wannierldos(T, gR)
            got = q.retrieveResult()
            _check(got[None], ref[None], prec, "hex drop-in energy %d" % e)
    finally:
        q.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_hex_spin_workload_kgrid(gpu, prec):
    """The 10-band workload's k-grid (K = 1024: G(k) workspace path, Stockham
    passes of 1024 points) and tiles, with an 8 x 8-cell map and one energy."""
    wl = _hex(1024, 8, E=1)
    print("case: hex 10 bands spin w=3 os=6 K=1024 n=8 E=1 prec=%s" % prec, flush=True)
    ref = _oracle(("hex", 1024, 8), wl)
    with cqpi.engine_for(wl, prec) as q:
        got = q.run(wl.energies, wl.eta)
    _check(got, ref, prec, "hex K=1024 n=8")


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_square_3orb_full_size(gpu, prec):
    """The 3-orbital workload at full size on its first and middle energies."""
    wl = synth.square_3orb()
    wl.energies = wl.energies[[0, 25]]
    print("case: square 3 orbitals K=512 n=128 w=2 os=4 E=2 prec=%s" % prec, flush=True)
    ref = _oracle("c2full", wl)
    with cqpi.engine_for(wl, prec) as q:
        got = q.run(wl.energies, wl.eta)
    _check(got, ref, prec, "square3 full")


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_square_1orb_full_size(gpu, prec):
    wl = synth.square_1orb()
    wl.energies = wl.energies[[0, 5]]
    print("case: square 1 orbital K=256 n=64 w=1 os=8 E=2 prec=%s" % prec, flush=True)
    ref = _oracle("c1full", wl)
    with cqpi.engine_for(wl, prec) as q:
        got = q.run(wl.energies, wl.eta)
    _check(got, ref, prec, "square1 full")
