"""GPU parity of the Josephson map (GPUQPI::wannierjosephson, the drop-in path
of src/qpihipkernel.cc:600-622) against the FP64 oracle restatement of
src/qpi.cc:953-1033, called through the C ABI.

G0(R) comes from a Bogoliubov-de Gennes model built from a seeded normal-state
model: H_BdG(k) = [[H(k), D], [D^H, -H(-k)^*]] with constant s-wave pairing D,
inverted per k and transformed by the pinned oracle transform, so the
particle-hole block the map reads is nonzero.

Precision modes: FP64 to 1e-10; FP32 and mixed (split-fp16 MFMA products
with FP32 accumulation) to 1e-4, in the metric of refmath.rel_errors.  The
wide-window case (w = 17) takes the per-lane gather kernels in FP64 and FP32
(the staged rows exceed the LDS) and the generic mixed kernel."""
import numpy as np
import pytest

import cqpi
import oracle
import refmath as rm
from cqpi import synth

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOL32 = 1e-4
TOL = {"fp64": TOL64, "fp32": TOL32, "mixed": TOL32}
_REF = {}


def _bdg(nb, K=16, n=8, seed=51, delta=0.08, omega=0.03, w=1):
    wl = synth.square_norb(nb, K=K, n=n, E=1, seed=seed, w=w, os_=2, eta=0.02)
    k = synth.kgrid(K)
    kx, ky = np.meshgrid(k, k, indexing="ij")
    H = synth.hk_numpy(wl.hops, nb, wl.efermi, kx, ky)
    Hm = synth.hk_numpy(wl.hops, nb, wl.efermi, -kx, -ky)
    b = 2 * nb
    Hb = np.zeros((K, K, b, b), np.complex128)
    Hb[..., :nb, :nb] = H
    Hb[..., nb:, nb:] = -np.conj(Hm)
    D = delta * np.eye(nb)
    Hb[..., :nb, nb:] = D
    Hb[..., nb:, :nb] = D.conj().T
    g0k = np.linalg.inv((omega + 1j * wl.eta) * np.eye(b) - Hb)
    gR = oracle.realspace(g0k)
    V = np.zeros((b, b), np.complex128)
    V[:nb, :nb] = wl.V
    V[nb:, nb:] = -np.conj(wl.V)
    T = oracle.tmatrix(V, gR[0, 0])
    return wl, gR, T


def _tip(omega=0.03, deltat=0.01, etat=0.02):
    # src/calcqpi.cc:705-709
    oc = complex(omega, etat)
    return complex(0.0, -np.copysign(1.0, omega) * np.pi * deltat) / np.sqrt(oc * oc - deltat * deltat)


def _run(wl, gR, T, tip, spin, prec="fp64"):
    b = gR.shape[2]
    print("case: josephson drop-in b=%d spin=%s K=%d n=%d prec=%s" % (b, spin, wl.K, wl.n, prec), flush=True)
    q = cqpi.AmdQPI(wl.W, wl.K, wl.n, wl.window, wl.oversamp, b, b // 2, spin, wl.tiles)
    try:
        if prec != "fp64":
            q.set_precision(prec)
        q.wannierjosephson(T, gR, tip)
        return q.retrieveResult()
    finally:
        q.close()


@pytest.mark.parametrize("nb,spin", [(1, False), (3, False), (6, True)])
def test_dropin_josephson_matches_oracle(gpu, nb, spin):
    wl, gR, T = _bdg(nb)
    tip = _tip()
    ref = oracle.wannierjosephson(gR, T, wl.tiles, wl.n, wl.window, wl.oversamp, tip, spin=spin)
    got = _run(wl, gR, T, tip, spin)
    assert np.abs(ref).max() > 0
    assert rm.rel_errors(got, ref)[0] < TOL64


def test_dropin_josephson_fp32(gpu):
    wl, gR, T = _bdg(3)
    tip = _tip(omega=-0.02)
    ref = oracle.wannierjosephson(gR, T, wl.tiles, wl.n, wl.window, wl.oversamp, tip)
    got = _run(wl, gR, T, tip, False, "fp32")
    assert rm.rel_errors(got, ref)[0] < TOL32


@pytest.mark.parametrize("nb,spin", [(1, False), (3, False), (6, True)])
def test_dropin_josephson_mixed(gpu, nb, spin):
    """The staged mixed-precision kernel with the b planes taken from the
    particle-hole columns o + maxband (spin: per spin block)."""
    wl, gR, T = _bdg(nb)
    tip = _tip(omega=-0.02)
    ref = oracle.wannierjosephson(gR, T, wl.tiles, wl.n, wl.window, wl.oversamp, tip, spin=spin)
    got = _run(wl, gR, T, tip, spin, "mixed")
    assert np.abs(ref).max() > 0
    assert rm.rel_errors(got, ref)[0] < TOL32


@pytest.mark.parametrize("prec", ["fp64", "fp32", "mixed"])
def test_dropin_josephson_wide_window(gpu, prec):
    """w = 17 with 6 bands: (8 + 2w)^2 staged window rows of 6 slots per map
    side exceed the LDS in FP64 (338 KB) and FP32 (169 KB), so the map runs
    on the per-lane gather kernel; mixed precision runs the generic (global
    B operand) kernel."""
    wl, gR, T = _bdg(3, K=48, w=17)
    tip = _tip()
    if "wide" not in _REF:  # ~13 s of oracle work, shared by the three modes
        _REF["wide"] = oracle.wannierjosephson(gR, T, wl.tiles, wl.n, wl.window, wl.oversamp, tip)
    ref = _REF["wide"]
    got = _run(wl, gR, T, tip, False, prec)
    assert np.abs(ref).max() > 0
    assert rm.rel_errors(got, ref)[0] < TOL[prec]


def test_dropin_josephson_then_ldos(gpu):
    """The two maps share one context: a Josephson call followed by an LDOS
    call gives each map its own periodic table."""
    wl, gR, T = _bdg(3)
    b = gR.shape[2]
    tip = _tip()
    q = cqpi.AmdQPI(wl.W, wl.K, wl.n, wl.window, wl.oversamp, b, b // 2, False, wl.tiles)
    try:
        q.wannierjosephson(T, gR, tip)
        gj = q.retrieveResult()
        q.wannierldos(T, gR)
        gl = q.retrieveResult()
    finally:
        q.close()
    rj = oracle.wannierjosephson(gR, T, wl.tiles, wl.n, wl.window, wl.oversamp, tip)
    rl = oracle.wannierldos(gR, T, wl.tiles, wl.n, wl.window, wl.oversamp, maxband=b // 2)
    assert rm.rel_errors(gj, rj)[0] < TOL64
    assert rm.rel_errors(gl, rl)[0] < TOL64


def test_josephson_rejects_what_it_cannot_compute(gpu):
    wl, gR, T = _bdg(1)
    b = gR.shape[2]
    # every band carries a tile: there is no particle-hole column o + maxband
    tiles = np.concatenate([wl.tiles, wl.tiles])
    q = cqpi.AmdQPI(wl.W, wl.K, wl.n, wl.window, wl.oversamp, b, b, False, tiles)
    try:
        with pytest.raises(cqpi.CqpiError):
            q.wannierjosephson(T, gR, 1.0)
    finally:
        q.close()


def _oracle_chain_josephson(wl, deltat, etat=None):
    """Per energy: oracle eigensolve, resolvent, transform and T-matrix of the
    superconducting model, then the Josephson map with the driver's tip
    factor (src/calcqpi.cc:705-709)."""
    ev, vec = oracle.calchamiltonian(wl.hops, wl.bands, wl.efermi, wl.K, scmodel=True)
    out = []
    for om in wl.energies:
        gR = oracle.realspace(oracle.greens_k(ev, vec, om, wl.eta))
        T = oracle.tmatrix(wl.V, gR[0, 0])
        tip = _tip(om, deltat, wl.eta if etat is None else etat)
        out.append(oracle.wannierjosephson(gR, T, wl.tiles, wl.n, wl.window, wl.oversamp, tip))
    return np.array(out)


@pytest.mark.parametrize("prec,etat", [("fp64", None), ("fp64", 0.004), ("fp32", None), ("mixed", None)])
def test_full_chain_josephson(gpu, prec, etat):
    """cqpi_run with the Josephson map selected: eigensolve, resolvent,
    transforms and T-matrix on the device, tip factor per energy."""
    wl = synth.bdg_square(3, K=32, n=8, E=4, seed=61)
    ref = _oracle_chain_josephson(wl, 0.01, etat)
    print("case: josephson full chain b=6 K=32 n=8 E=4 prec=%s etat=%s" % (prec, etat), flush=True)
    with cqpi.engine_for(wl, prec) as q:
        q.set_josephson(True, deltat=0.01, etat=float("nan") if etat is None else etat)
        got = q.run(wl.energies, wl.eta)
    tol = TOL[prec]
    for e in range(len(wl.energies)):
        assert np.abs(ref[e]).max() > 0
        assert rm.rel_errors(got[e], ref[e])[0] < tol, e


def test_full_chain_josephson_not_differentiated(gpu):
    wl = synth.bdg_square(1, K=16, n=4, E=2, seed=62)
    with cqpi.engine_for(wl, "fp64") as q:
        q.set_keep_windows(True)
        q.set_josephson(True, deltat=0.01)
        maps = q.run(wl.energies, wl.eta)
        with pytest.raises(cqpi.CqpiError):
            q.gradient(np.ones_like(maps))
