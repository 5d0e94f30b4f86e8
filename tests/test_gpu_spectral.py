"""GPU parity of the spectral-function maps (GPUQPI::spf and ::uspf, the
whole-chain output=spf|uspf) and of the density of states against the FP64
oracle (src/qpi.cc:657-681, :731-775, :856-862), called through the C ABI."""
import numpy as np
import pytest

import cqpi
import oracle
import refmath as rm
from cqpi import synth

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def _model(b, K, seed):
    if b in synth.ORB_SETS:
        return synth.square_norb(b, K=K, n=K, E=1, seed=seed)
    wl = synth.square_norb(5, K=K, n=K, E=1, seed=seed)
    keep = (wl.hops["o1"] < b) & (wl.hops["o2"] < b)
    return synth.Workload(name="%d of 5" % b, bands=b, hops=wl.hops[keep], K=K, n=K, window=1, oversamp=1,
                          energies=wl.energies, eta=wl.eta, V=wl.V[:b, :b], tiles=wl.tiles[:b])


def _g0k(wl, om, eta):
    ev, V = oracle.calchamiltonian(wl.hops, wl.bands, wl.efermi, wl.K)
    return oracle.greens_k(ev, V, om, eta)


@pytest.mark.parametrize("mode,n,spin", [("spf", 16, False), ("spf", 10, False), ("uspf", 16, False),
                                         ("uspf", 12, False), ("uspf", 16, True)])
def test_dropin_spectral_matches_oracle(gpu, mode, n, spin):
    b = 4 if spin else 3
    wl = _model(b, 16, 100 + b)
    rng = np.random.default_rng(7)
    pos = rng.uniform(-0.5, 0.5, (b, 2))
    pre = rng.uniform(0.5, 1.5, b)
    g0k = _g0k(wl, 0.07, 0.02)
    print("case: %s drop-in b=%d K=16 n=%d spin=%s" % (mode, b, n, spin), flush=True)
    q = cqpi.AmdQPI.spectral(16, n, b, b, spin, pos, pre)
    try:
        getattr(q, mode)(g0k)
        g0k[:] = 0  # copied before the call returned
        got = q.retrieveResult()
    finally:
        q.close()
    g0k = _g0k(wl, 0.07, 0.02)
    ref = oracle.spf(g0k, n) if mode == "spf" else oracle.uspf(g0k, n, pos, pre, spin=spin)
    assert np.abs(ref).max() > 0
    assert rm.rel_errors(got, ref)[0] < TOL64


@pytest.mark.parametrize("mode", ["spf", "uspf"])
def test_full_chain_spectral_matches_oracle(gpu, mode):
    wl = _model(3, 32, 105)
    pos = np.array([[0.0, 0.0], [0.5, 0.0], [0.0, 0.5]])
    pre = np.array([1.0, 0.8, 1.2])
    om = np.array([-0.2, 0.05, 0.3])
    q = cqpi.AmdQPI.spectral(32, 32, 3, 3, False, pos, pre)
    try:
        q.set_model(wl.hops, wl.efermi, False)
        q.set_spectral(mode)
        got = q.run(om, 0.02)
    finally:
        q.close()
    for e, w in enumerate(om):
        g0k = _g0k(wl, w, 0.02)
        ref = oracle.spf(g0k, 32) if mode == "spf" else oracle.uspf(g0k, 32, pos, pre)
        assert rm.rel_errors(got[e], ref)[0] < TOL64, e


def test_dos_matches_oracle(gpu):
    wl = _model(3, 32, 106)
    om = np.linspace(-0.5, 0.5, 7)
    with cqpi.engine_for(wl, "fp64") as q:
        got = q.dos(om, 0.03)
    ref = np.array([oracle.dos(oracle.realspace(_g0k(wl, w, 0.03))) for w in om])
    assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref))
    # maxband < bands: the trace over the first maxband orbitals
    q = cqpi.AmdQPI.spectral(32, 32, 3, 2)
    try:
        q.set_model(wl.hops, wl.efermi, False)
        got2 = q.dos(om, 0.03)
    finally:
        q.close()
    ref2 = np.array([oracle.dos(oracle.realspace(_g0k(wl, w, 0.03)), 2) for w in om])
    assert np.max(np.abs(got2 - ref2)) < 1e-12 * np.max(np.abs(ref2))


def test_spectral_context_rejects_other_uses(gpu):
    wl = _model(3, 16, 107)
    g0k = _g0k(wl, 0.0, 0.02)
    q = cqpi.AmdQPI.spectral(16, 16, 3)
    try:
        q.set_model(wl.hops, wl.efermi, False)
        with pytest.raises(cqpi.CqpiError):
            q.run([0.1], 0.02)  # no map selected
        with pytest.raises(cqpi.CqpiError):
            q.wannierldos(np.eye(3), g0k)
    finally:
        q.close()
    with pytest.raises(cqpi.CqpiError):
        cqpi.AmdQPI.spectral(16, 20, 3)  # n > kpoints
    q = cqpi.AmdQPI(wl.W, wl.K, wl.n, wl.window, wl.oversamp, 3, 3, False, wl.tiles)
    try:
        with pytest.raises(cqpi.CqpiError):
            q.spf(g0k)
        with pytest.raises(cqpi.CqpiError):
            q.set_spectral("spf")
    finally:
        q.close()
