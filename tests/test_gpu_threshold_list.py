"""GPU parity of the threshold-list contraction (GPUQPI::wannierldoslist,
src/qpihipkernel.cc:624-636) against the FP64 oracle restatement of
src/qpi.cc:1088-1135, called through the C ABI, with lists built by the
driver's rule (src/qpi.cc:139-211)."""
import numpy as np
import pytest

import cqpi
import oracle
import refmath as rm
from cqpi import synth

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def _g0(wl, omega):
    ev, V = oracle.calchamiltonian(wl.hops, wl.bands, wl.efermi, wl.K)
    gR = oracle.realspace(oracle.greens_k(ev, V, omega, wl.eta))
    return gR, oracle.tmatrix(wl.V, gR[0, 0])


def _dropin(wl, flags, counts, gR, T, prec="fp64"):
    q = cqpi.AmdQPI.from_list(wl.W, wl.K, wl.n, wl.window, wl.oversamp, wl.bands, flags, counts)
    try:
        q.set_precision(prec)
        q.wannierldoslist(T, gR)
        gR[:] = 0  # the caller may overwrite g0 once the call returned
        return q.retrieveResult()
    finally:
        q.close()


@pytest.mark.parametrize("case", ["3orb-w2-os4", "1orb-w1-os8", "5orb-w1-os2"])
@pytest.mark.parametrize("threshold", [0.05, -1.0])
def test_dropin_list_matches_oracle(gpu, case, threshold):
    b, w, os_ = {"3orb-w2-os4": (3, 2, 4), "1orb-w1-os8": (1, 1, 8), "5orb-w1-os2": (5, 1, 2)}[case]
    wl = synth.square_norb(b, K=32, n=10, E=1, seed=80 + b, w=w, os_=os_, eta=0.01)
    gR, T = _g0(wl, 0.04)
    flags, counts = cqpi.make_threshold_list(wl.tiles, w, os_, threshold)
    print("case: threshold list %s threshold=%g entries=%d" % (case, threshold, len(flags)), flush=True)
    ref = oracle.wannierldoslist(gR, T, flags, counts, wl.n, w, os_)
    got = _dropin(wl, flags, counts, gR.copy(), T)
    assert np.abs(ref).max() > 0
    assert rm.rel_errors(got, ref)[0] < TOL64
    if threshold < 0:
        # every term kept: the dense map
        dense = oracle.wannierldos(gR, T, wl.tiles, wl.n, w, os_)
        assert rm.rel_errors(got, dense)[0] < TOL64


def test_dropin_list_spin_and_duplicates(gpu):
    """A spin list (both halves), with some entries repeated: every entry
    adds its own term."""
    wl = synth.square_norb(5, K=32, n=8, E=1, seed=86, w=1, os_=2, eta=0.01)
    tiles = wl.tiles[:4]
    wl4 = synth.Workload(name="4 of 5", bands=4, hops=wl.hops[(wl.hops["o1"] < 4) & (wl.hops["o2"] < 4)],
                         K=wl.K, n=wl.n, window=1, oversamp=2, energies=wl.energies, eta=wl.eta,
                         V=wl.V[:4, :4], tiles=tiles)
    gR, T = _g0(wl4, 0.02)
    flags, counts = cqpi.make_threshold_list(tiles, 1, 2, 0.03, spin=True)
    # repeat the first entry of every non-empty list
    offs = np.concatenate([[0], np.cumsum(counts)]).astype(int)
    parts, cnt = [], counts.copy()
    for p in range(len(counts)):
        seg = flags[offs[p]:offs[p + 1]]
        if len(seg):
            seg = np.concatenate([seg, seg[:1]])
            cnt[p] += 1
        parts.append(seg)
    flags2 = np.concatenate(parts)
    ref = oracle.wannierldoslist(gR, T, flags2, cnt, wl4.n, 1, 2)
    got = _dropin(wl4, flags2, cnt, gR.copy(), T)
    assert rm.rel_errors(got, ref)[0] < TOL64


def test_dropin_spinarr_list(gpu):
    """A list built by the spin-label rule (src/qpi.cc:213-253, driver key
    spinarr): orbitals 0 and 2 share label 1, orbital 1 has label 2, orbital
    3 none; the map against the oracle's list contraction of the same list."""
    wl = synth.square_norb(4, K=32, n=8, E=1, seed=88, w=1, os_=2, eta=0.01)
    gR, T = _g0(wl, 0.03)
    flags, counts = cqpi.make_threshold_list_spinarr(wl.tiles, 1, 2, 0.01, [1, 2, 1, 0])
    print("case: spinarr list entries=%d" % len(flags), flush=True)
    assert len(flags) > 0 and not (flags["o1"] == 3).any()
    ref = oracle.wannierldoslist(gR, T, flags, counts, wl.n, 1, 2)
    got = _dropin(wl, flags, counts, gR.copy(), T)
    assert np.abs(ref).max() > 0
    assert rm.rel_errors(got, ref)[0] < TOL64


def test_full_chain_list(gpu):
    """cqpi_run on a list context: the device chain up to T, then the list map."""
    wl = synth.square_norb(3, K=32, n=8, E=3, seed=87, w=2, os_=4, eta=0.01, elo=-0.1, ehi=0.1)
    flags, counts = cqpi.make_threshold_list(wl.tiles, 2, 4, 0.04)
    q = cqpi.AmdQPI.from_list(wl.W, wl.K, wl.n, wl.window, wl.oversamp, wl.bands, flags, counts)
    try:
        q.set_model(wl.hops, wl.efermi, wl.scmodel)
        q.set_impurity(wl.V)
        got = q.run(wl.energies, wl.eta)
    finally:
        q.close()
    for e, om in enumerate(wl.energies):
        gR, T = _g0(wl, om)
        ref = oracle.wannierldoslist(gR, T, flags, counts, wl.n, 2, 4)
        assert rm.rel_errors(got[e], ref)[0] < TOL64, e


def test_list_context_rejects_other_uses(gpu):
    wl = synth.square_norb(3, K=16, n=4, E=1, seed=88, w=1, os_=2)
    flags, counts = cqpi.make_threshold_list(wl.tiles, 1, 2, 0.05)
    gR, T = _g0(wl, 0.03)
    q = cqpi.AmdQPI.from_list(wl.W, wl.K, wl.n, 1, 2, 3, flags, counts)
    try:
        with pytest.raises(cqpi.CqpiError):
            q.wannierldos(T, gR)
        with pytest.raises(cqpi.CqpiError):
            q.set_precision("mixed")   # no matrix product to put on MFMA
    finally:
        q.close()
    bad = flags.copy()
    bad["o1"][0] = 3
    with pytest.raises(cqpi.CqpiError):
        cqpi.AmdQPI.from_list(wl.W, wl.K, wl.n, 1, 2, 3, bad, counts)
    with pytest.raises(cqpi.CqpiError):
        cqpi.AmdQPI.from_list(wl.W, wl.K, wl.n, 1, 2, 17, flags, counts)
    # a plain context has no list
    q = cqpi.AmdQPI(wl.W, wl.K, wl.n, 1, 2, 3, 3, False, wl.tiles)
    try:
        with pytest.raises(cqpi.CqpiError):
            q.wannierldoslist(T, gR)
    finally:
        q.close()


def _hex_small():
    wl = synth.hex_5d_spin(K=16, n=6, E=1)
    wl.window, wl.oversamp = 1, 2
    wl.tiles = np.concatenate([synth.gaussian_tiles(["xz", "yz", "xy", "x2y2", "z2"], 0.4, 0.5, 1, 2, "hex")] * 2)
    return wl


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("case", ["6orb", "hex10-spin", "3orb"])
@pytest.mark.parametrize("threshold", [0.05, -1.0])
def test_dropin_list_many_bands_and_fp32(gpu, case, prec, threshold):
    """Band counts beyond the register kernel (6; 10 with spin halves, the
    hexagonal 5 d x spin model) and the FP32 list map (windows and T narrowed)
    against the oracle: FP64 1e-10, FP32 1e-4."""
    if case == "6orb":
        wl = synth.square_norb(6, K=32, n=8, E=1, seed=90, w=1, os_=2, eta=0.02)
        spin = False
    elif case == "3orb":
        wl = synth.square_norb(3, K=32, n=10, E=1, seed=83, w=2, os_=4, eta=0.01)
        spin = False
    else:
        wl = _hex_small()
        spin = True
    gR, T = _g0(wl, 0.03)
    flags, counts = cqpi.make_threshold_list(wl.tiles, wl.window, wl.oversamp, threshold, spin=spin)
    print("case: threshold list %s %s threshold=%g entries=%d" % (case, prec, threshold, len(flags)), flush=True)
    ref = oracle.wannierldoslist(gR, T, flags, counts, wl.n, wl.window, wl.oversamp)
    got = _dropin(wl, flags, counts, gR.copy(), T, prec)
    tol = TOL64 if prec == "fp64" else 1e-4
    assert np.abs(ref).max() > 0
    assert rm.rel_errors(got, ref)[0] < tol
    if threshold < 0:
        dense = oracle.wannierldos(gR, T, wl.tiles, wl.n, wl.window, wl.oversamp, spin=spin)
        assert rm.rel_errors(got, dense)[0] < tol


def test_full_chain_list_many_bands_fp32(gpu):
    """cqpi_run on a 6-band list context in FP32 (FP32 transform lines, FP32
    list map) against the oracle chain at 1e-4."""
    wl = synth.square_norb(6, K=64, n=8, E=2, seed=91, w=1, os_=2, eta=0.02, elo=-0.1, ehi=0.1)
    flags, counts = cqpi.make_threshold_list(wl.tiles, 1, 2, 0.04)
    q = cqpi.AmdQPI.from_list(wl.W, wl.K, wl.n, wl.window, wl.oversamp, wl.bands, flags, counts)
    try:
        q.set_model(wl.hops, wl.efermi, wl.scmodel)
        q.set_impurity(wl.V)
        q.set_precision("fp32")
        got = q.run(wl.energies, wl.eta)
    finally:
        q.close()
    for e, om in enumerate(wl.energies):
        gR, T = _g0(wl, om)
        ref = oracle.wannierldoslist(gR, T, flags, counts, wl.n, 1, 2)
        assert rm.rel_errors(got[e], ref)[0] < 1e-4, e
