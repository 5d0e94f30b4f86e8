"""Pair symmetry of the transform passes: for models with H(k)^T = H(-k)
(real hoppings) the passes transform only the orbital pairs r <= c and write
G0(R)_cr = G0(-R)_rc.  The maps must match the oracle (which transforms every
pair, as src/qpi.cc:557-578 does) to the usual bounds, and the maps with the
symmetry switched off to rounding."""
import numpy as np
import pytest

import cqpi
import oracle
import refmath as rm
from cqpi import synth

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-10, "fp32": 1e-4, "mixed": 1e-4}
# on against off: FP64 differs by the rounding of the other pair's transform;
# FP32 lines by FP32 rounding of the same
TOL_ONOFF = {"fp64": 1e-12, "fp32": 2e-5, "mixed": 2e-5}


def _maps(wl, prec, sym):
    print("case: %s K=%d n=%d w=%d os=%d E=%d prec=%s sym=%d" % (
        wl.name, wl.K, wl.n, wl.window, wl.oversamp, len(wl.energies), prec, sym), flush=True)
    with cqpi.engine_for(wl, prec) as q:
        q.set_pair_symmetry(sym)
        rho = q.run(wl.energies, wl.eta)
        return rho, q.pair_symmetry()


def _close(got, ref, tol):
    for e in range(ref.shape[0]):
        mx, rms = rm.rel_errors(got[e], ref[e])
        assert mx < tol, "energy %d: max_rel %.3e rms_rel %.3e" % (e, mx, rms)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "mixed"])
def test_square_3orb_on_off_and_oracle(gpu, prec):
    # the 3-orbital workload's model, tiles, window and oversampling on a
    # 64^2 grid: one chunk of the 6 pairs r <= c
    wl = synth.square_3orb(K=64, n=16, E=3)
    on, (model, used) = _maps(wl, prec, True)
    assert model and used
    off, (model, used) = _maps(wl, prec, False)
    assert model and not used
    _close(on, off, TOL_ONOFF[prec])
    _close(on, oracle.cldos_run(wl), TOL[prec])


@pytest.mark.parametrize("b,K", [(2, 128), (4, 256), (5, 64)])
def test_band_counts_and_grids(gpu, b, K):
    # 3, 10 and 15 transformed pairs; 4 and 5 bands split them over chunks
    wl = synth.square_norb(b, K=K, n=8, E=2, seed=60 + b, w=1, os_=2, eta=0.02)
    on, (model, used) = _maps(wl, "fp64", True)
    assert model and used
    _close(on, oracle.cldos_run(wl), TOL["fp64"])
    on32, _ = _maps(wl, "fp32", True)
    off32, _ = _maps(wl, "fp32", False)
    _close(on32, off32, TOL_ONOFF["fp32"])


def test_workspace_path_k1024(gpu):
    # K = 1024: G0(k) through the workspace, read per transformed pair
    wl = synth.square_norb(3, K=1024, n=4, E=2, seed=71, w=1, os_=2, eta=0.02)
    for prec in ("fp64", "fp32"):
        on, (_, used) = _maps(wl, prec, True)
        assert used
        off, _ = _maps(wl, prec, False)
        _close(on, off, TOL_ONOFF[prec])


def test_complex_hoppings_keep_every_pair(gpu):
    # spin-orbit hoppings are complex: H(k)^T != H(-k), no symmetry
    wl = synth.hex_5d_spin(K=16, n=6, E=2)
    wl.window, wl.oversamp = 1, 2
    wl.tiles = np.concatenate([synth.gaussian_tiles(["xz", "yz", "xy", "x2y2", "z2"], 0.4, 0.5, 1, 2, "hex")] * 2)
    rho, (model, used) = _maps(wl, "fp64", True)
    assert not model and not used
    _close(rho, oracle.cldos_run(wl), TOL["fp64"])


def test_one_asymmetric_term_keeps_every_pair(gpu):
    # a real model with one hopping whose transpose at -R differs
    wl = synth.square_3orb(K=32, n=8, E=2)
    hops = wl.hops.copy()
    i = int(np.flatnonzero((hops["o1"] != hops["o2"]) & ((hops["rx"] != 0) | (hops["ry"] != 0)))[0])
    hops["tre"][i] += 0.05
    wl.hops = hops
    rho, (model, used) = _maps(wl, "fp64", True)
    assert not model and not used
    _close(rho, oracle.cldos_run(wl), TOL["fp64"])


def test_planes_path_keeps_every_pair(gpu):
    # mixed mode with a shape outside the T-folded contraction writes fp16
    # planes, which have no mirrored copies: every pair is transformed
    wl = synth.square_norb(3, K=64, n=8, E=2, seed=81, w=1, os_=2, eta=0.02)
    rho, (model, used) = _maps(wl, "mixed", True)
    assert model and not used
    _close(rho, oracle.cldos_run(wl), TOL["mixed"])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_particle_hole_model(gpu, prec):
    # the particle-hole model of real normal-state hoppings and real s-wave
    # pairing has H(k)^T = H(-k) as well: 21 of 36 pairs, LDOS and Josephson maps
    wl = synth.bdg_square(3, K=64, n=8, E=3, seed=61)
    on, (model, used) = _maps(wl, prec, True)
    assert model and used
    off, _ = _maps(wl, prec, False)
    _close(on, off, TOL_ONOFF[prec])
    _close(on, oracle.cldos_run(wl), TOL[prec])
    with cqpi.engine_for(wl, prec) as q:
        q.set_josephson(True, deltat=0.01)
        jon = q.run(wl.energies, wl.eta)
        assert q.pair_symmetry() == (True, True)
        q.set_pair_symmetry(False)
        joff = q.run(wl.energies, wl.eta)
    _close(jon, joff, TOL_ONOFF[prec])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_threshold_list_context(gpu, prec):
    # the list contraction reads the same window: symmetric passes, list map
    # against the oracle's list contraction of the oracle window
    wl = synth.square_norb(3, K=64, n=8, E=2, seed=87, w=2, os_=4, eta=0.01, elo=-0.1, ehi=0.1)
    flags, counts = cqpi.make_threshold_list(wl.tiles, 2, 4, 0.04)
    maps = {}
    for sym in (True, False):
        print("case: list context K=64 prec=%s sym=%d" % (prec, sym), flush=True)
        q = cqpi.AmdQPI.from_list(wl.W, wl.K, wl.n, wl.window, wl.oversamp, wl.bands, flags, counts)
        try:
            q.set_model(wl.hops, wl.efermi, wl.scmodel)
            q.set_impurity(wl.V)
            q.set_precision(prec)
            q.set_pair_symmetry(sym)
            maps[sym] = q.run(wl.energies, wl.eta)
            assert q.pair_symmetry() == (True, sym)
        finally:
            q.close()
    _close(maps[True], maps[False], TOL_ONOFF[prec])
    ev, V = oracle.calchamiltonian(wl.hops, wl.bands, wl.efermi, wl.K)
    for e, om in enumerate(wl.energies):
        gR = oracle.realspace(oracle.greens_k(ev, V, om, wl.eta))
        T = oracle.tmatrix(wl.V, gR[0, 0])
        ref = oracle.wannierldoslist(gR, T, flags, counts, wl.n, 2, 4)
        mx, rms = rm.rel_errors(maps[True][e], ref)
        assert mx < TOL[prec], "energy %d: max_rel %.3e" % (e, mx)
