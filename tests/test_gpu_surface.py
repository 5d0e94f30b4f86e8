"""GPU parity of the surface / bulk Green's function (green=surface|bulk,
tmatrix::calcsurfacegreensfunction, src/qpi.cc:488-554) against the FP64
oracle chain surfacegreen -> realspace -> tmatrix -> wannierldos (and the
threshold-list, SPF / USPF and DOS outputs that use it), called through the
C ABI (cqpi_set_surface + cqpi_run / cqpi_dos)."""
import numpy as np
import pytest

import cqpi
import oracle
import refmath as rm
from cqpi import synth

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOL_RED = 1e-4


def _oracle_maps(wl, surface, epserr, shops=None):
    out = []
    for om in wl.energies:
        g, it = oracle.surfacegreen(wl.hops, wl.bands, wl.efermi, wl.K, om, wl.eta, epserr=epserr,
                                    surface=surface, shops=shops, scmodel=wl.scmodel)
        assert it > 1
        gR = oracle.realspace(g)
        T = oracle.tmatrix(wl.V, gR[0, 0])
        out.append(oracle.wannierldos(gR, T, wl.tiles, wl.n, wl.window, wl.oversamp, wl.maxband, wl.spin))
    return np.array(out)


def _run(wl, mode, epserr, precision="fp64", shops=None):
    print("case: %s %s K=%d n=%d b=%d E=%d prec=%s shops=%s" % (
        wl.name, mode, wl.K, wl.n, wl.bands, len(wl.energies), precision, shops is not None), flush=True)
    with cqpi.engine_for(wl, precision) as q:
        q.set_surface(mode, epserr, shops)
        return q.run(wl.energies, wl.eta)


def _assert_close(got, ref, tol):
    for e in range(ref.shape[0]):
        mx, rms = rm.rel_errors(got[e], ref[e])
        assert mx < tol, "energy %d: max_rel %.3e rms_rel %.3e" % (e, mx, rms)


@pytest.mark.parametrize("mode", ["surface", "bulk"])
@pytest.mark.parametrize("b", [1, 3])
def test_surface_chain_matches_oracle(gpu, mode, b):
    wl = synth.layered(b, K=64, n=16, E=3, seed=40 + b, w=1, os_=2, eta=0.02)
    _assert_close(_run(wl, mode, 1e-10), _oracle_maps(wl, mode == "surface", 1e-10), TOL64)


@pytest.mark.parametrize("mode", ["surface", "bulk"])
@pytest.mark.parametrize("b", [4, 5, 7, 8, 10])
def test_surface_chain_many_bands(gpu, mode, b):
    """4 bands on the per-thread kernel; 5 to 10 on the lane-group kernel
    (a row of each block per lane, no scratch), e.g. BdG models of 4 and 5
    orbitals."""
    wl = synth.layered(b, K=32, n=8, E=2, seed=60 + b, w=1, os_=2, eta=0.03)
    _assert_close(_run(wl, mode, 1e-10), _oracle_maps(wl, mode == "surface", 1e-10), TOL64)


@pytest.mark.parametrize("b", [5, 10])
def test_surface_many_bands_fp32_default_epserr(gpu, b):
    wl = synth.layered(b, K=64, n=8, E=2, seed=70 + b, w=1, os_=2, eta=0.03)
    _assert_close(_run(wl, "surface", 1e-5, "fp32"), _oracle_maps(wl, True, 1e-5), TOL_RED)


def test_surface_layer_model(gpu):
    """stbfile=: the surface layer's own in-plane terms (here the on-site
    energies shifted) for eps_s; the bulk and coupling blocks from the model."""
    wl = synth.layered(3, K=64, n=16, E=2, seed=47, w=1, os_=2, eta=0.02)
    shops = np.array(wl.hops[wl.hops["rz"] == 0], synth.HOP_DTYPE)
    on = (shops["rx"] == 0) & (shops["ry"] == 0) & (shops["o1"] == shops["o2"])
    shops["tre"][on] += 0.07
    _assert_close(_run(wl, "surface", 1e-10, shops=shops),
                  _oracle_maps(wl, True, 1e-10, shops=shops), TOL64)


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_surface_reduced_precision(gpu, precision):
    wl = synth.layered(3, K=128, n=16, E=2, seed=48, w=1, os_=2, eta=0.02)
    _assert_close(_run(wl, "surface", 1e-10, precision), _oracle_maps(wl, True, 1e-10), TOL_RED)


def test_surface_default_epserr(gpu):
    """epserr = 1e-5 (the driver's default): the decimation stops at the same
    step as the oracle's, so the maps agree to rounding."""
    wl = synth.layered(3, K=64, n=16, E=2, seed=49, w=1, os_=2, eta=0.03)
    _assert_close(_run(wl, "bulk", 1e-5), _oracle_maps(wl, False, 1e-5), 1e-8)


def test_normal_mode_restores_the_eigen_path(gpu):
    wl = synth.layered(3, K=64, n=16, E=2, seed=50, w=1, os_=2, eta=0.02)
    with cqpi.engine_for(wl) as q:
        q.set_surface("surface", 1e-10)
        a = q.run(wl.energies, wl.eta)
        q.set_surface("normal")
        b = q.run(wl.energies, wl.eta)
    _assert_close(a, _oracle_maps(wl, True, 1e-10), TOL64)
    _assert_close(b, oracle.cldos_run(wl), TOL64)


def test_surface_rejections(gpu):
    wl = synth.layered(3, K=64, n=16, E=1, seed=51, w=1, os_=2, eta=0.02)
    q = cqpi.AmdQPI(wl.W, wl.K, wl.n, wl.window, wl.oversamp, wl.bands, wl.maxband, wl.spin, wl.tiles)
    try:
        with pytest.raises(cqpi.CqpiError):
            q.set_surface("surface")  # before set_model
        q.set_model(wl.hops, wl.efermi, False)
        with pytest.raises(cqpi.CqpiError):
            q.set_surface("surface", 0.0)
        with pytest.raises(cqpi.CqpiError):
            q.set_surface(7)
    finally:
        q.close()
    wl10 = synth.bdg_square(5, K=16, n=4, E=1, seed=52)
    with cqpi.engine_for(wl10) as q:
        q.set_surface("surface")  # 10 bands (a BdG model of 5 orbitals): accepted
    with cqpi.engine_for(wl) as q:
        q.set_keep_windows(True)
        q.set_surface("bulk", 1e-10)
        q.run(wl.energies, wl.eta)
        with pytest.raises(cqpi.CqpiError):
            q.gradient(np.ones((1, wl.W * wl.W)))  # the adjoint differentiates the eigen path only


def test_surface_general_grid(gpu):
    """A k grid outside the power-of-two Stockham range (K = 48): the LDS-line
    ky pass reads G(k) from the decimation workspace."""
    wl = synth.layered(3, K=48, n=12, E=2, seed=53, w=1, os_=2, eta=0.02)
    _assert_close(_run(wl, "surface", 1e-10), _oracle_maps(wl, True, 1e-10), TOL64)


def test_surface_threshold_list(gpu):
    wl = synth.layered(3, K=64, n=8, E=2, seed=54, w=1, os_=2, eta=0.02)
    flags, counts = cqpi.make_threshold_list(wl.tiles, 1, 2, 0.05)
    q = cqpi.AmdQPI.from_list(wl.W, wl.K, wl.n, wl.window, wl.oversamp, wl.bands, flags, counts)
    try:
        q.set_model(wl.hops, wl.efermi, wl.scmodel)
        q.set_impurity(wl.V)
        q.set_surface("bulk", 1e-10)
        got = q.run(wl.energies, wl.eta)
    finally:
        q.close()
    for e, om in enumerate(wl.energies):
        g, _ = oracle.surfacegreen(wl.hops, 3, wl.efermi, wl.K, om, wl.eta, epserr=1e-10, surface=False)
        gR = oracle.realspace(g)
        ref = oracle.wannierldoslist(gR, oracle.tmatrix(wl.V, gR[0, 0]), flags, counts, wl.n, 1, 2)
        assert rm.rel_errors(got[e], ref)[0] < TOL64, e


@pytest.mark.parametrize("mode", ["spf", "uspf"])
def test_surface_spectral_maps(gpu, mode):
    wl = synth.layered(3, K=32, n=32, E=2, seed=55, w=1, os_=1, eta=0.02)
    pos = np.array([[0.0, 0.0], [0.5, 0.0], [0.0, 0.5]])
    pre = np.array([1.0, 0.8, 1.2])
    q = cqpi.AmdQPI.spectral(32, 32, 3, 3, False, pos, pre)
    try:
        q.set_model(wl.hops, wl.efermi, False)
        q.set_surface("surface", 1e-10)
        q.set_spectral(mode)
        got = q.run(wl.energies, wl.eta)
    finally:
        q.close()
    for e, om in enumerate(wl.energies):
        g, _ = oracle.surfacegreen(wl.hops, 3, wl.efermi, 32, om, wl.eta, epserr=1e-10, surface=True)
        ref = oracle.spf(g, 32) if mode == "spf" else oracle.uspf(g, 32, pos, pre)
        assert rm.rel_errors(got[e], ref)[0] < TOL64, e


@pytest.mark.parametrize("mode", ["surface", "bulk"])
def test_surface_dos(gpu, mode):
    wl = synth.layered(3, K=40, n=8, E=1, seed=56, w=1, os_=2, eta=0.03)
    om = np.linspace(-0.4, 0.4, 5)
    with cqpi.engine_for(wl) as q:
        q.set_surface(mode, 1e-10)
        got = q.dos(om, 0.03)
    ref = np.array([oracle.dos(oracle.realspace(oracle.surfacegreen(wl.hops, 3, wl.efermi, 40, w, 0.03,
                                                                     epserr=1e-10, surface=mode == "surface")[0]))
                    for w in om])
    assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref))


def test_surface_step_limit_reports_non_convergence(gpu):
    """A step limit below what the k-points need (forced from the host; every
    thread still ends after that many steps): the run and the DOS report
    CQPI_E_CONVERGE, and the default limit runs again afterwards."""
    wl = synth.layered(3, K=64, n=16, E=2, seed=57, w=1, os_=2, eta=0.02)
    with cqpi.engine_for(wl) as q:
        q.set_surface("surface", 1e-10)
        q.set_surface_steps(2)
        with pytest.raises(cqpi.CqpiError) as ei:
            q.run(wl.energies, wl.eta)
        assert ei.value.code == 7  # CQPI_E_CONVERGE
        with pytest.raises(cqpi.CqpiError):
            q.dos([0.0], 0.02)
        with pytest.raises(cqpi.CqpiError):
            q.set_surface_steps(0)
        q.set_surface_steps(500)
        got = q.run(wl.energies, wl.eta)
    _assert_close(got, _oracle_maps(wl, True, 1e-10), TOL64)


def test_surface_josephson_map(gpu):
    """output=josephson with green=surface: the Josephson map of a layered
    superconducting model from the surface Green's function (scmodel: the
    -/+ efermi diagonal of the particle and hole halves)."""
    wl = synth.bdg_square(3, K=64, n=16, E=2, seed=58, w=1, os_=2, eta=0.02)
    rng = np.random.default_rng(59)
    M = 0.1 * (rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6)))
    extra = [(0.0, 0.0, -1.0, a, c, M[a, c].real, M[a, c].imag, 0, 0) for a in range(6) for c in range(6)]
    extra += [(0.0, 0.0, 1.0, a, c, M[c, a].real, -M[c, a].imag, 0, 0) for a in range(6) for c in range(6)]
    wl.hops = np.concatenate([np.asarray(wl.hops, synth.HOP_DTYPE), np.array(extra, synth.HOP_DTYPE)])
    deltat, etat = 0.01, 0.015
    print("case: josephson + surface, BdG 6 bands, K=64, n=16, E=2", flush=True)
    with cqpi.engine_for(wl) as q:
        q.set_josephson(True, deltat, etat)
        q.set_surface("surface", 1e-10)
        got = q.run(wl.energies, wl.eta)
    for e, om in enumerate(wl.energies):
        g, it = oracle.surfacegreen(wl.hops, 6, wl.efermi, wl.K, om, wl.eta, epserr=1e-10, surface=True,
                                    scmodel=True)
        assert it > 1
        gR = oracle.realspace(g)
        oc = complex(om, etat)
        tip = complex(0.0, -np.copysign(1.0, om) * np.pi * deltat) / np.sqrt(oc * oc - deltat * deltat)
        ref = oracle.wannierjosephson(gR, oracle.tmatrix(wl.V, gR[0, 0]), wl.tiles, wl.n, wl.window, wl.oversamp,
                                      tip)
        assert np.abs(ref).max() > 0
        assert rm.rel_errors(got[e], ref)[0] < TOL64, e
