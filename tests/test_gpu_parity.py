"""GPU parity: the HIP path against the FP64 oracle (restated reference
algorithm), called through the C ABI.  FP64 mode must agree to 1e-10 relative
(max over the map of |rho - rho_oracle| / max(|rho_oracle|, 1e-3 max|rho_oracle|))."""
import numpy as np
import pytest

import cqpi
import oracle
import refmath as rm
from cqpi import synth

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def _run(wl, precision="fp64"):
    print("case: %s K=%d n=%d w=%d os=%d E=%d spin=%s prec=%s" % (
        wl.name, wl.K, wl.n, wl.window, wl.oversamp, len(wl.energies), wl.spin, precision), flush=True)
    with cqpi.engine_for(wl, precision) as q:
        return q.run(wl.energies, wl.eta)


def _assert_close(got, ref, tol=TOL64):
    for e in range(ref.shape[0]):
        mx, rms = rm.rel_errors(got[e], ref[e])
        assert mx < tol, "energy %d: max_rel %.3e rms_rel %.3e" % (e, mx, rms)


def test_full_chain_3orb_small(gpu):
    wl = synth.square_norb(3, K=32, n=8, E=3, seed=11, w=1, os_=2, elo=-0.1, ehi=0.1, eta=0.01)
    _assert_close(_run(wl), oracle.cldos_run(wl))


def test_full_chain_config1_reduced(gpu):
    wl = synth.square_1orb(K=64, n=16, E=3)
    _assert_close(_run(wl), oracle.cldos_run(wl))


def test_full_chain_config2_shapes_reduced(gpu):
    # config-2 model, tiles, window and oversampling with a smaller grid and map
    wl = synth.square_3orb(K=64, n=16, E=3)
    _assert_close(_run(wl), oracle.cldos_run(wl))


def test_full_chain_spin_hex_reduced(gpu):
    # config-3 model family (hexagonal, 10 bands, SOC, full Hermitian V, spin blocks)
    wl = synth.hex_5d_spin(K=16, n=6, E=2)
    wl.window, wl.oversamp = 1, 2
    wl.tiles = np.concatenate([synth.gaussian_tiles(["xz", "yz", "xy", "x2y2", "z2"], 0.4, 0.5, 1, 2, "hex")] * 2)
    _assert_close(_run(wl), oracle.cldos_run(wl))


@pytest.mark.parametrize("b", [6, 7])
def test_full_chain_lane_group_eigensolve(gpu, b):
    # 6..12 bands take the lane-group Jacobi (odd counts with a decoupled padding row)
    wl = synth.square_norb(b, K=16, n=6, E=2, seed=20 + b, w=1, os_=2, eta=0.02)
    _assert_close(_run(wl), oracle.cldos_run(wl))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_full_chain_wide_window(gpu, prec):
    # w = 31: the FP64 window rows of an 8 x 8-cell block ((8+2w)^2 slots, two
    # blocks) exceed the LDS, so the FP64 contraction gathers them per lane;
    # the FP32 rows still fit and are staged
    wl = synth.square_norb(1, K=128, n=4, E=1, seed=41, w=31, os_=1, eta=0.02)
    _assert_close(_run(wl, prec), oracle.cldos_run(wl), TOL64 if prec == "fp64" else TOLMIX)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_full_chain_spin_hex_gk_path(gpu, prec):
    # 10 bands on a 64^2 grid: G(k) is formed for all 100 pairs in a workspace
    # and read back by the line transform (bands beyond the register path)
    wl = synth.hex_5d_spin(K=64, n=6, E=2)
    wl.window, wl.oversamp = 1, 2
    wl.tiles = np.concatenate([synth.gaussian_tiles(["xz", "yz", "xy", "x2y2", "z2"], 0.4, 0.5, 1, 2, "hex")] * 2)
    _assert_close(_run(wl, prec), oracle.cldos_run(wl), TOL64 if prec == "fp64" else TOLMIX)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_gk_path_energy_blocks(gpu, prec):
    """10 bands on a 1024^2 grid: the G(k) workspace holds 5 (FP32) or 2 (FP64)
    energies, so a run of 6 energies goes through several blocks; each
    energy's map must equal the map of a run of that energy alone."""
    wl = synth.hex_5d_spin(K=1024, n=4, E=6)
    print("case: G(k) workspace blocks K=1024 b=10 E=6 prec=%s" % prec, flush=True)
    with cqpi.engine_for(wl, prec) as q:
        allE = q.run(wl.energies, wl.eta)
        for e in range(len(wl.energies)):
            one = q.run(wl.energies[e:e + 1], wl.eta)
            assert np.array_equal(one[0], allE[e]), "energy %d differs between block and single runs" % e


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_energy_chunk_split_over_two_streams(gpu, prec):
    # 9 energies: the run splits them 5 + 4 over the context's two streams
    wl = synth.square_norb(3, K=64, n=8, E=9, seed=16, w=2, os_=4, eta=0.01)
    got = _run(wl, prec)
    _assert_close(got, oracle.cldos_run(wl), TOL64 if prec == "fp64" else TOLMIX)


def test_non_power_of_two_kgrid(gpu):
    wl = synth.square_norb(3, K=24, n=8, E=2, seed=12, w=1, os_=2, eta=0.02)
    _assert_close(_run(wl), oracle.cldos_run(wl))


def test_map_larger_than_kgrid_odd_n(gpu):
    # n > K: the map repeats periodic images; odd n exercises the window bounds
    wl = synth.square_norb(1, K=8, n=11, E=2, seed=13, w=2, os_=2, eta=0.05)
    _assert_close(_run(wl), oracle.cldos_run(wl))


def test_degenerate_bands(gpu):
    # two identical decoupled orbitals: every band is exactly two-fold degenerate
    base = synth.square_1orb(K=32, n=8, E=2)
    b1 = {(int(h["rx"]), int(h["ry"])): complex(h["tre"], h["tim"]) for h in base.hops}
    blocks = {R: np.diag([t, t]) for R, t in b1.items()}
    wl = synth.Workload(name="doubled", bands=2, hops=synth.hops_from_blocks(blocks, 2), K=32, n=8,
                        window=1, oversamp=2, energies=np.array([-0.2, 0.3]), eta=0.02,
                        V=np.array([[0.7, 0.2], [0.2, -0.4]], complex),
                        tiles=np.concatenate([synth.gaussian_tiles(["s"], 0.35, 0.5, 1, 2)] * 2))
    _assert_close(_run(wl), oracle.cldos_run(wl))


def test_g0_window_and_tmatrix(gpu):
    wl = synth.square_norb(3, K=32, n=8, E=2, seed=14, w=1, os_=2, eta=0.01)
    with cqpi.engine_for(wl) as q:
        q.run(wl.energies, wl.eta)
        h = q.window_half()
        ev, V = oracle.calchamiltonian(wl.hops, wl.bands, wl.efermi, wl.K)
        for e, om in enumerate(wl.energies):
            gR = oracle.realspace(oracle.greens_k(ev, V, om, wl.eta))
            win = q.g0_window(e)
            idx = (np.arange(2 * h + 1) - h) % wl.K
            ref = gR[np.ix_(idx, idx)]
            assert np.max(np.abs(win - ref)) < 1e-12 * np.max(np.abs(ref))
            T = q.tmatrix(e)
            Tref = oracle.tmatrix(wl.V, gR[0, 0])
            assert np.max(np.abs(T - Tref)) < 1e-12 * np.max(np.abs(Tref))


@pytest.mark.parametrize("K,prec", [(1024, "fp64"), (512, "fp64"), (256, "fp64"), (128, "fp64"), (24, "fp64"),
                                    (1024, "fp32"), (512, "fp32"), (128, "fp32"), (64, "fp32")])
def test_transform_matches_rocfft(gpu, K, prec):
    """The in-house k->R transform against rocFFT (torch.fft on a ROCm device
    runs hipFFT/rocFFT): G0(R) = K^-2 sum_k G0(k) e^{+2 pi i k.R} is ifft2 over
    the k-grid.  G0(k) is built on the host from H(k) by a dense inverse.  The
    FP32 transform (reduced-precision modes) is held to 2e-6 of the largest
    entry: FP32 lines of up to 1024 points carry ~1e-7 relative rounding per
    pass."""
    import torch
    tol = 1e-12 if prec == "fp64" else 2e-6
    wl = synth.square_norb(3, K=K, n=8, E=2, seed=15, w=1, os_=2, eta=0.01)
    kg = synth.kgrid(K)
    H = synth.hk_numpy(wl.hops, wl.bands, wl.efermi, kg[:, None], kg[None, :])
    print("case: rocFFT cross-check K=%d b=%d prec=%s" % (K, wl.bands, prec), flush=True)
    with cqpi.engine_for(wl, prec) as q:
        q.run(wl.energies, wl.eta)
        h = q.window_half()
        idx = (np.arange(2 * h + 1) - h) % K
        for e, om in enumerate(wl.energies):
            Gk = np.linalg.inv((om + 1j * wl.eta) * np.eye(wl.bands) - H)
            gR = torch.fft.ifft2(torch.from_numpy(Gk).to("cuda"), dim=(0, 1)).cpu().numpy()
            ref = gR[np.ix_(idx, idx)]
            win = q.g0_window(e)
            assert np.max(np.abs(win - ref)) < tol * np.max(np.abs(ref))


def test_mixed_run_keeps_central_window_only(gpu):
    """A mixed-precision run writes the FP64 window only for |R| <= 2w (the
    T-matrix and the bulk term read no more): retrieving the whole window is
    refused, and with kept windows it is whole and matches the FP64 run's to
    the FP32 transform's accuracy."""
    wl = synth.square_norb(3, K=64, n=16, E=2, seed=17, w=1, os_=2, eta=0.01)
    print("case: mixed-mode window retrieval K=64", flush=True)
    with cqpi.engine_for(wl, "fp64") as q:
        q.run(wl.energies, wl.eta)
        ref = [q.g0_window(e) for e in range(2)]
    with cqpi.engine_for(wl, "mixed") as q:
        q.run(wl.energies, wl.eta)
        with pytest.raises(cqpi.CqpiError):
            q.g0_window(0)
        q.set_keep_windows(True)
        q.run(wl.energies, wl.eta)
        for e in range(2):
            win = q.g0_window(e)
            assert np.max(np.abs(win - ref[e])) < 2e-6 * np.max(np.abs(ref[e]))


def test_dropin_wannierldos_matches_oracle(gpu):
    """GPUQPI::wannierldos path: host G0(R) and T in, one map out."""
    wl = synth.square_3orb(K=32, n=12, E=1)
    ev, V = oracle.calchamiltonian(wl.hops, wl.bands, wl.efermi, wl.K)
    gR = oracle.realspace(oracle.greens_k(ev, V, wl.energies[0], wl.eta))
    T = oracle.tmatrix(wl.V, gR[0, 0])
    ref = oracle.wannierldos(gR, T, wl.tiles, wl.n, wl.window, wl.oversamp)
    q = cqpi.AmdQPI(wl.W, wl.K, wl.n, wl.window, wl.oversamp, wl.bands, wl.maxband, wl.spin, wl.tiles)
    try:
        q.wannierldos(T, gR)
        gR[:] = 0  # the caller may overwrite g0 once wannierldos returned
        got = q.retrieveResult()
        assert "HIP GPU" in q.printinfo()
    finally:
        q.close()
    mx, _ = rm.rel_errors(got, ref)
    assert mx < TOL64


def test_dropin_repeated_calls_and_zero_tile(gpu):
    """Two energies through the drop-in path, with one orbital's tile zero
    (the reference pads missing orbitals with zero functions)."""
    wl = synth.square_3orb(K=32, n=8, E=2)
    tiles = wl.tiles.copy()
    tiles[1] = 0.0
    ev, V = oracle.calchamiltonian(wl.hops, wl.bands, wl.efermi, wl.K)
    q = cqpi.AmdQPI(wl.W, wl.K, wl.n, wl.window, wl.oversamp, wl.bands, wl.maxband, False, tiles)
    try:
        for om in wl.energies:
            gR = oracle.realspace(oracle.greens_k(ev, V, om, wl.eta))
            T = oracle.tmatrix(wl.V, gR[0, 0])
            ref = oracle.wannierldos(gR, T, tiles, wl.n, wl.window, wl.oversamp,
                                     wfzero=np.array([0, 1, 0], np.int32))
            q.wannierldos(T, gR)
            mx, _ = rm.rel_errors(q.retrieveResult(), ref)
            assert mx < TOL64
    finally:
        q.close()


def test_no_impurity_gives_periodic_map(gpu):
    """V = 0: rho is the lattice-periodic bulk term in every cell."""
    wl = synth.square_3orb(K=32, n=8, E=1)
    wl.V = np.zeros((3, 3), complex)
    rho = _run(wl)[0].reshape(wl.W, wl.W)  # [y, x]
    cell = rho[: wl.oversamp, : wl.oversamp]
    tiled = np.tile(cell, (wl.n, wl.n))
    assert np.max(np.abs(rho - tiled)) < 1e-12 * np.max(np.abs(rho))


def test_singular_tmatrix_reported(gpu):
    """1 - V G0(0) singular: reported as an error, not a silent result."""
    wl = synth.square_1orb(K=16, n=4, E=1)
    ev, V = oracle.calchamiltonian(wl.hops, 1, 0.0, wl.K)
    g00 = oracle.realspace(oracle.greens_k(ev, V, 0.2, 0.05))[0, 0, 0, 0]
    wl.V = np.array([[1.0 / g00]])  # makes 1 - V G0(0) exactly zero up to rounding
    wl.energies = np.array([0.2])
    wl.eta = 0.05
    with cqpi.engine_for(wl) as q:
        try:
            out = q.run(wl.energies, wl.eta)
        except cqpi.CqpiError as e:
            assert e.code == 4
        else:
            # rounding may leave a tiny nonzero pivot; then the map is huge but finite
            assert np.all(np.isfinite(out))


# ---------------------------------------------------------------------------
# reduced-precision contractions (split-fp16 MFMA; FP32), FP64 elsewhere: 1e-4 relative
# ---------------------------------------------------------------------------
TOLMIX = 1e-4


def _reduced_case(case):
    if case == "c2":
        wl = synth.square_3orb(K=64, n=16, E=3)
    elif case == "c1":
        wl = synth.square_1orb(K=64, n=16, E=3)
    elif case == "hexspin":
        wl = synth.hex_5d_spin(K=16, n=6, E=2)
        wl.window, wl.oversamp = 1, 2
        wl.tiles = np.concatenate([synth.gaussian_tiles(["xz", "yz", "xy", "x2y2", "z2"], 0.4, 0.5, 1, 2, "hex")] * 2)
    elif case == "oddn":
        wl = synth.square_norb(3, K=16, n=13, E=2, seed=31, w=2, os_=4, eta=0.01)
    else:
        wl = synth.square_norb(5, K=32, n=10, E=2, seed=32, w=2, os_=4, eta=0.01)
    return wl


@pytest.mark.parametrize("prec", ["mixed", "fp32"])
@pytest.mark.parametrize("case", ["c2", "c1", "hexspin", "oddn", "b5"])
def test_reduced_precision_matches_oracle(gpu, case, prec):
    wl = _reduced_case(case)
    got = _run(wl, prec)
    ref = oracle.cldos_run(wl)
    for e in range(ref.shape[0]):
        mx, rms = rm.rel_errors(got[e], ref[e])
        print("%s %s energy %d: max_rel %.3e rms_rel %.3e" % (prec, case, e, mx, rms))
        assert mx < TOLMIX, (case, e, mx, rms)


def test_mixed_run_after_dropin_call_on_one_context(gpu):
    """The drop-in call rewrites the per-energy plane scales; a later whole-chain
    run on the same context must not reuse stale ones (and repeats exactly)."""
    wl = synth.square_3orb(K=32, n=8, E=2)
    print("case: mixed run / drop-in / run on one context", flush=True)
    ev, V = oracle.calchamiltonian(wl.hops, wl.bands, wl.efermi, wl.K)
    gR = oracle.realspace(oracle.greens_k(ev, V, wl.energies[0], wl.eta))
    T = oracle.tmatrix(wl.V, gR[0, 0])
    with cqpi.engine_for(wl, "mixed") as q:
        first = q.run(wl.energies, wl.eta)
        q.wannierldos(T, gR)
        q.retrieveResult()
        again = q.run(wl.energies, wl.eta)
        assert np.array_equal(first, again)
        ref = oracle.cldos_run(wl)
        for e in range(ref.shape[0]):
            assert rm.rel_errors(again[e], ref[e])[0] < TOLMIX


@pytest.mark.parametrize("prec", ["mixed", "fp32"])
def test_reduced_precision_dropin_path(gpu, prec):
    wl = synth.square_3orb(K=32, n=12, E=1)
    ev, V = oracle.calchamiltonian(wl.hops, wl.bands, wl.efermi, wl.K)
    gR = oracle.realspace(oracle.greens_k(ev, V, wl.energies[0], wl.eta))
    T = oracle.tmatrix(wl.V, gR[0, 0])
    ref = oracle.wannierldos(gR, T, wl.tiles, wl.n, wl.window, wl.oversamp)
    q = cqpi.AmdQPI(wl.W, wl.K, wl.n, wl.window, wl.oversamp, wl.bands, wl.maxband, wl.spin, wl.tiles)
    try:
        q.set_precision(prec)
        q.wannierldos(T, gR)
        got = q.retrieveResult()
    finally:
        q.close()
    assert rm.rel_errors(got, ref)[0] < TOLMIX


def test_mixed_dropin_wide_window_packs_in_short_chunks(gpu):
    """A window of more than 256 rows (n = 264: 267 rows) stages 7 of the 9
    pairs per pack workgroup, to stay within 64 KB of LDS, and 2 in the last
    chunk; os = 2 is not a T-folded shape, so the drop-in mixed map packs fp16
    planes.  It matches the FP64 drop-in map to the mixed tolerance."""
    wl = synth.square_norb(3, K=512, n=264, E=1, seed=93, w=1, os_=2, eta=0.02)
    ev, V = oracle.calchamiltonian(wl.hops, wl.bands, wl.efermi, wl.K)
    gR = oracle.realspace(oracle.greens_k(ev, V, wl.energies[0], wl.eta))
    T = oracle.tmatrix(wl.V, gR[0, 0])
    maps = {}
    for prec in ("fp64", "mixed"):
        q = cqpi.AmdQPI(wl.W, wl.K, wl.n, wl.window, wl.oversamp, wl.bands, wl.maxband, wl.spin, wl.tiles)
        try:
            q.set_precision(prec)
            assert 2 * q.window_half() + 1 == 267
            q.wannierldos(T, gR)
            maps[prec] = q.retrieveResult()
        finally:
            q.close()
    assert np.abs(maps["fp64"]).max() > 0
    assert rm.rel_errors(maps["mixed"], maps["fp64"])[0] < TOLMIX
