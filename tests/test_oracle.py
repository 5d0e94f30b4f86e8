"""Pins of the FP64 oracle (CPU).  The reference holds no golden vectors, so
each oracle function is checked against mathematics that fixes it:
LAPACK inverses, numpy's inverse FFT, closed forms of the one-band model, a
brute-force dense inverse of the periodic supercell with the impurity, and
the factorised contraction.  Each check is chosen so that a dropped term, a
wrong sign or a transposed index fails it."""
import numpy as np
import pytest

import oracle
import refmath as rm
from cqpi import synth


@pytest.fixture(scope="module")
def small3():
    wl = synth.square_norb(3, K=8, n=8, E=1, seed=7, w=1, os_=2, eta=0.01)
    wl.efermi = 0.1
    wl.V = np.array([[0.5, 0.1 - 0.2j, 0], [0.1 + 0.2j, 0.3, 0.05], [0, 0.05, 0.2]])
    return wl


def test_setmatrix_closed_form_one_band():
    wl = synth.square_1orb()
    for kx, ky in [(0.0, 0.0), (0.25, -0.125), (0.5, 0.5), (-0.3, 0.1)]:
        H = oracle.setmatrix(wl.hops, 1, 0.2, kx, ky)
        eps = -0.5 + 2 * (-1.0) * (np.cos(2 * np.pi * kx) + np.cos(2 * np.pi * ky)) \
            + 4 * 0.3 * np.cos(2 * np.pi * kx) * np.cos(2 * np.pi * ky) - 0.2
        assert abs(H[0, 0] - eps) < 1e-13


def test_setmatrix_phase_sign_and_anti():
    # a single complex hopping at R=(1,0): +2 pi i phase, conjugated when 'anti'
    hops = np.zeros(1, synth.HOP_DTYPE)
    hops[0] = (1.0, 0.0, 0.0, 0, 1, 0.3, 0.1, 0, 0)
    H = oracle.setmatrix(hops, 2, 0.0, 0.1, 0.0)
    assert abs(H[0, 1] - (0.3 + 0.1j) * np.exp(2j * np.pi * 0.1)) < 1e-15
    hops[0]["anti"] = 1
    H = oracle.setmatrix(hops, 2, 0.0, 0.1, 0.0)
    assert abs(H[0, 1] - (0.3 + 0.1j) * np.exp(-2j * np.pi * 0.1)) < 1e-15


def test_eigh_reconstructs_and_matches_lapack():
    rng = np.random.default_rng(3)
    for b in (1, 2, 3, 5, 10):
        A = rng.normal(size=(b, b)) + 1j * rng.normal(size=(b, b))
        H = A + A.conj().T
        ev, V = oracle.eigh(H)
        assert np.allclose(V @ np.diag(ev) @ V.conj().T, H, atol=1e-13 * np.abs(H).max())
        assert np.allclose(V.conj().T @ V, np.eye(b), atol=1e-14)
        assert np.allclose(np.sort(ev), np.linalg.eigvalsh(H), atol=1e-13 * np.abs(H).max())


def test_eigh_degenerate():
    H = np.diag([1.0, 1.0, -2.0, 1.0]).astype(complex)
    U = np.linalg.qr(np.random.default_rng(1).normal(size=(4, 4)) + 0j)[0]
    H = U @ H @ U.conj().T
    ev, V = oracle.eigh(H)
    assert np.allclose(np.sort(ev), [-2, 1, 1, 1], atol=1e-13)
    assert np.allclose(V @ np.diag(ev) @ V.conj().T, H, atol=1e-13)


def test_greens_k_matches_inverse(small3):
    wl = small3
    ev, V = oracle.calchamiltonian(wl.hops, 3, wl.efermi, wl.K)
    g = oracle.greens_k(ev, V, 0.05, wl.eta)
    ref = rm.g0k_inv(wl.hops, 3, wl.efermi, wl.K, 0.05, wl.eta)
    assert np.max(np.abs(g - ref)) < 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("K", [8, 16, 6, 12])
def test_realspace_matches_numpy_ifft2(K):
    rng = np.random.default_rng(K)
    g = rng.normal(size=(K, K, 2, 2)) + 1j * rng.normal(size=(K, K, 2, 2))
    got = oracle.realspace(g)
    ref = np.fft.ifft2(g, axes=(0, 1))
    assert np.max(np.abs(got - ref)) < 1e-14 * np.max(np.abs(ref)) * K


def test_realspace_sum_rule_and_one_band_closed_form():
    # sum_R G0(R) = G0(k=0); one band: G0(R) = 1/K^2 sum_k e^{2 pi i kR}/(z - eps_k)
    wl = synth.square_1orb(K=16)
    ev, V = oracle.calchamiltonian(wl.hops, 1, 0.0, 16)
    gk = oracle.greens_k(ev, V, 0.1, 0.05)
    gR = oracle.realspace(gk)
    assert abs(gR.sum(axis=(0, 1))[0, 0] - gk[0, 0, 0, 0]) < 1e-13
    k = rm.kgrid(16)
    KX, KY = np.meshgrid(k, k, indexing="ij")
    eps = -0.5 - 2 * (np.cos(2 * np.pi * KX) + np.cos(2 * np.pi * KY)) \
        + 1.2 * np.cos(2 * np.pi * KX) * np.cos(2 * np.pi * KY)
    for R in [(0, 0), (1, 0), (2, -3)]:
        val = np.sum(np.exp(2j * np.pi * (KX * R[0] + KY * R[1])) / (0.1 + 0.05j - eps)) / 256
        assert abs(gR[R[0] % 16, R[1] % 16, 0, 0] - val) < 1e-14


def test_tmatrix_dyson_identity():
    rng = np.random.default_rng(5)
    b = 4
    V = rng.normal(size=(b, b)) + 1j * rng.normal(size=(b, b))
    V = V + V.conj().T
    G = rng.normal(size=(b, b)) + 1j * rng.normal(size=(b, b))
    T = oracle.tmatrix(V, G)
    assert np.allclose(T, V + V @ G @ T, atol=1e-12)
    # one band closed form
    T1 = oracle.tmatrix(np.array([[1.0]]), np.array([[0.2 - 0.3j]]))
    assert abs(T1[0, 0] - 1.0 / (1 - (0.2 - 0.3j))) < 1e-15


def test_contraction_matches_factorised(small3):
    wl = small3
    ev, V = oracle.calchamiltonian(wl.hops, 3, wl.efermi, wl.K)
    gR = oracle.realspace(oracle.greens_k(ev, V, 0.05, wl.eta))
    T = oracle.tmatrix(wl.V, gR[0, 0])
    got = oracle.wannierldos(gR, T, wl.tiles, wl.n, wl.window, wl.oversamp)
    ref = rm.ldos_factorised(gR, T, wl.tiles, wl.n, wl.window, wl.oversamp)
    assert rm.rel_errors(got, ref)[0] < 1e-12


def test_contraction_spin_blocks_factorised():
    wl = synth.square_norb(3, K=6, n=6, E=1, seed=21, w=1, os_=2)
    # 4 bands, spin: blocks {0,1} and {2,3}
    rng = np.random.default_rng(2)
    gR = rng.normal(size=(6, 6, 4, 4)) + 1j * rng.normal(size=(6, 6, 4, 4))
    T = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    tiles = rng.normal(size=(4, 6, 6))
    got = oracle.wannierldos(gR, T, tiles, 6, 1, 2, maxband=4, spin=True)
    ref = rm.ldos_factorised(gR, T, tiles, 6, 1, 2, maxband=4, spin=True)
    assert rm.rel_errors(got, ref)[0] < 1e-12


def _josephson_inputs(b, seed):
    rng = np.random.default_rng(seed)
    K, n, w, os_ = 6, 6, 1, 2
    gR = rng.normal(size=(K, K, b, b)) + 1j * rng.normal(size=(K, K, b, b))
    T = rng.normal(size=(b, b)) + 1j * rng.normal(size=(b, b))
    tiles = rng.normal(size=(b // 2, (2 * w + 1) * os_, (2 * w + 1) * os_))
    return gR, T, tiles, n, w, os_


def _hole_swap(b):
    """Column permutation o <-> o + b/2 of the particle and hole blocks."""
    mb = b // 2
    return np.array([o + mb if o < mb else o - mb for o in range(b)])


@pytest.mark.parametrize("b,spin", [(4, False), (6, False), (8, True)])
def test_josephson_is_ldos_of_the_hole_block(b, spin):
    """With tip = 1 the Josephson map (column o1 + maxband) equals the LDOS map
    of G0 Pi with T' = Pi^T T, Pi swapping the particle and hole columns:
    (G0 Pi T' G0 Pi)_{o2,o1} = (G0 T G0)_{o2,o1+maxband} and the periodic term
    G0 Pi(R2-R1)_{o2,o1} = G0(R2-R1)_{o2,o1+maxband}.  The LDOS oracle is
    pinned by the supercell inverse; a wrong column, a dropped periodic term or
    a transposed product fails here."""
    gR, T, tiles, n, w, os_ = _josephson_inputs(b, 30 + b)
    pi = _hole_swap(b)
    got = oracle.wannierjosephson(gR, T, tiles, n, w, os_, 1.0, spin=spin)
    ref = oracle.wannierldos(gR[..., pi], T[pi, :], tiles, n, w, os_, maxband=b // 2, spin=spin)
    assert rm.rel_errors(got, ref)[0] < 1e-12


@pytest.mark.parametrize("phi", [0.7, 2.0, -1.3])
def test_josephson_tip_phase(phi):
    """z = (G0 T G0 + G0)_{o2,o1+maxband} turns into e^{i phi} z under
    G0 -> e^{i phi} G0, T -> e^{-i phi} T, so the map with tip = 1 on the
    rotated inputs equals the map with tip = e^{i phi} on the original ones;
    a conjugated tip or Re in place of Im fails this."""
    gR, T, tiles, n, w, os_ = _josephson_inputs(4, 41)
    ph = np.exp(1j * phi)
    got = oracle.wannierjosephson(gR, T, tiles, n, w, os_, ph)
    ref = oracle.wannierjosephson(gR * ph, T / ph, tiles, n, w, os_, 1.0)
    assert rm.rel_errors(got, ref)[0] < 1e-12
    # real scaling of the tip scales the map
    got2 = oracle.wannierjosephson(gR, T, tiles, n, w, os_, 2.5 * ph)
    assert np.allclose(got2, 2.5 * got, rtol=1e-13, atol=1e-13 * np.abs(got).max())


def test_whole_path_matches_supercell_bruteforce(small3):
    """Dense inverse of the K x K periodic supercell with the impurity on site 0
    pins eigensolve, resolvent, transform, T-matrix and contraction together."""
    wl = small3
    om = 0.05
    got = oracle.cldos_run(wl, energies=[om])[0]
    ref = rm.ldos_bruteforce(wl.hops, 3, wl.efermi, wl.K, wl.n, wl.window, wl.oversamp, om, wl.eta,
                             wl.V, wl.tiles)
    mx, rms = rm.rel_errors(got, ref)
    assert mx < 1e-11, (mx, rms)


def test_bruteforce_detects_planted_mistake(small3):
    """A wrong sign of the bulk term in the restatement would be caught."""
    wl = small3
    ev, V = oracle.calchamiltonian(wl.hops, 3, wl.efermi, wl.K)
    gR = oracle.realspace(oracle.greens_k(ev, V, 0.05, wl.eta))
    T = oracle.tmatrix(wl.V, gR[0, 0])
    good = oracle.wannierldos(gR, T, wl.tiles, wl.n, wl.window, wl.oversamp)
    # bulk-only map (T=0) subtracted twice emulates a flipped bulk sign
    bulk = oracle.wannierldos(gR, np.zeros_like(T), wl.tiles, wl.n, wl.window, wl.oversamp)
    planted = good - 2 * bulk
    ref = rm.ldos_bruteforce(wl.hops, 3, wl.efermi, wl.K, wl.n, wl.window, wl.oversamp, 0.05, wl.eta,
                             wl.V, wl.tiles)
    assert rm.rel_errors(planted, ref)[0] > 1e-3


def test_getbias():
    assert oracle.getbias(-0.1, 0.1, 51, 0) == -0.1
    assert abs(oracle.getbias(-0.1, 0.1, 51, 50) - 0.1) < 1e-17
    assert oracle.getbias(0.3, 0.5, 1, 0) == 0.3


def test_bdg_workload_is_particle_hole_symmetric():
    """The superconducting test model: Hermitian H(k), hole block -H(-k)^*
    through the 'anti' phase and the +mu diagonal, pairing in the off-diagonal
    blocks; the oracle's H(k) (scmodel) agrees with the numpy one."""
    wl = synth.bdg_square(3, K=8, n=4, E=1, seed=63)
    k = synth.kgrid(8)
    kx, ky = np.meshgrid(k, k, indexing="ij")
    H = synth.hk_numpy(wl.hops, 6, wl.efermi, kx, ky, scmodel=True)
    Hm = synth.hk_numpy(wl.hops, 6, wl.efermi, -kx, -ky, scmodel=True)
    assert np.abs(H - np.conj(np.swapaxes(H, -1, -2))).max() < 1e-14
    assert np.abs(H[..., 3:, 3:] + np.conj(Hm[..., :3, :3])).max() < 1e-14
    assert np.abs(H[..., :3, 3:] - wl.meta["delta"] * np.eye(3)).max() < 1e-14
    ev, vec = oracle.calchamiltonian(wl.hops, 6, wl.efermi, 8, scmodel=True)
    Hr = np.einsum("...in,...n,...jn->...ij", vec, ev, np.conj(vec))
    assert np.abs(Hr - H).max() < 1e-12


@pytest.mark.parametrize("spin", [False, True])
def test_full_threshold_list_is_the_dense_map(spin):
    """A threshold below zero keeps every (R1, R2, o1, o2, sub-pixel) term, so
    the list contraction must equal the dense contraction (pinned by the
    supercell inverse); this also pins the builder's enumeration and factors."""
    b = 4 if spin else 3
    rng = np.random.default_rng(70 + b)
    K, n, w, os_ = 6, 6, 1, 2
    gR = rng.normal(size=(K, K, b, b)) + 1j * rng.normal(size=(K, K, b, b))
    T = rng.normal(size=(b, b)) + 1j * rng.normal(size=(b, b))
    tiles = rng.normal(size=(b, 6, 6))
    flags, counts = oracle.mkthresholdlist(tiles, w, os_, -1.0, spin=spin)
    assert len(flags) == (2 * w + 1) ** 4 * os_ ** 2 * (b * b if not spin else 2 * (b // 2) ** 2)
    got = oracle.wannierldoslist(gR, T, flags, counts, n, w, os_)
    ref = oracle.wannierldos(gR, T, tiles, n, w, os_, maxband=b, spin=spin)
    assert rm.rel_errors(got, ref)[0] < 1e-12


def test_threshold_list_selects_by_magnitude():
    """The list at threshold t is exactly the full list's entries with
    |factor| > t * max(w^2), pair by pair, in the same order."""
    wl = synth.square_norb(3, K=8, n=6, E=1, seed=7, w=1, os_=2)
    full, cfull = oracle.mkthresholdlist(wl.tiles, 1, 2, -1.0)
    maxval = np.max(wl.tiles ** 2)
    offs = np.concatenate([[0], np.cumsum(cfull)]).astype(int)
    for t in (0.01, 0.05, 0.2):
        sel, csel = oracle.mkthresholdlist(wl.tiles, 1, 2, t)
        keep = np.abs(full["factor"]) > t * maxval
        assert np.array_equal(sel, full[keep])
        assert np.array_equal(csel, [keep[offs[p]:offs[p + 1]].sum() for p in range(len(cfull))])
        assert 0 < len(sel) < len(full)


def test_threshold_list_contraction_is_linear_in_the_list():
    """Splitting every pair's list in two gives two maps that sum to the map of
    the whole list (each entry adds its own term)."""
    wl = synth.square_norb(3, K=8, n=6, E=1, seed=7, w=1, os_=2, eta=0.01)
    ev, V = oracle.calchamiltonian(wl.hops, 3, wl.efermi, wl.K)
    gR = oracle.realspace(oracle.greens_k(ev, V, 0.05, wl.eta))
    T = oracle.tmatrix(wl.V, gR[0, 0])
    flags, counts = oracle.mkthresholdlist(wl.tiles, 1, 2, 0.02)
    offs = np.concatenate([[0], np.cumsum(counts)]).astype(int)
    parts = [([], []), ([], [])]
    for p in range(len(counts)):
        seg = flags[offs[p]:offs[p + 1]]
        h = len(seg) // 2
        for k, piece in enumerate((seg[:h], seg[h:])):
            parts[k][0].append(piece)
            parts[k][1].append(len(piece))
    whole = oracle.wannierldoslist(gR, T, flags, counts, 6, 1, 2)
    split = sum(oracle.wannierldoslist(gR, T, np.concatenate(f), np.array(c, np.uintp), 6, 1, 2)
                for f, c in parts)
    assert rm.rel_errors(split, whole)[0] < 1e-12


def test_spf_one_band_closed_form_and_centering():
    """One band: A(k) = eta / ((w - e(k))^2 + eta^2), shown with k = 0 at
    pixel (n/2, n/2) (the map is the k grid rolled by n/2)."""
    wl = synth.square_1orb(K=16, n=16, E=1)
    om, eta = 0.3, 0.05
    ev, V = oracle.calchamiltonian(wl.hops, 1, wl.efermi, 16)
    g0k = oracle.greens_k(ev, V, om, eta)
    got = oracle.spf(g0k, 16).reshape(16, 16).T        # [x, y]
    k = synth.kgrid(16)
    kx, ky = np.meshgrid(k, k, indexing="ij")
    e = synth.hk_numpy(wl.hops, 1, wl.efermi, kx, ky)[..., 0, 0].real
    A = eta / ((om - e) ** 2 + eta ** 2)
    assert np.allclose(got, np.roll(np.roll(A, 8, 0), 8, 1), rtol=1e-12, atol=0)
    assert abs(got[8, 8] - A[0, 0]) < 1e-12 * A[0, 0]


def test_dos_is_the_mean_spectral_weight():
    """G0(R=0) is the k average of G0(k): the DOS equals the mean of the SPF map
    over the whole grid (n = K)."""
    wl = synth.square_norb(3, K=16, n=16, E=1, seed=9)
    ev, V = oracle.calchamiltonian(wl.hops, 3, wl.efermi, 16)
    g0k = oracle.greens_k(ev, V, 0.1, 0.02)
    for mb in (3, 2):
        assert abs(oracle.dos(oracle.realspace(g0k), mb) - oracle.spf(g0k, 16, mb).mean()) < 1e-12


def test_uspf_one_band_at_the_origin_is_the_spf():
    wl = synth.square_1orb(K=16, n=16, E=1)
    ev, V = oracle.calchamiltonian(wl.hops, 1, wl.efermi, 16)
    g0k = oracle.greens_k(ev, V, -0.2, 0.03)
    u = oracle.uspf(g0k, 16, np.zeros((1, 2)), np.ones(1))
    assert np.allclose(u, oracle.spf(g0k, 16), rtol=1e-13, atol=0)


@pytest.mark.parametrize("spin", [False, True])
def test_uspf_position_phase_is_a_gauge_shift(spin):
    """Orbital positions x_o enter as e^{2 pi i k.(x_k - x_l)}: the map of model A
    with positions x equals the map at zero positions of model B whose hops
    (o1, o2, R) move to R + x_o2 - x_o1 (H_B = U^H H_A U, U = diag e^{2 pi i k.x}).
    Pixels on the zone edge (index 0) pick k = -1/2 in the phase but +1/2 on
    the grid, so they are left out."""
    b = 4 if spin else 2
    rng = np.random.default_rng(5 + b)
    onsite = rng.normal(0, 0.3, (b, b))
    onsite = 0.5 * (onsite + onsite.T)
    blocks = synth.hermitian_blocks(synth.square_shells(2), b, rng,
                                    lambda sh: rng.normal(0, 0.3, sh) + 1j * rng.normal(0, 0.1, sh),
                                    onsite.astype(np.complex128))
    hA = synth.hops_from_blocks(blocks, b)
    pos = rng.uniform(-0.5, 0.5, (b, 2))
    z = rng.uniform(0.5, 1.5, b)
    hB = hA.copy()
    hB["rx"] += pos[hA["o2"], 0] - pos[hA["o1"], 0]
    hB["ry"] += pos[hA["o2"], 1] - pos[hA["o1"], 1]
    K = 16
    gA = oracle.greens_k(*oracle.calchamiltonian(hA, b, 0.05, K), 0.1, 0.02)
    gB = oracle.greens_k(*oracle.calchamiltonian(hB, b, 0.05, K), 0.1, 0.02)
    uA = oracle.uspf(gA, K, pos, z, spin=spin).reshape(K, K)
    uB = oracle.uspf(gB, K, np.zeros((b, 2)), z, spin=spin).reshape(K, K)
    assert np.allclose(uA[1:, 1:], uB[1:, 1:], rtol=1e-11, atol=1e-12 * np.abs(uA).max())
    assert not np.allclose(uA, oracle.uspf(gA, K, np.zeros((b, 2)), z, spin=spin).reshape(K, K))


def _layered_1band(t_perp, rzsign=-1):
    """Square-lattice one-band layers (rz = 0) coupled along z by t_perp at
    R = (0, 0, -1) and its conjugate at (0, 0, +1)."""
    wl = synth.square_1orb(K=8, n=8, E=1)
    h = list(wl.hops)
    extra = np.zeros(2, synth.HOP_DTYPE)
    extra["rz"] = [rzsign, -rzsign]
    extra["tre"] = [np.real(t_perp), np.real(t_perp)]
    extra["tim"] = [np.imag(t_perp), -np.imag(t_perp)]
    return wl, np.concatenate([np.array(h, synth.HOP_DTYPE), extra])


def test_surface_green_without_coupling_is_the_plain_resolvent():
    wl = synth.square_norb(3, K=8, n=8, E=1, seed=12)
    for surface in (True, False):
        g, it = oracle.surfacegreen(wl.hops, 3, 0.05, 8, 0.1, 0.03, surface=surface)
        ref = oracle.greens_k(*oracle.calchamiltonian(wl.hops, 3, 0.05, 8), 0.1, 0.03)
        assert it == 1 and np.allclose(g, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("surface", [True, False])
def test_surface_green_semi_infinite_chain_closed_form(surface):
    """One band per layer, interlayer hopping t: per k the layers form a
    semi-infinite chain with on-site e(k), whose surface and bulk Green's
    functions are (z-e -/+ sqrt((z-e)^2 - 4|t|^2)) / (2|t|^2) and
    1/sqrt((z-e)^2 - 4|t|^2) on the retarded branch (Im g < 0)."""
    t = 0.4 * np.exp(0.3j)
    wl, hops = _layered_1band(t)
    om, eta = 0.2, 0.05
    g, it = oracle.surfacegreen(hops, 1, wl.efermi, 8, om, eta, epserr=1e-14, surface=surface)
    assert it > 1
    k = synth.kgrid(8)
    kx, ky = np.meshgrid(k, k, indexing="ij")
    e = synth.hk_numpy(wl.hops, 1, wl.efermi, kx, ky)[..., 0, 0]
    z = om + 1j * eta - e
    root = np.sqrt(z * z - 4 * abs(t) ** 2)
    if surface:
        cands = [(z - root) / (2 * abs(t) ** 2), (z + root) / (2 * abs(t) ** 2)]
    else:
        cands = [1 / root, -1 / root]
    ref = np.where(np.imag(cands[0]) < 0, cands[0], cands[1])
    assert np.allclose(g[..., 0, 0], ref, rtol=1e-10, atol=0)


@pytest.mark.parametrize("t_perp", [0.0, 0.4 * np.exp(0.3j)])
def test_surface_green_with_surface_model_closed_form(t_perp):
    """stbfile (src/qpi.cc:488-554): the surface layer's in-plane Hamiltonian
    comes from the surface model, here the bulk band shifted by delta, while
    the layers below keep the bulk band e(k).  The surface site then sees the
    semi-infinite bulk chain below it through the interlayer hopping t:
      G_s = 1 / (z - e(k) - delta - |t|^2 g0),
      g0 = (z - e - sqrt((z - e)^2 - 4|t|^2)) / (2|t|^2)  (Im g0 < 0),
    and for uncoupled layers (t = 0) G_s = 1/(z - e(k) - delta).  The bulk
    Green's function does not see the surface model."""
    delta = 0.07
    wl, hops = _layered_1band(t_perp) if t_perp else (synth.square_1orb(K=8, n=8, E=1), None)
    if hops is None:
        hops = np.array(list(wl.hops), synth.HOP_DTYPE)
    shift = np.zeros(1, synth.HOP_DTYPE)
    shift["tre"] = delta
    shops = np.concatenate([np.array(list(wl.hops), synth.HOP_DTYPE), shift])
    om, eta = 0.2, 0.05
    g, it = oracle.surfacegreen(hops, 1, wl.efermi, 8, om, eta, epserr=1e-14, surface=True, shops=shops)
    k = synth.kgrid(8)
    kx, ky = np.meshgrid(k, k, indexing="ij")
    e = synth.hk_numpy(wl.hops, 1, wl.efermi, kx, ky)[..., 0, 0]
    z = om + 1j * eta - e
    t2 = abs(t_perp) ** 2
    if t2:
        root = np.sqrt(z * z - 4 * t2)
        c0, c1 = (z - root) / (2 * t2), (z + root) / (2 * t2)
        g0 = np.where(np.imag(c0) < 0, c0, c1)
        ref = 1 / (z - delta - t2 * g0)
    else:
        ref = 1 / (z - delta)
    assert np.allclose(g[..., 0, 0], ref, rtol=1e-10, atol=0)
    gb, _ = oracle.surfacegreen(hops, 1, wl.efermi, 8, om, eta, epserr=1e-14, surface=False, shops=shops)
    gb0, _ = oracle.surfacegreen(hops, 1, wl.efermi, 8, om, eta, epserr=1e-14, surface=False)
    assert np.array_equal(gb, gb0)


def test_bands_histogram_flat_band_closed_form():
    """A flat band e0 (on-site term only): every k adds the same Lorentzian
    w/2/((bias - e0)^2 + w^2/4)/pi, w = 2 (ue - le)/layers, to the three layers
    around round((e0 - le)/(ue - le)(L - 1)), os^2 times per pixel
    (src/tightbinding.cc:1158-1193); nothing elsewhere.  e0 sits exactly half
    way between two layers, where C's round() goes up."""
    le, ue, L, pix, os_ = -0.1, 0.1, 11, 3, 2
    e0 = le + 4.5 * (ue - le) / (L - 1)
    hops = np.zeros(1, synth.HOP_DTYPE)
    hops["tre"] = e0
    h = oracle.bands_histogram(hops, 1, 0.0, pix, os_, le, ue, L)
    w = 2 * (ue - le) / L
    ref = np.zeros(L)
    for nl in (4, 5, 6):
        x = le + (ue - le) * nl / (L - 1) - e0
        ref[nl] = os_ * os_ * w / 2 / (x * x + w * w / 4) / np.pi
    assert np.allclose(h, ref[None, None, :], rtol=1e-12, atol=0)


def test_bands_histogram_dispersive_band_on_the_reference_k_grid():
    """The histogram's k-grid (src/tightbinding.cc:1158-1193): sub-pixel io of
    ospix = pix*os samples k = io/(ospix-1) - 1/2, the ends included, and
    pixel (x, y) = (io/os, jo/os) takes the band energies of (kx(io), ky(jo)).
    With the anisotropic band e(k) = 2 tx cos 2pi kx + 2 ty cos 2pi ky and
    ospix = 3 the samples are kx, ky in {-1/2, 0, 1/2}, cos in {-1, 1, -1}, so
    pixel (x, y) holds one energy 2 tx c(x) + 2 ty c(y) in {-0.3, -0.1, 0.1,
    0.3}, each the centre of a layer: that layer gets 2/(pi w) and its two
    neighbours w/2/(bin^2 + w^2/4)/pi.  A grid io/ospix - 1/2, swapped axes or
    a dropped sub-pixel would move energies off the layer centres or between
    pixels."""
    le, ue, L = -1.0, 1.0, 21
    tx, ty = 0.1, 0.05
    hops = np.zeros(4, synth.HOP_DTYPE)
    for i, (rx, ry, t) in enumerate([(1, 0, tx), (-1, 0, tx), (0, 1, ty), (0, -1, ty)]):
        hops[i]["rx"], hops[i]["ry"], hops[i]["tre"] = rx, ry, t
    h = oracle.bands_histogram(hops, 1, 0.0, 3, 1, le, ue, L)
    binsize = (ue - le) / (L - 1)
    w = 2 * (ue - le) / L
    c = [-1.0, 1.0, -1.0]
    for x in range(3):
        for y in range(3):
            e = 2 * tx * c[x] + 2 * ty * c[y]
            lay = int(round((e - le) / binsize))
            ref = np.zeros(L)
            ref[lay] = 2 / (np.pi * w)
            ref[lay - 1] = ref[lay + 1] = w / 2 / (binsize ** 2 + w * w / 4) / np.pi
            assert np.allclose(h[x, y], ref, rtol=1e-9, atol=1e-12), (x, y, e)
    # the same grid summed over 3 x 3 sub-pixels of one pixel
    h1 = oracle.bands_histogram(hops, 1, 0.0, 1, 3, le, ue, L)
    assert np.allclose(h1[0, 0], h.sum(axis=(0, 1)), rtol=1e-12, atol=0)


def test_cell_block_contraction_equals_full_map_on_the_block():
    """or_wannierldos_cells computes exactly the full loop nest's pixels of
    its cells and leaves the rest 0 (the sample the bench compares the
    10-band workload on)."""
    wl = synth.square_norb(3, K=16, n=8, E=1, seed=61, w=1, os_=2, eta=0.02)
    ev, V = oracle.calchamiltonian(wl.hops, wl.bands, wl.efermi, wl.K)
    gR = oracle.realspace(oracle.greens_k(ev, V, wl.energies[0], wl.eta))
    T = oracle.tmatrix(wl.V, gR[0, 0])
    full = oracle.wannierldos(gR, T, wl.tiles, wl.n, wl.window, wl.oversamp).reshape(wl.W, wl.W)  # [y, x]
    part = oracle.wannierldos_cells(gR, T, wl.tiles, wl.n, wl.window, wl.oversamp, (2, 5, 3, 7)).reshape(wl.W, wl.W)
    os_ = wl.oversamp
    xs, ys = slice(2 * os_, 5 * os_), slice(3 * os_, 7 * os_)
    assert np.array_equal(part[ys, xs], full[ys, xs])
    mask = np.ones_like(part, bool)
    mask[ys, xs] = False
    assert not np.any(part[mask])
