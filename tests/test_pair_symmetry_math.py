"""The identity behind the transform passes' pair symmetry, on the oracle:
for real hoppings of a Hermitian model H(k)^T = H(-k), so that
G0(R)_cr = G0(-R)_rc at every energy; for complex (spin-orbit) hoppings it
fails, so the identity is not vacuous."""
import numpy as np

import oracle
from cqpi import synth


def _window(wl, omega):
    ev, V = oracle.calchamiltonian(wl.hops, wl.bands, wl.efermi, wl.K)
    return oracle.realspace(oracle.greens_k(ev, V, omega, wl.eta))  # [x, y, r, c]


def _mirror_transposed(g):
    # m[x, y, r, c] = g[-x, -y, c, r] (indices mod K)
    m = np.roll(np.flip(g, axis=(0, 1)), 1, axis=(0, 1))
    return np.swapaxes(m, 2, 3)


def test_real_model_satisfies_pair_identity():
    wl = synth.square_3orb(K=16, n=4, E=1)
    for omega in (-0.05, 0.02):
        g = _window(wl, omega)
        dev = np.abs(g - _mirror_transposed(g)).max() / np.abs(g).max()
        assert dev < 1e-13, dev


def test_complex_model_breaks_pair_identity():
    wl = synth.hex_5d_spin(K=12, n=4, E=1)
    g = _window(wl, 0.03)
    dev = np.abs(g - _mirror_transposed(g)).max() / np.abs(g).max()
    assert dev > 1e-3, dev
