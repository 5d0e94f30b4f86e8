"""Independent numpy mathematics used to pin the oracle (tests only).

None of this restates the reference's loop structure; it computes the same
quantities by different routes:
  * G0(k) = inv(w + i eta - H(k)) with LAPACK (no eigenvectors),
  * G0(R) = numpy.fft.ifft2(G0(k)) (e^{+2 pi i k.R}, 1/K^2),
  * T from numpy.linalg.solve,
  * the LDOS from a dense inverse of the K x K periodic supercell with the
    impurity on site 0 (brute force; pins eigen/resolvent/transform/T/contraction
    together), and
  * the LDOS from the factorised contraction rho = -Im(a^T T b) - bulk.
"""
from __future__ import annotations

import numpy as np


def hk(hops, b, efermi, kx, ky, scmodel=False):
    import importlib
    synth = importlib.import_module("cqpi.synth")
    return synth.hk_numpy(hops, b, efermi, kx, ky, scmodel)


def kgrid(K):
    i = np.arange(K, dtype=np.float64)
    return np.where(i > (K >> 1), (i - K) / K, i / K)


def g0k_inv(hops, b, efermi, K, omega, eta, scmodel=False):
    k = kgrid(K)
    KX, KY = np.meshgrid(k, k, indexing="ij")
    H = hk(hops, b, efermi, KX, KY, scmodel)
    z = (omega + 1j * eta) * np.eye(b)
    return np.linalg.inv(z - H)


def g0R_fft(g0k):
    return np.fft.ifft2(g0k, axes=(0, 1))


def tmatrix(V, g00):
    b = V.shape[0]
    return np.linalg.solve(np.eye(b) - V @ g00, V)


def supercell_H(hops, b, efermi, K, scmodel=False):
    """H_sc[(r,a),(r',c)] = t_ac(r'-r) (mod K) - efermi delta: the real-space
    operator whose Bloch transform is H(k) = sum_R t(R) e^{+2 pi i k.R}."""
    N = K * K
    H = np.zeros((N, b, N, b), np.complex128)
    scmax = b >> 1 if scmodel else b
    for i in range(b):
        for s in range(N):
            H[s, i, s, i] = -efermi if i < scmax else efermi
    xs, ys = np.meshgrid(np.arange(K), np.arange(K), indexing="ij")
    src = (xs * K + ys).ravel()
    for h in hops:
        if h["anti"]:
            raise NotImplementedError
        rx, ry = int(h["rx"]), int(h["ry"])
        dst = (((xs + rx) % K) * K + (ys + ry) % K).ravel()
        t = complex(h["tre"], h["tim"])
        np.add.at(H, (src, h["o1"], dst, h["o2"]), t)
    return H.reshape(N * b, N * b)


def ldos_bruteforce(hops, b, efermi, K, n, window, oversamp, omega, eta, V, tiles,
                    maxband=None, spin=False):
    """rho(pixel) = -Im sum_{R1,R2,o1,o2} w_o1(x-R1) w_o2(x-R2) G(pos+R2, pos+R1)_{o2,o1}
    with G = inv(z - H_sc - V at site 0) on the K x K periodic supercell."""
    if maxband is None:
        maxband = b
    N = K * K
    H = supercell_H(hops, b, efermi, K)
    H[0:b, 0:b] += V
    G = np.linalg.inv((omega + 1j * eta) * np.eye(N * b) - H).reshape(N, b, N, b)
    W = n * oversamp
    n2 = n >> 1
    w = window
    rho = np.zeros((W, W))
    if spin:
        sb = maxband >> 1
        sets = [list(range(0, sb)), list(range(sb, 2 * sb))]
    else:
        sets = [list(range(maxband))]
    for i in range(n):
        for j in range(n):
            pos = (i - n2, j - n2)
            for ii in range(oversamp):
                for jj in range(oversamp):
                    x, y = ii + w * oversamp, jj + w * oversamp
                    acc = 0j
                    for r1x in range(-w, w + 1):
                        for r1y in range(-w, w + 1):
                            s1 = ((pos[0] + r1x) % K) * K + (pos[1] + r1y) % K
                            for r2x in range(-w, w + 1):
                                for r2y in range(-w, w + 1):
                                    s2 = ((pos[0] + r2x) % K) * K + (pos[1] + r2y) % K
                                    for S in sets:
                                        for o1 in S:
                                            f1 = tiles[o1, x - r1x * oversamp, y - r1y * oversamp]
                                            for o2 in S:
                                                f2 = tiles[o2, x - r2x * oversamp, y - r2y * oversamp]
                                                acc += f1 * f2 * G[s2, o2, s1, o1]
                    rho[i * oversamp + ii, j * oversamp + jj] = -acc.imag
    return rho.T.ravel()  # x + y*W


def ldos_factorised(g0R, T, tiles, n, window, oversamp, maxband=None, spin=False):
    """rho = -Im sum_sigma a^T T b - bulk, with
    a_p(r) = sum_{R2,o2} w_o2(r;R2) G0(pos+R2)[o2,p],
    b_q(r) = sum_{R1,o1} G0(-pos-R1)[q,o1] w_o1(r;R1),
    bulk(r) = sum w_o1(r;R1) w_o2(r;R2) Im G0(R2-R1)[o2,o1]."""
    K = g0R.shape[0]
    b = g0R.shape[2]
    if maxband is None:
        maxband = b
    w, os_ = window, oversamp
    W = n * os_
    n2 = n >> 1
    if spin:
        sb = maxband >> 1
        sets = [np.arange(0, sb), np.arange(sb, 2 * sb)]
    else:
        sets = [np.arange(maxband)]
    rho = np.zeros((W, W))
    Rs = [(rx, ry) for rx in range(-w, w + 1) for ry in range(-w, w + 1)]
    for ii in range(os_):
        for jj in range(os_):
            x, y = ii + w * os_, jj + w * os_
            wt = np.array([[tiles[o, x - rx * os_, y - ry * os_] for o in range(maxband)] for rx, ry in Rs])
            # bulk
            bulk = 0.0
            for S in sets:
                for a1, (r1x, r1y) in enumerate(Rs):
                    for a2, (r2x, r2y) in enumerate(Rs):
                        g = g0R[(r2x - r1x) % K, (r2y - r1y) % K]
                        bulk += np.einsum("i,j,ji->", wt[a1, S], wt[a2, S], g[np.ix_(S, S)].imag)
            for i in range(n):
                for j in range(n):
                    px, py = i - n2, j - n2
                    val = 0.0
                    for S in sets:
                        a = np.zeros(b, complex)
                        bb = np.zeros(b, complex)
                        for a_, (rx, ry) in enumerate(Rs):
                            ga = g0R[(px + rx) % K, (py + ry) % K]
                            gb = g0R[(-px - rx) % K, (-py - ry) % K]
                            a += wt[a_, S] @ ga[S, :]
                            bb += gb[:, S] @ wt[a_, S]
                        val += (a @ T @ bb).imag
                    rho[i * os_ + ii, j * os_ + jj] = -(val + bulk)
    return rho.T.ravel()


def rel_errors(x, ref, floor=1e-3):
    """max_rel = max|x-ref| / max(|ref|, floor*max|ref|); rms_rel = ||x-ref||/||ref||."""
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    m = np.max(np.abs(ref))
    den = np.maximum(np.abs(ref), floor * m)
    max_rel = float(np.max(np.abs(x - ref) / den)) if m > 0 else float(np.max(np.abs(x - ref)))
    nr = np.linalg.norm(ref)
    rms_rel = float(np.linalg.norm(x - ref) / nr) if nr > 0 else float(np.linalg.norm(x - ref))
    return max_rel, rms_rel
