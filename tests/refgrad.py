"""Plain PyTorch FP64 reference of the continuum-LDOS chain (tests only).

Differentiable end to end with torch autograd, so it provides the FP64
reference gradient for the hand-written adjoint.  It computes the same
quantities by a different route from the library and the oracle: H(k) as a sum
of one-hot phase tensors, G0(k) by torch.linalg.inv, G0(R) by torch.fft.ifft2,
T by torch.linalg.solve and the contraction factorised as -Im(a^T T b) - bulk.
Square lattices without spin blocks.  Maps are returned as (E, W*W), x + y*W.
"""
from __future__ import annotations

import numpy as np
import torch


def kgrid(K):
    i = np.arange(K, dtype=np.float64)
    return np.where(i > (K >> 1), (i - K) / K, i / K)


def ldos(tre, tim, mu, Vre, Vim, hops, b, K, n, w, os_, omegas, eta, tiles):
    k = kgrid(K)
    KX, KY = np.meshgrid(k, k, indexing="ij")
    nh = len(hops)
    # phase of each hopping P[h, kx, ky], added into its orbital pair o1*b + o2
    P = np.empty((nh, K, K), np.complex128)
    pair = np.empty(nh, np.int64)
    for i, hp in enumerate(hops):
        arg = 2 * np.pi * (hp["rx"] * KX + hp["ry"] * KY)
        if hp["anti"]:
            arg = -arg
        P[i] = np.exp(1j * arg)
        pair[i] = hp["o1"] * b + hp["o2"]
    P = torch.from_numpy(P)
    t = torch.complex(tre, tim)
    eye = torch.eye(b, dtype=torch.complex128)
    Hp = torch.zeros((K, K, b * b), dtype=torch.complex128).index_add(
        2, torch.from_numpy(pair), (t[:, None, None] * P).permute(1, 2, 0))
    H = Hp.reshape(K, K, b, b) - torch.complex(mu, torch.zeros_like(mu)) * eye
    V = torch.complex(Vre, Vim)
    h = max((n >> 1) + w, 2 * w)
    nw = 2 * h + 1
    idx = torch.from_numpy((np.arange(nw) - h) % K)
    tl = torch.from_numpy(np.asarray(tiles, np.float64)).to(torch.complex128)
    n2 = n >> 1
    W = n * os_
    out = []
    for om in omegas:
        z = complex(om, eta)
        G = torch.linalg.inv(z * eye - H)
        GR = torch.fft.ifft2(G, dim=(0, 1))
        Gw = GR[idx][:, idx]
        T = torch.linalg.solve(eye - V @ Gw[h, h], V)
        a = torch.zeros((n, n, os_, os_, b), dtype=torch.complex128)
        bb = torch.zeros((n, n, os_, os_, b), dtype=torch.complex128)
        for rx in range(-w, w + 1):
            for ry in range(-w, w + 1):
                wt = tl[:, (w - rx) * os_:(w - rx + 1) * os_, (w - ry) * os_:(w - ry + 1) * os_]  # o,ii,jj
                sa = Gw[h - n2 + rx:h - n2 + rx + n, h - n2 + ry:h - n2 + ry + n]  # cells x,y: G(pos+R)
                sb = torch.flip(Gw[h + n2 - rx - n + 1:h + n2 - rx + 1, h + n2 - ry - n + 1:h + n2 - ry + 1],
                                dims=(0, 1))  # G(-pos-R)
                a = a + torch.einsum("oij,xyop->xyijp", wt, sa)
                bb = bb + torch.einsum("oij,xyqo->xyijq", wt, sb)
        imp = torch.einsum("xyijp,pq,xyijq->xyij", a, T, bb).imag
        bulk = torch.zeros((os_, os_), dtype=torch.float64)
        for r1x in range(-w, w + 1):
            for r1y in range(-w, w + 1):
                w1 = tl[:, (w - r1x) * os_:(w - r1x + 1) * os_, (w - r1y) * os_:(w - r1y + 1) * os_].real
                for r2x in range(-w, w + 1):
                    for r2y in range(-w, w + 1):
                        w2 = tl[:, (w - r2x) * os_:(w - r2x + 1) * os_, (w - r2y) * os_:(w - r2y + 1) * os_].real
                        g = Gw[r2x - r1x + h, r2y - r1y + h].imag  # [o2][o1]
                        bulk = bulk + torch.einsum("aij,bij,ba->ij", w1, w2, g)
        rho = -(imp + bulk[None, None])  # x, y, ii, jj
        out.append(rho.permute(1, 3, 0, 2).reshape(W * W))  # (y, jj, x, ii) -> Y*W + X
    return torch.stack(out)
