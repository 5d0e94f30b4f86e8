"""Gradient of an LDOS-map loss w.r.t. hopping amplitudes, mu and V.

L = sum (rho - rho_target)^2 with the target computed from perturbed
parameters.  The hand-written adjoint (cqpi_gradient) is checked against
 * torch autograd through a plain FP64 PyTorch implementation (refgrad.py),
   itself pinned to the oracle on the CPU, and
 * central finite differences of the device forward (h = 1e-6, FP64) along
   Hermitian directions (a hopping and its (-R, swapped-orbital) partner move
   together, so H(k) stays Hermitian)."""
import numpy as np
import pytest
import torch

import cqpi
import oracle
import refgrad
import refmath as rm
from cqpi import synth


def _job(E=3):
    wl = synth.square_norb(3, K=32, n=8, E=E, seed=41, w=2, os_=4, elo=-0.05, ehi=0.05, eta=0.01)
    wl.efermi = 0.05
    wl.V = np.array([[0.5, 0.1 - 0.05j, 0.0], [0.1 + 0.05j, 0.4, 0.02j], [0.0, -0.02j, 0.3]])
    target = synth.perturbed(wl)
    return wl, target


def _torch_params(wl):
    return (torch.tensor(wl.hops["tre"], dtype=torch.float64, requires_grad=True),
            torch.tensor(wl.hops["tim"], dtype=torch.float64, requires_grad=True),
            torch.tensor(wl.efermi, dtype=torch.float64, requires_grad=True),
            torch.tensor(wl.V.real.copy(), dtype=torch.float64, requires_grad=True),
            torch.tensor(wl.V.imag.copy(), dtype=torch.float64, requires_grad=True))


def _torch_rho(params, wl):
    return refgrad.ldos(*params, wl.hops, wl.bands, wl.K, wl.n, wl.window, wl.oversamp, wl.energies, wl.eta,
                        wl.tiles)


def test_torch_reference_matches_oracle():
    wl, _ = _job()
    wl.energies = wl.energies[:2]
    with torch.no_grad():
        rho = _torch_rho(_torch_params(wl), wl).numpy()
    ref = oracle.cldos_run(wl)
    for e in range(2):
        assert rm.rel_errors(rho[e], ref[e])[0] < 1e-11


def test_torch_reference_gradient_matches_its_finite_differences():
    """The reference gradient itself: autograd vs central differences of the
    same torch forward for mu and one V entry."""
    wl, target = _job()
    wl.energies = wl.energies[:1]
    tgt = torch.from_numpy(oracle.cldos_run(synth.Workload(**{**target.__dict__, "energies": wl.energies})))
    p = _torch_params(wl)
    L = ((_torch_rho(p, wl) - tgt) ** 2).sum()
    L.backward()
    h = 1e-6
    with torch.no_grad():
        for which, get in (("mu", lambda q: q[2]), ("V00", lambda q: q[3][0, 0])):
            qp = [x.detach().clone() for x in p]
            qm = [x.detach().clone() for x in p]
            if which == "mu":
                qp[2] += h
                qm[2] -= h
            else:
                qp[3][0, 0] += h
                qm[3][0, 0] -= h
            fd = (((_torch_rho(qp, wl) - tgt) ** 2).sum() - ((_torch_rho(qm, wl) - tgt) ** 2).sum()) / (2 * h)
            g = p[2].grad if which == "mu" else p[3].grad[0, 0]
            assert abs(fd - g) < 1e-6 * max(1.0, abs(float(g))), (which, float(fd), float(g))


def _gpu_loss_and_grad(q, wl, tgt):
    q.set_model(wl.hops, wl.efermi)
    q.set_impurity(wl.V)
    q.set_keep_windows(True)
    rho = q.run(wl.energies, wl.eta)
    L = float(((rho - tgt) ** 2).sum())
    return L, q.gradient(2.0 * (rho - tgt))


@pytest.mark.gpu
@pytest.mark.parametrize("E", [3, 9])
def test_gradient_matches_torch_autograd(gpu, E):
    # E = 9: the forward splits the energies over two streams
    wl, target = _job(E)
    print("case: gradient, 3-orbital square lattice K=32 n=8 w=2 os=4 E=%d" % E, flush=True)
    tgt = oracle.cldos_run(target)
    with cqpi.engine_for(wl) as q:
        _, (gh, gmu, gV) = _gpu_loss_and_grad(q, wl, tgt)
    p = _torch_params(wl)
    L = ((_torch_rho(p, wl) - torch.from_numpy(tgt)) ** 2).sum()
    L.backward()
    ref_h = np.stack([p[0].grad.numpy(), p[1].grad.numpy()], axis=1)
    assert np.linalg.norm(gh - ref_h) < 1e-9 * np.linalg.norm(ref_h)
    assert abs(gmu - float(p[2].grad)) < 1e-9 * max(1.0, abs(float(p[2].grad)))
    ref_V = p[3].grad.numpy() + 1j * p[4].grad.numpy()
    assert np.linalg.norm(gV - ref_V) < 1e-9 * np.linalg.norm(ref_V)


@pytest.mark.gpu
@pytest.mark.parametrize("b,K", [(4, 32), (5, 32), (5, 256), (2, 1024)])
def test_gradient_multiband_matches_torch_autograd(gpu, b, K):
    # 4 and 5 bands take the two-kernel ky adjoint through the Gbar(k)
    # workspace; 2 bands at K = 1024 the one-kernel path with 4 k-points a thread
    wl = synth.square_norb(b, K=K, n=8, E=2, seed=43 + b, w=2, os_=4, elo=-0.05, ehi=0.05, eta=0.01)
    wl.efermi = 0.02
    target = synth.perturbed(wl)
    print("case: gradient, %d-orbital square lattice K=%d n=8 E=2" % (b, K), flush=True)
    tgt = oracle.cldos_run(target)
    with cqpi.engine_for(wl) as q:
        _, (gh, gmu, gV) = _gpu_loss_and_grad(q, wl, tgt)
    p = _torch_params(wl)
    L = ((_torch_rho(p, wl) - torch.from_numpy(tgt)) ** 2).sum()
    L.backward()
    ref_h = np.stack([p[0].grad.numpy(), p[1].grad.numpy()], axis=1)
    assert np.linalg.norm(gh - ref_h) < 1e-9 * np.linalg.norm(ref_h)
    assert abs(gmu - float(p[2].grad)) < 1e-9 * max(1.0, abs(float(p[2].grad)))
    ref_V = p[3].grad.numpy() + 1j * p[4].grad.numpy()
    assert np.linalg.norm(gV - ref_V) < 1e-9 * np.linalg.norm(ref_V)


@pytest.mark.gpu
def test_gradient_energy_blocks_add_up(gpu):
    # 3 bands at K = 1024: the two-kernel ky adjoint, whose 1 GiB workspace
    # holds 7 energies of Gbar(k); 8 energies run as blocks of 7 and 1, whose
    # sums must equal the gradients of the two energy subsets run apart
    wl = synth.square_norb(3, K=1024, n=8, E=8, seed=47, w=2, os_=4, elo=-0.05, ehi=0.05, eta=0.01)
    print("case: gradient energy blocks, 3-orbital K=1024 n=8 E=8", flush=True)
    tgt = np.zeros((8, wl.W * wl.W))  # L = sum rho^2
    parts = []
    with cqpi.engine_for(wl) as q:
        for sl in (slice(0, 8), slice(0, 7), slice(7, 8)):
            w2 = synth.Workload(**{**wl.__dict__, "energies": wl.energies[sl]})
            parts.append(_gpu_loss_and_grad(q, w2, tgt[sl])[1])
    full, a, b = parts
    for i in range(3):
        ref = a[i] + b[i]
        assert np.linalg.norm(np.asarray(full[i] - ref)) <= 1e-10 * np.linalg.norm(np.asarray(ref)), i


@pytest.mark.gpu
def test_gradient_matches_central_differences(gpu):
    wl, target = _job()
    print("case: gradient vs finite differences, K=32 n=8 E=3", flush=True)
    tgt = oracle.cldos_run(target)
    h = 1e-6
    hops = wl.hops
    key = {(int(x["rx"]), int(x["ry"]), int(x["o1"]), int(x["o2"])): i for i, x in enumerate(hops)}
    with cqpi.engine_for(wl) as q:
        _, (gh, gmu, gV) = _gpu_loss_and_grad(q, wl, tgt)

        def loss(w2):
            q.set_model(w2.hops, w2.efermi)
            q.set_impurity(w2.V)
            q.set_keep_windows(False)
            rho = q.run(w2.energies, w2.eta)
            return float(((rho - tgt) ** 2).sum())

        checks = []
        # Hermitian hopping directions: (R, o1, o2) with its partner (-R, o2, o1)
        for i in (5, 40, 77):
            x = hops[i]
            j = key[(-int(x["rx"]), -int(x["ry"]), int(x["o2"]), int(x["o1"]))]
            for part in ("tre", "tim"):
                if part == "tim" and i == j:
                    continue
                sgn = 1.0 if part == "tre" else -1.0
                wp = synth.Workload(**{**wl.__dict__, "hops": hops.copy()})
                wm = synth.Workload(**{**wl.__dict__, "hops": hops.copy()})
                wp.hops[i][part] += h
                wm.hops[i][part] -= h
                if j != i:
                    wp.hops[j][part] += sgn * h
                    wm.hops[j][part] -= sgn * h
                col = 0 if part == "tre" else 1
                g = gh[i, col] + (sgn * gh[j, col] if j != i else 0.0)
                checks.append(((loss(wp) - loss(wm)) / (2 * h), g, "hop%d.%s" % (i, part)))
        wp = synth.Workload(**{**wl.__dict__, "efermi": wl.efermi + h})
        wm = synth.Workload(**{**wl.__dict__, "efermi": wl.efermi - h})
        checks.append(((loss(wp) - loss(wm)) / (2 * h), gmu, "mu"))
        for (r, c, part) in ((0, 0, "re"), (0, 1, "re"), (0, 1, "im")):
            d = np.zeros((3, 3), complex)
            if r == c:
                d[r, c] = h
                g = gV[r, c].real
            elif part == "re":
                d[r, c] = d[c, r] = h
                g = gV[r, c].real + gV[c, r].real
            else:
                d[r, c] = 1j * h
                d[c, r] = -1j * h
                g = gV[r, c].imag - gV[c, r].imag
            wp = synth.Workload(**{**wl.__dict__, "V": wl.V + d})
            wm = synth.Workload(**{**wl.__dict__, "V": wl.V - d})
            checks.append(((loss(wp) - loss(wm)) / (2 * h), g, "V%d%d.%s" % (r, c, part)))
    for fd, g, name in checks:
        print("%-12s fd %+.9e  adjoint %+.9e" % (name, fd, g))
        assert abs(fd - g) < 1e-5 * max(abs(g), 1e-3), name


@pytest.mark.gpu
def test_autograd_function(gpu):
    from cqpi.autograd import cldos

    wl, target = _job()
    print("case: autograd.Function, K=32 n=8 E=3", flush=True)
    tgt = torch.from_numpy(oracle.cldos_run(target)).to("cuda")
    dev = "cuda"
    p = [x.detach().to(dev).requires_grad_(True) for x in _torch_params(wl)]
    with cqpi.engine_for(wl) as q:
        rho = cldos(*p, q, wl.hops, wl.energies, wl.eta)
        L = ((rho - tgt) ** 2).sum()
        L.backward()
        _, (gh, gmu, gV) = _gpu_loss_and_grad(q, wl, tgt.cpu().numpy())
    assert np.allclose(p[0].grad.cpu().numpy(), gh[:, 0], rtol=1e-12, atol=1e-14)
    assert np.allclose(p[1].grad.cpu().numpy(), gh[:, 1], rtol=1e-12, atol=1e-14)
    assert abs(float(p[2].grad) - gmu) < 1e-12 * max(1.0, abs(gmu))
    assert np.allclose(p[3].grad.cpu().numpy() + 1j * p[4].grad.cpu().numpy(), gV, rtol=1e-12, atol=1e-14)
