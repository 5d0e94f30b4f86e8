"""cqpi_wannierldos_device, the entry point that allocates nothing, against
the drop-in cqpi_wannierldos (host G0(R) and T per energy) on the same inputs:
two calls in a row on one context, the second with fewer energies, in every
precision; FP64 to 1e-10, FP32 and mixed to 1e-4 (and both against the
oracle).  A request beyond the fixed scratch is refused before any launch."""
import numpy as np
import pytest
import torch

import cqpi
import oracle
import refmath as rm
from cqpi import synth

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-10, "fp32": 1e-4, "mixed": 1e-4}


def _inputs(wl, h):
    """Per energy: G0(R) on the grid, its window (slot - h wrapped onto the
    grid) and T, from the oracle."""
    ev, V = oracle.calchamiltonian(wl.hops, wl.bands, wl.efermi, wl.K)
    idx = (np.arange(2 * h + 1) - h) % wl.K
    out = []
    for om in wl.energies:
        gR = oracle.realspace(oracle.greens_k(ev, V, om, wl.eta))
        T = oracle.tmatrix(wl.V, gR[0, 0])
        out.append((gR, np.ascontiguousarray(gR[np.ix_(idx, idx)]), T))
    return out


@pytest.mark.parametrize("prec", ["fp64", "mixed", "fp32"])
def test_device_entry_matches_dropin(gpu, prec):
    wl = synth.square_3orb(K=32, n=12, E=3)
    print("case: no-alloc entry point vs drop-in, 3 orbitals K=32 n=12 prec=%s" % prec, flush=True)
    q = cqpi.AmdQPI(wl.W, wl.K, wl.n, wl.window, wl.oversamp, wl.bands, wl.maxband, wl.spin, wl.tiles)
    try:
        q.set_precision(prec)
        h = q.window_half()
        data = _inputs(wl, h)
        ref = []
        for gR, _, T in data:
            q.wannierldos(T, gR)
            ref.append(q.retrieveResult())
        for sel in ([0, 1, 2], [2]):   # the second call is smaller
            win = torch.from_numpy(np.stack([data[e][1] for e in sel])).to("cuda")
            Ts = torch.from_numpy(np.stack([data[e][2] for e in sel])).to("cuda")
            out = torch.full((len(sel), wl.W * wl.W), float("nan"), dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            q.wannierldos_device(Ts.data_ptr(), win.data_ptr(), len(sel), out.data_ptr())
            torch.cuda.synchronize()
            got = out.cpu().numpy()
            for i, e in enumerate(sel):
                mx, rms = rm.rel_errors(got[i], ref[e])
                ora = oracle.wannierldos(data[e][0], data[e][2], wl.tiles, wl.n, wl.window, wl.oversamp)
                mo, _ = rm.rel_errors(got[i], ora)
                print("%s call %s energy %d: vs drop-in %.3e, vs oracle %.3e, identical %s"
                      % (prec, sel, e, mx, mo, np.array_equal(got[i], ref[e])), flush=True)
                assert np.all(np.isfinite(got[i]))
                assert mx < TOL[prec] and mo < TOL[prec]
    finally:
        q.close()


def test_device_entry_refuses_beyond_scratch(gpu):
    wl = synth.square_3orb(K=32, n=12, E=1)
    q = cqpi.AmdQPI(wl.W, wl.K, wl.n, wl.window, wl.oversamp, wl.bands, wl.maxband, wl.spin, wl.tiles)
    try:
        q.set_precision("mixed")
        buf = torch.zeros(16, dtype=torch.float64, device="cuda")
        # 65535 energies of fp16 planes exceed the fixed scratch: refused on
        # the host before any launch
        with pytest.raises(cqpi.CqpiError) as ei:
            q.wannierldos_device(buf.data_ptr(), buf.data_ptr(), 65535, buf.data_ptr())
        assert ei.value.code == 1 and "scratch" in str(ei.value)
        with pytest.raises(cqpi.CqpiError):
            q.wannierldos_device(buf.data_ptr(), buf.data_ptr(), 0, buf.data_ptr())
    finally:
        q.close()
