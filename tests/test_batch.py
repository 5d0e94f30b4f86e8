"""Batches of independent instances in one captured HIP graph, and the
autotuner: both must leave the numbers unchanged.

A batch runs every instance's chain (eigensolve, resolvent, transforms,
T-matrix, contraction) from one graph launch; each replay must reproduce the
eager per-instance run bit for bit (same kernels, same inputs), and the FP64
batch must match the oracle to 1e-10 relative."""
import numpy as np
import pytest

import cqpi
import oracle
import refmath as rm
from cqpi import synth

pytestmark = pytest.mark.gpu


def _instances():
    # uneven shapes: orbitals {1,3,5}, k-grids {16,32,64,24,128}, maps, energies;
    # the 3- and 5-orbital instances at K = 64 and 128 transform only the
    # orbital pairs r <= c (real hoppings)
    spec = [(1, 32, 8, 3, 2, 4), (3, 64, 12, 4, 2, 4), (5, 16, 6, 2, 1, 2), (3, 24, 9, 3, 1, 2),
            (1, 64, 16, 2, 1, 8), (3, 32, 8, 5, 2, 4), (5, 128, 8, 2, 2, 4)]
    out = []
    for i, (b, K, n, E, w, os_) in enumerate(spec):
        out.append(synth.square_norb(b, K=K, n=n, E=E, seed=7000 + i, w=w, os_=os_, eta=0.02))
    return out


def _eager(wl, precision):
    with cqpi.engine_for(wl, precision) as q:
        return q.run(wl.energies, wl.eta)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_graph_matches_eager_and_oracle(gpu, precision):
    wls = _instances()
    print("case: batch of %d instances, %s" % (len(wls), precision), flush=True)
    eager = [_eager(wl, precision) for wl in wls]
    with cqpi.Batch(wls, precision) as b:
        for rep in range(3):
            ms = b.run()
            assert ms > 0
            for i, wl in enumerate(wls):
                got = b.maps(i)
                # replays against the eager run of the same instance, exactly
                assert np.array_equal(got, eager[i]), (rep, i)
            # eager work on another context between replays
            _eager(wls[0], precision)
        info = b.info()
        assert info["graph_nodes"] >= len(wls)
    tol = 1e-10 if precision == "fp64" else 1e-4
    for i in range(len(wls)):
        ref = oracle.cldos_run(wls[i])
        for e in range(ref.shape[0]):
            mx, rms = rm.rel_errors(eager[i][e], ref[e])
            assert mx < tol, (i, e, mx, rms)


@pytest.mark.skip(reason="a captured per-instance fork/join (energy chunks split over an "
                         "instance's two streams inside the batch graph) ended in a host "
                         "segmentation fault in its first run, also for this 3-instance batch "
                         "and with non-blocking extra streams; batch instances run on one "
                         "stream until its cause is shown")
def test_batch_with_split_energy_chunks(gpu):
    # 8 or more energies: each instance's chunk would be split over its two
    # streams, joined to the batch branch inside the captured graph
    spec = [(3, 32, 8, 12, 2, 4), (1, 64, 16, 8, 1, 8), (5, 16, 6, 9, 1, 2)]
    wls = [synth.square_norb(b, K=K, n=n, E=E, seed=7300 + i, w=w, os_=os_, eta=0.02)
           for i, (b, K, n, E, w, os_) in enumerate(spec)]
    print("case: batch of %d instances with split chunks" % len(wls), flush=True)
    for precision in ("fp64", "mixed"):
        eager = [_eager(wl, precision) for wl in wls]
        with cqpi.Batch(wls, precision) as b:
            for rep in range(3):
                b.run()
                for i in range(len(wls)):
                    assert np.array_equal(b.maps(i), eager[i]), (precision, rep, i)
                _eager(wls[1], precision)


def test_batch_reports_singular_instance(gpu):
    wl = synth.square_1orb(K=16, n=4, E=1)
    ev, V = oracle.calchamiltonian(wl.hops, 1, 0.0, wl.K)
    g00 = oracle.realspace(oracle.greens_k(ev, V, 0.2, 0.05))[0, 0, 0, 0]
    wl.V = np.array([[1.0 / g00]])
    wl.energies = np.array([0.2])
    wl.eta = 0.05
    good = synth.square_norb(3, K=16, n=6, E=2, seed=7100, w=1, os_=2, eta=0.02)
    print("case: batch with a singular instance", flush=True)
    try:
        with cqpi.Batch([good, wl], "fp64") as b:
            b.run()
    except cqpi.CqpiError as e:
        assert e.code == 4
    else:
        pass  # rounding may leave a tiny nonzero pivot


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_autotune_leaves_results_unchanged(gpu, precision):
    # 10 energies: the two-stream split applies, so the streams knob is timed
    wl = synth.square_norb(3, K=64, n=16, E=10, seed=7200, w=2, os_=4, eta=0.01)
    print("case: autotune %s" % precision, flush=True)
    base = _eager(wl, precision)
    with cqpi.engine_for(wl, precision) as q:
        tpw, eb = q.autotune(wl.energies, wl.eta, reps=2)
        assert tpw in (0, 2, 4, 8) and 0 <= eb <= len(wl.energies)
        assert q.tuning() == (tpw, eb)
        ext = q.tuning_ext()
        assert ext[:2] == (tpw, eb) and ext[2] in (0, 1, 2) and ext[3] in (0, 1)
        assert np.array_equal(q.run(wl.energies, wl.eta), base)
        for t, e in ((2, 1), (4, 3), (8, 0)):
            q.set_tuning(t, e)
            assert np.array_equal(q.run(wl.energies, wl.eta), base), (t, e)
        # every waves-per-row and stream choice gives the same maps
        for js in (1, 2):
            for st in (1, 2):
                q.set_tuning_ext(4, 0, js, st)
                assert q.tuning_ext() == (4, 0, js, st)
                assert np.array_equal(q.run(wl.energies, wl.eta), base), (js, st)
    # a second context of the same shape reuses the remembered tuning
    with cqpi.engine_for(wl, precision) as q2:
        assert q2.autotune(wl.energies, wl.eta, reps=2) == (tpw, eb)
        assert q2.tuning_ext() == ext


def test_batch_autotune_then_run(gpu):
    wls = _instances()[:3]
    print("case: batch autotune", flush=True)
    eager = [_eager(wl, "mixed") for wl in wls]
    with cqpi.Batch(wls, "mixed") as b:
        b.autotune(reps=1)
        b.run()
        for i in range(len(wls)):
            assert np.array_equal(b.maps(i), eager[i])
