"""GPU parity of the band-structure histogram (bsfile=,
tightbind::calcbst3dhisthr, src/tightbinding.cc:1158-1193) against the
oracle, through the C ABI and through the driver."""
import subprocess

import numpy as np
import pytest

import cqpi
import oracle
from cqpi import synth
from test_driver import DRIVER

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("b,scmodel", [(3, False), (5, False), (6, True)])
def test_bands_histogram_matches_oracle(gpu, b, scmodel):
    wl = synth.bdg_square(3, K=16, n=4, E=1, seed=71) if scmodel else synth.square_norb(b, K=16, n=4, E=1, seed=70)
    print("case: band histogram b=%d scmodel=%s pix=12 os=3" % (wl.bands, scmodel), flush=True)
    with cqpi.engine_for(wl) as q:
        got = q.bands_histogram(12, 3, -0.3, 0.3, 13)
    ref = oracle.bands_histogram(wl.hops, wl.bands, wl.efermi, 12, 3, -0.3, 0.3, 13, scmodel=wl.scmodel)
    assert np.abs(ref).max() > 0
    assert np.max(np.abs(got - ref)) < 1e-11 * np.abs(ref).max()


def test_driver_bsfile_matches_oracle(tmp_path, gpu):
    wl = synth.square_3orb(K=16, n=4, E=1)
    p = synth.write_inputs(str(tmp_path), wl, extra={"bsfile": '"bs.idl"', "bslattice": "10", "bsoversamp": "2",
                                                     "bslayers": "9", "bsenergies": "(-0.2,0.2)"},
                           output="nomode")
    r = subprocess.run([DRIVER, p], capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr + r.stdout
    assert "Bandstructure calculation" in r.stdout
    data, hdr = synth.read_idl(str(tmp_path / "bs.idl"))
    assert data.shape == (10, 10, 9) and hdr["bias"] == [-0.2, 0.2] and hdr["offset"] == [-0.5, -0.5]
    hops, b, _ = oracle.read_hr(str(tmp_path / "model.dat"))
    ref = oracle.bands_histogram(hops, 3, wl.efermi, 10, 2, -0.2, 0.2, 9)
    assert np.max(np.abs(data - ref)) < 1e-6 * np.abs(ref).max()   # float32 file
