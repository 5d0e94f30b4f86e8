"""CPU checks of the measurement plumbing: kernel-name shortening in the
profile summariser and bench.py's lookup of measured traffic in the committed
profile summary."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from cqpi.pmcsummary import short_name  # noqa: E402


@pytest.mark.parametrize("raw,short", [
    ("void (anonymous namespace)::k_res_ky_s<HIP_vector_type<float, 2u>, 3, 512>(Geom, double const*)",
     "k_res_ky_s<f32, 3, 512>"),
    ("void (anonymous namespace)::k_contract<double, 3, 4>(Geom, HIP_vector_type<double, 2u> const*)",
     "k_contract<double, 3, 4>"),
    ("(anonymous namespace)::k_tmatrix(Geom, HIP_vector_type<double, 2u> const*)", "k_tmatrix"),
    ("_ZN12_GLOBAL__N_16k_kx_sI15HIP_vector_typeIdLj2EELi512EEEv4Geom7MixGeomPKS2_S5_idPS2_PDF16_",
     "k_kx_s<f64, 512>"),
    ("_ZN12_GLOBAL__N_111k_cmix_aregILi1ELi4ELi4ELi4ELi8ELi3EEEv4Geom7MixGeomPKDF16_PK15HIP_vector_typeIdLj2EE",
     "k_cmix_areg<1, 4, 4, 4, 8, 3>"),
])
def test_short_name(raw, short):
    assert short_name(raw) == short


def test_bench_reads_measured_traffic():
    import json
    import bench
    f32 = bench.measured_traffic("k_res_ky_s", "mixed")
    f64 = bench.measured_traffic("k_res_ky_s", "fp64")
    assert f32 and f64 and f32[0] != f64[0]   # the line type selects the row
    assert f32[1] == 51                        # the probe's energies are recorded
    txt = open(bench.TRAFFIC_MD["square_3orb"]).read()
    rows = json.loads(txt.split("```json", 1)[1].split("```", 1)[0])["kernels"]
    w32 = [r["write_bytes"] for k, r in rows.items() if k.startswith("k_res_ky_s<f32")]
    w64 = [r["write_bytes"] for k, r in rows.items() if k.startswith("k_res_ky_s<f64")]
    assert w32 and w64 and w32[0] < 0.6 * w64[0]   # FP32 rows: half the bytes written
    assert bench.measured_traffic("k_cmix_tf", "mixed")[0] > 0
    assert bench.measured_traffic("no_such_kernel", "mixed") is None


def test_fused_transform_bytes_follow_the_window_of_the_reference():
    import bench
    from cqpi import synth
    wl = synth.square_3orb()
    nk = 128 + 2 * 2 + 2
    assert bench.fused_transform_bytes(wl, 51) == 512 * 512 * (16 * 9 + 8 * 3) + 51 * 16 * 9 * nk * nk
