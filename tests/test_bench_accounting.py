"""bench.py's work accounting and its command-line contract (CPU only).

The roofline's `achieved` is algorithmic work per launch / measured launch
time, so the face and byte counts it uses are checked here against counts
made by hand, and the one-GPU contract (`--gpus N` with N > 1 exits with a
message, before any GPU work) is run as a subprocess.
"""
import os
import subprocess
import sys
from types import SimpleNamespace

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def params(n, bc, viscous=1):
    return SimpleNamespace(IM=n[0], JM=n[1], KM=n[2], bc=bc, viscous=viscous)


def test_face_counts_periodic_and_reference_axes():
    # periodic axis of n nodes: faces -1 .. n-1 (n + 1 per line), every node
    # updated; reference axis: faces 0 .. n-2 (n - 1), nodes 1 .. n-2 updated
    p = params((8, 6, 5), (1, 0, 1))
    faces, upd = bench.face_counts(p)
    assert upd == [8, 4, 5]
    assert faces == [(8 + 1) * 4 * 5, (6 - 1) * 8 * 5, (5 + 1) * 8 * 4]


def test_march_flops_cover_every_face_once():
    p = params((16, 16, 16), (1, 1, 1))
    faces, _ = bench.face_counts(p)
    eta = bench.kernel_work("march_eta", p)
    zeta = bench.kernel_work("march_zeta", p)
    assert eta[0] == zeta[0] == "flop"
    assert eta[1] + zeta[1] == sum(faces) * bench.FLOPS_PER_FACE
    assert eta[1] == 16 * 16 * 17 * bench.FLOPS_PER_FACE  # eta-faces only


def test_wrap_bytes_follow_the_fields_a_step_refreshes():
    n = 16
    shell = (n + 8) ** 3 - n ** 3
    visc = bench.kernel_work("bc", params((n, n, n), (1, 1, 1), viscous=1))
    invisc = bench.kernel_work("bc", params((n, n, n), (1, 1, 1), viscous=0))
    # stages 1-2: Q + u,v,w,T (viscous) or Q alone; stage 3: Q + all primitives
    assert visc == ("byte", int(shell * (9 + 9 + 11) / 3 * 16))
    assert invisc == ("byte", int(shell * (5 + 5 + 10) / 3 * 16))


def test_dt_partials_match_the_flux_pass_tiling():
    # one maximum per flux-pass workgroup: 60-node tiles along i, 4 rows
    # along j, 8 nodes along k over the viscous range [-2, N+1]
    p = params((256, 256, 256), (1, 1, 1))
    kind, b = bench.kernel_work("dt", p)
    assert kind == "byte"
    assert b == ((260 + 59) // 60) * (260 // 4) * ((260 + 7) // 8) * 8


@pytest.mark.parametrize("n", [2, 8])
def test_more_than_one_gpu_exits_with_a_message(n):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_step_hbm_sums_launches_per_step():
    # each kernel's committed bytes per launch times its launches per step,
    # over the step time; a kernel without a committed figure gives None
    p = params((256, 256, 256), (1, 1, 1))
    t = {"march_eta": 2e9, "march_zeta": 7e9, "visc_flux": 5.5e9, "bc": 0.3e9, "dt": 1e5}
    saved = bench.measured_traffic
    bench.measured_traffic = lambda name, p_, wl: t.get(name)
    try:
        kt = {"march_eta": (1.8, 6), "march_zeta": (4.8, 6), "visc_flux": (3.1, 6),
              "bc": (0.2, 6), "dt": (0.04, 2), "indicator": (0.0, 0)}
        r = bench.step_hbm(kt, 2, 10.0, p, "tgv")
        per_step = 3 * (2e9 + 7e9 + 5.5e9 + 0.3e9) + 1e5
        assert r["bytes_per_step"] == pytest.approx(per_step)
        assert r["achieved"] == pytest.approx(per_step / 10e-3 / 1e9)
        assert r["frac"] == pytest.approx(r["achieved"] / bench.PEAK_HBM_GBS)
        kt["visc_div"] = (0.6, 6)
        assert bench.step_hbm(kt, 2, 10.0, p, "tgv") is None
    finally:
        bench.measured_traffic = saved
