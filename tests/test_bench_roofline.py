"""The benchmark's roofline arithmetic on the CPU: the algorithmic bytes of
the 1024^2, 10x10, 41^2 workload from the closed form (inputs and outputs
of the three operators once), the launch-level figures over frames per
launch, the step-level figures over frames per step, and the committed
counter summary read only at the frames per launch it was measured at."""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope='module')
def bench():
    spec = importlib.util.spec_from_file_location('bench_under_test', os.path.join(ROOT, 'bench.py'))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_operator_bytes_closed_form(bench):
    # x (1024^2): read by the forward and the f-adjoint, written (x-bar) by
    # the x-adjoint; y / g (984^2): written by the forward, read by both
    # adjoints; f (100 PSFs of 41^2): read by the forward and the x-adjoint,
    # written (f-bar) by the f-adjoint.  Each moves 3 x 4 bytes per value.
    x, y, f = 1024 * 1024, 984 * 984, 100 * 41 * 41
    assert bench.operator_bytes() == 12.0 * (x + y + f) == 26_219_184.0


def test_roofline_launch_and_step_levels(bench):
    kernels = {'cols_adj2': dict(avg_us=500.0, total_us_per_step=1000.0),
               'cols_fwd': dict(avg_us=350.0, total_us_per_step=700.0)}
    r = bench.operator_roofline(kernels, 32, 2.9, 64)
    assert r['kernel'] == 'cols_adj2'
    by = bench.operator_bytes() * 32
    assert r['alg_bytes_per_launch'] == by
    assert r['achieved'] == pytest.approx(by / 500e-6 / 1e9, abs=0.05)
    assert r['frac'] == pytest.approx(r['achieved'] / 8000.0, abs=1e-4)
    # the step carries 64 frames in 2.9 ms
    assert r['step']['hbm_gbs'] == pytest.approx(bench.operator_bytes() * 64 / 2.9e-3 / 1e9, abs=0.05)
    assert r['valu']['alg_flops'] == bench.operator_flops() * 64
    assert r['valu']['achieved_tflops'] == pytest.approx(bench.operator_flops() * 64 / 2.9e-3 / 1e12, abs=0.005)


def test_measured_traffic_only_at_its_frames_per_launch(bench):
    t = bench.measured_traffic('cols_adj2', 32)
    assert isinstance(t, int) and 1.0e9 < t < 3.0e9
    assert bench.measured_traffic('cols_adj2', 16) is None
    assert bench.measured_traffic('no_such_kernel', 32) is None
