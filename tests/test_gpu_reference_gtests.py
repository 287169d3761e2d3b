"""The reference's own known-answer tests (test/gradient_test.cpp:20-150,
test/metrics_test.cpp:77-601), run through the GPU library's C ABI:

  cbns_calc_gradient          f = i^2 + j^2 + k^2 and f = i j k on 50x100x150:
                              max |error| below the gtests' 1e-5, and bit for
                              bit the oracle's gradient
  cbns_set_mesh +             tanh-stretched box, rotated box, conical shell:
  cbns_get_metrics            per-component MSE within the gtests' bounds
                              (2e-2, 1e-8, 1e-3), and bit for bit the oracle's
                              metrics and Jacobian
"""
import numpy as np
import pytest

import ref_cases
from cbns3d_cuda_amd.block import BC_REF, Block, calc_gradient, make_params

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", [ref_cases.gradient_fun01, ref_cases.gradient_fun02])
def test_gradient_reference_gtests(oracle, case):
    f, exp, tol = case()
    print("gradient gtest", case.__name__, f.shape, flush=True)
    got = calc_gradient(f)
    ref = oracle.gradient_nodes(f)
    for g, e, r in zip(got, exp, ref):
        err = np.abs(g - e).max()
        print("  max |error| %.3e (gtest bound %.0e)" % (err, tol), flush=True)
        assert err < tol
        assert np.array_equal(g, r)


@pytest.mark.parametrize("case", [ref_cases.metrics_rectilinear, ref_cases.metrics_rotation,
                                  ref_cases.metrics_conical_shell])
def test_metrics_reference_gtests(oracle, case):
    (x, y, z), exp, tol = case()
    KM, JM, IM = x.shape
    print("metrics gtest", case.__name__, x.shape, flush=True)
    b = Block(make_params(IM, JM, KM, bc=(BC_REF,) * 3))
    try:
        b.calc_metrics(x, y, z)
        got = b.get_metrics()
    finally:
        b.free()
    mse = ref_cases.metric_mse(got, exp)
    print("  per-component MSE", ["%.2e" % m for m in mse], "bound %.0e" % tol, flush=True)
    assert max(mse) <= tol
    o = oracle.Oracle(oracle.make_params(IM, JM, KM, bc=(0, 0, 0)))
    o.set_mesh(x, y, z)
    for c in range(10):
        assert np.array_equal(got[c], o.inner(o.met[c])), c
