"""cbns_calc_gradient_noalloc (the entry point for callers that cannot
allocate device memory) against cbns_calc_gradient through the C ABI: bit for
bit on two calls in a row, the second with a smaller input, and the
reference's gradient gtest answers (test/gradient_test.cpp:20-150)."""
import numpy as np
import pytest

import ref_cases
from cbns3d_cuda_amd.block import (NOALLOC_MAX_NODES, CbnsError, ERR_SIZE, calc_gradient,
                                   calc_gradient_noalloc)

pytestmark = pytest.mark.gpu


def test_noalloc_gradient_matches_allocating_entry_point():
    f1, exp, tol = ref_cases.gradient_fun01()      # 50 x 100 x 150
    print("noalloc gradient, call 1:", f1.shape, flush=True)
    got1 = calc_gradient_noalloc(f1)
    ref1 = calc_gradient(f1)
    for g, r, e in zip(got1, ref1, exp):
        assert np.array_equal(g, r)
        assert np.abs(g - e).max() < tol
    # a smaller input: the array keeps the first call's larger contents
    rng = np.random.default_rng(5)
    KM, JM, IM = 23, 31, 17
    k, j, i = np.meshgrid(np.arange(KM, dtype=float), np.arange(JM, dtype=float),
                          np.arange(IM, dtype=float), indexing="ij")
    f2 = np.sin(0.3 * i + rng.uniform()) * np.cos(0.2 * j) + 0.01 * k ** 3
    print("noalloc gradient, call 2:", f2.shape, flush=True)
    got2 = calc_gradient_noalloc(f2)
    ref2 = calc_gradient(f2)
    for g, r in zip(got2, ref2):
        assert np.array_equal(g, r)


def test_noalloc_gradient_size_limit():
    n = int(round(NOALLOC_MAX_NODES ** (1.0 / 3.0))) + 2
    with pytest.raises(CbnsError) as e:
        calc_gradient_noalloc(np.zeros((n, n, n)))
    assert e.value.code == ERR_SIZE
