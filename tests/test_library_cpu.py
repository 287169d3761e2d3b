"""The C ABI library without a GPU: it loads, exports every entry point
include/cbns.h declares, and its parameter struct has the layout the header
defines (checked against a gcc-compiled probe of the header)."""
import ctypes
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from cbns3d_cuda_amd import block

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cbns.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cbns_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = block.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert sorted(block.EXPORTS) == names


def test_params_layout_matches_header():
    src = r"""
    #include <stdio.h>
    #include <stddef.h>
    #include "cbns.h"
    int main(void) {
      printf("%zu %zu %zu %zu %zu\n", sizeof(cbns_params), offsetof(cbns_params, bc),
             offsetof(cbns_params, CFL), offsetof(cbns_params, T_inf),
             offsetof(cbns_params, angle_attack));
      return 0;
    }
    """
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "p.c")
        exe = os.path.join(d, "p")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        got = [int(t) for t in subprocess.run([exe], capture_output=True, text=True,
                                               check=True).stdout.split()]
    P = block.Params
    exp = [ctypes.sizeof(P), P.bc.offset, P.CFL.offset, P.T_inf.offset, P.angle_attack.offset]
    assert got == exp


def test_oracle_params_layout_matches_library(oracle):
    a, b = oracle.Params, block.Params
    assert ctypes.sizeof(a) == ctypes.sizeof(b)
    for (na, _), (nb, _) in zip(a._fields_, b._fields_):
        assert na == nb and getattr(a, na).offset == getattr(b, nb).offset


def test_case_meshes_have_positive_jacobian(oracle):
    from cbns3d_cuda_amd import cases
    from helpers import oracle_block
    for case in (cases.isentropic_vortex(16), cases.taylor_green(16), cases.c_grid_channel()):
        o = oracle_block(oracle, case)
        assert o.inner(o.met[9]).min() > 0.0, case.name


def test_taylor_green_initial_diagnostics(oracle):
    """<rho |u|^2>/2 = 1/8 and <rho |omega|^2>/2 = 3/8 for the Taylor-Green
    initial field (rho = 1 + O(M^2)), on the perturbed curvilinear mesh."""
    from cbns3d_cuda_amd import cases, diagnostics
    from helpers import oracle_block
    case = cases.taylor_green(48)
    o = oracle_block(oracle, case)
    met = o.inner(o.met)
    prim = case.prim
    assert diagnostics.kinetic_energy(prim, met) == pytest.approx(0.125, rel=2e-3)
    assert diagnostics.enstrophy(prim, met) == pytest.approx(0.375, rel=5e-3)


def test_noalloc_gradient_rejects_sizes_beyond_its_capacity():
    """cbns_calc_gradient_noalloc refuses an input larger than its fixed
    staging array before any device work (so this runs without a GPU)."""
    L = block.lib()
    n = int(round((block.NOALLOC_MAX_NODES + 1) ** (1.0 / 3.0))) + 1
    assert n ** 3 > block.NOALLOC_MAX_NODES
    dummy = np.zeros(1)
    p = dummy.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert L.cbns_calc_gradient_noalloc(n, n, n, p, p, p, p) == block.ERR_SIZE
    assert L.cbns_calc_gradient_noalloc(4, 5, 5, p, p, p, p) == 1  # CBNS_ERR_ARG
