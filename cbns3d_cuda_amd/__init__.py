"""cbns3d_cuda_amd -- MI355X-native explicit SSP-RK3 step of a curvilinear
single-block compressible Navier-Stokes solver (characteristic WENO5 with
Roe-averaged eigenvectors, Lax-Friedrichs face flux, 4th-order central
viscous terms), behind the C ABI of include/cbns.h.

The compute is hand-written HIP for gfx950 in csrc/, built into libcbns.so
next to this file; block.py mirrors the reference's block3d_cuda launcher API
over that ABI.
"""
from .block import (BC_PERIODIC, BC_REF, BUF_Q, BUF_QP, FORM_FP32, FORM_FP64, G, NG_REF,
                    Block, CbnsError, Params, lib, make_params)

__all__ = ["Block", "CbnsError", "Params", "make_params", "lib", "G", "NG_REF", "BC_REF",
           "BC_PERIODIC", "BUF_Q", "BUF_QP", "FORM_FP64", "FORM_FP32"]
