"""bielcc_amd - MI355X (gfx950) implementation of BIELcc's Stratton-Chu
K / R / K0 path: cell-pair quadrature kernels in HIP behind a C ABI
(include/bielcc_hip.h), with a Python host layer.

The HIP library is loaded lazily by `Context()`; importing the package does
not touch the GPU.
"""
from .api import (BielccError, Context, DeviceArray, Mesh, Plan, NumParam, KernelParam, IntegralParam, num_param,
                  kernel_param, OPS, KERNELS, DEFAULT_NUM_PARAM, PREC_FP64, PREC_MIXED)
from . import mesh, rwg, tune

__all__ = ["BielccError", "Context", "DeviceArray", "Mesh", "Plan", "NumParam", "KernelParam", "IntegralParam",
           "num_param", "kernel_param", "OPS", "KERNELS", "DEFAULT_NUM_PARAM", "PREC_FP64", "PREC_MIXED",
           "mesh", "rwg", "tune"]
