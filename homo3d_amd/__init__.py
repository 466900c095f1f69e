"""MI355X-native periodic homogenization solver (Python binding of the homo3d C ABI).

The compute path is libhomo3d.so (hand-written gfx950 HIP kernels + C++ drivers, built in-tree by
``homo3d_amd.build``).  There is no CPU fallback: importing the binding raises if the library is
missing, and every call raises ``Homo3dError`` on a non-zero status.

The classes mirror the reference's host API (lavenklau/homo3d):
  * ``HomoConfig``        cfg::HomoConfig (cmdline.h:39-60)
  * ``Homogenization``    homo::Homogenization (homogenization/homogenization.h:17-61)
  * ``OCOptimizer``       homo::OCOptimizer (optimization/OCOptimizer.h:7-20)
  * ``MMAOptimizer``      homo::MMAOptimizer (optimization/mmaOptimizer.h:12-34)
"""
import ctypes
import os

import numpy as np

from .api import (HOMO3D_OK, HOMO3D_NOT_CONVERGED, HOMO3D_DIVERGED, HOMO3D_BAD_ARG, HOMO3D_HIP_ERROR,
                  HOMO3D_NO_SUPPORT, SOLVER_MG, SOLVER_PCG, FIELD_U, FIELD_F, FIELD_R, SYM, Homo3dError,
                  lib, lib_path, DeviceBuffer, HomoConfig, Homogenization, OCOptimizer, MMAOptimizer, WarmStartNet, Batch, randc_density)

__all__ = ["HOMO3D_OK", "HOMO3D_NOT_CONVERGED", "HOMO3D_DIVERGED", "HOMO3D_BAD_ARG", "HOMO3D_HIP_ERROR",
           "HOMO3D_NO_SUPPORT", "SOLVER_MG", "SOLVER_PCG", "FIELD_U", "FIELD_F", "FIELD_R", "SYM",
           "Homo3dError", "lib", "lib_path", "DeviceBuffer", "HomoConfig", "Homogenization", "OCOptimizer", "MMAOptimizer", "WarmStartNet", "Batch",
           "randc_density"]
