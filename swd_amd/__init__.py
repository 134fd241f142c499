"""swd_amd - MI355X (gfx950) engine for the SWD hot path.

The compute lives in ``libswd.so`` (HIP kernels + C ABI, include/swd_abi.h);
the modules here mirror the reference's MATLAB/MEX entry points over ctypes:
``fd`` (eigen_staggerfd_*, eigen_e2drtm_*, staggerfd_eigen(d), e2drtm_eigen(d)).
"""
from ._lib import SwdError, gpu_available, load  # noqa: F401

__all__ = ["SwdError", "gpu_available", "load"]
