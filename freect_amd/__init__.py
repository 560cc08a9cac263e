"""freect_amd -- MI355X-native WFBP engine (drop-in for FreeCT's fct_wfbp path).

Native code: freect_amd/lib/libfct_wfbp_gpu.so (HIP kernels + C ABI, see
include/fct_wfbp_gpu.h).  Python modules here are the ctypes mirror of that
ABI (wfbp), scan/phantom descriptions (scans) and the torch.autograd harness
of the adjoint (autograd).
"""
from . import scans  # noqa: F401

__all__ = ["scans"]
