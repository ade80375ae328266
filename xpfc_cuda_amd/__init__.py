"""xpfc_cuda_amd -- MI355X-native XPFC time stepping.

Python host mirror of the reference's operator surface (globals.cuh:98-145:
read_input, initialize/randn, power_functions, convolve_fields, euler_step,
scaling_step, FE_calc, Avg_calc, sse_calculation, write_vtk) over the C ABI of
``libxpfc.so`` (``include/xpfc.h``).  Every numeric call runs on the GPU
through that library; if it is missing, importing the bindings raises.
"""
from .api import (  # noqa: F401
    XPFC_FP64, XPFC_FP32_MIXED, XPFC_DCF_SUM, XPFC_DCF_MAX,
    XPFCError, Params, Sample, Context, BatchContext, MixedContext, lib, load_library, library_path,
    read_input, randn, philox4x32_10, reference_steps, instance_bytes, copy_bandwidth,
)

__all__ = ["XPFC_FP64", "XPFC_FP32_MIXED", "XPFC_DCF_SUM", "XPFC_DCF_MAX",
           "XPFCError", "Params", "Sample", "Context", "BatchContext", "MixedContext", "lib", "load_library",
           "library_path", "read_input", "randn", "philox4x32_10", "reference_steps",
           "instance_bytes", "copy_bandwidth"]
