"""rsr_amd: MI355X (gfx950) fitness-evaluation engine for RSR.

C ABI: include/rsr_amd.h, implemented by librsr_amd.so (HIP kernels + C++ runtime).
Python surface: rsr_amd.api (ctypes), rsr_amd.isa (reference instruction set),
rsr_amd.workloads (synthetic datasets and populations).
"""
from . import isa  # noqa: F401
