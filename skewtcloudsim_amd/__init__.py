"""MI355X-native implementation of finntie/skewTCloudSim's simulation step.

The step (environmentGPU::updateGPU) runs as hand-written HIP kernels for
gfx950 in libskt.so; this package is the Python mirror of its C ABI
(include/skt.h). Nothing here computes on the CPU.
"""
from ._lib import FIELD, FIELDS, NEUMANN, DIRICHLET, CUSTOM, PCG_REFERENCE, PCG_STANDARD, load_library
from .environment import (EnsembleGPU, EnvironmentGPU, ReadTable, SktError, default_config, make_sounding,
                          terrain)
from .serving import Server, job_footprint, run_poisson

__all__ = ["FIELD", "FIELDS", "NEUMANN", "DIRICHLET", "CUSTOM", "PCG_REFERENCE", "PCG_STANDARD", "load_library",
           "EnsembleGPU", "EnvironmentGPU", "ReadTable", "SktError", "default_config", "make_sounding", "terrain",
           "Server", "job_footprint", "run_poisson"]
