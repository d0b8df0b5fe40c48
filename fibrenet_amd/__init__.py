"""fibrenet_amd - MI355X (gfx950) engine for FibreNet's particle-dynamics step.

The step runs in hand-written HIP kernels inside libfibrenet_amd.so, behind a
C ABI (include/fibrenet.h) with the call surface of FibreNet's
OpenclManager/SimulationManager.  This package holds the ctypes bindings and
the synthetic-network generator.
"""
from . import records
from .api import (Engine, Ensemble, FibreNetError, KERNEL_NAMES, LIB_PATH, Model, Simulation, bind_standard, lib,
                  run_loop_granular)

__all__ = ["Engine", "Ensemble", "FibreNetError", "KERNEL_NAMES", "LIB_PATH", "Model", "Simulation",
           "bind_standard", "lib", "records", "run_loop_granular"]
