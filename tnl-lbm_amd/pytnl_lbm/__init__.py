"""pytnl_lbm: the MI355X (gfx950) lattice update of TNL-LBM, exposed to PyTorch.

Mirrors the reference's Python module surface (TNL-LBM
pytnl_lbm/pytnl_lbm.cpp:15-73): a lattice block whose step is the fused
collide+stream+BC kernel, initial conditions, macroscopic output and FP64
reductions, the simulation driver (State with its hooks, Lattice units,
Actions counters, execute), outputs and checkpoints.  Every operation runs in libtnl_lbm_amd.so (HIP, gfx950);
nothing falls back to the CPU.
"""
from . import _abi
from ._abi import ADE_GEO, GEO, LBMError
from .lattice import D3Q7Lattice, D3Q27Lattice, from_storage, i16_class_ranges, to_storage
from .fields import (channel_geometry, duct_flow_fields, duct_profile, gaussian_blobs, inflow_control, smooth_perturbation,
                     tgv_fields, turbulence_fields)
from .adjoint import (CheckpointedRun, CoupledScalarRun, density_velocity, lattice_run, populations, scalar_run,
                      top_layer_variance)
from . import io
from .io import load_checkpoint, macro_view, save_checkpoint, write_vti
from .state import (OUT2D, OUT3D, OUT3DCUT, PRINT, PROBE1, PROBE2, PROBE3, SAVESTATE, STAT2_RESET, STAT_RESET, Actions,
                    Lattice, State, UniformDataWriter, counter, execute, getMacroView)

__all__ = ["_abi", "GEO", "ADE_GEO", "LBMError", "D3Q27Lattice", "D3Q7Lattice", "to_storage", "from_storage", "i16_class_ranges",
           "tgv_fields", "turbulence_fields", "gaussian_blobs", "channel_geometry", "duct_flow_fields", "duct_profile", "inflow_control",
           "smooth_perturbation", "CheckpointedRun", "density_velocity", "lattice_run", "populations", "CoupledScalarRun",
           "scalar_run", "top_layer_variance", "io", "save_checkpoint", "load_checkpoint", "macro_view", "write_vti",
           "Actions", "Lattice", "State", "UniformDataWriter", "counter", "execute", "getMacroView", "STAT_RESET",
           "STAT2_RESET", "PRINT", "OUT2D", "OUT3D", "OUT3DCUT", "PROBE1", "PROBE2", "PROBE3", "SAVESTATE"]
