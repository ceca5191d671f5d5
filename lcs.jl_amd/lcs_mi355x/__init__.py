"""MI355X (gfx950) backend for the LCS.jl per-time-step hot path.

All computation runs in hand-written HIP kernels inside the in-tree shared
library ``liblcs_mi355x.so`` (C ABI: include/lcs_mi355x.h).  This package is
the host-side mirror of the reference's configuration and phase API.
"""
from . import config, lcsio
from ._lib import LCSArgumentError, LCSError, LIBPATH, load
from .config import (Config, ConstantPowerLF, EnergyPreserveRCF, FlowParams, Grid, HSMAC, NoForcing,
                     ParticleConfig, RestartInit, SpectrumInit, TimeStep, from_dict, readconfig, to_dict)
from .core import (Batch, Box, autotune, boundary, boxmean, couple, fill_ghost, flow_stat, flow_timestep, forcing, localize,
                   get_tiles, makeindex, motion, particle_density, particle_stat, particle_timestep, particle_update, read_log, rhs, rhs_update, Server,
                   set_state, set_tiles, step, step_config, update)
from .simulate import build, simulate, stat

__all__ = [
    "Batch", "Box", "Config", "ConstantPowerLF", "EnergyPreserveRCF", "FlowParams", "Grid", "HSMAC", "LCSArgumentError",
    "LCSError", "LIBPATH", "NoForcing", "ParticleConfig", "RestartInit", "SpectrumInit", "TimeStep", "autotune", "boundary", "boxmean",
    "build", "config", "couple", "fill_ghost", "flow_stat", "flow_timestep", "forcing", "from_dict", "get_tiles", "lcsio", "load", "localize",
    "makeindex", "motion", "particle_density", "particle_stat", "particle_timestep", "particle_update", "read_log", "readconfig", "rhs", "rhs_update",
    "Server", "set_state", "set_tiles", "simulate", "stat", "step", "step_config", "to_dict", "update",
]
