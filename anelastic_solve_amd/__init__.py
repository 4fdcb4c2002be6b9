"""MI355X-native pressure-projection Poisson solver (gfx950 HIP + C ABI).

Drop-in for the reference's multigrid pressure solver ``mgMK`` and the
projection step around it; see ``include/amg.h`` for the C ABI and
``anelastic_solve_amd.mg`` for the reference-named Python interface.
"""
from . import _capi
from .grid import (Axis, Grid, axis_from_faces, box_grid, load_grid_dir, load_grid_file, tanh_axis,
                   uniform_axis, write_grid_file)
from .flowbc import FlowBC
from .mg import (AMGError, BatchSolver, PressureSolver, Server, Pgrad, Psource, default_config,
                 divergence, ghost, mg_end, mg_init, mg_read_params, mg_solver, var_advance_CN_PL)

__all__ = [
    "Axis", "Grid", "axis_from_faces", "box_grid", "load_grid_dir", "load_grid_file", "tanh_axis",
    "uniform_axis", "write_grid_file",
    "AMGError", "BatchSolver", "PressureSolver", "Server", "Pgrad", "Psource", "default_config",
    "divergence", "ghost", "mg_end", "mg_init", "mg_read_params", "mg_solver", "var_advance_CN_PL",
    "FlowBC", "load_library",
]


def load_library():
    """Load ``libanelastic_mg.so`` (raises ImportError if it was not built)."""
    return _capi.load()
