"""MI355X-native Cartesian finite-volume step for TORCH.

Host-side mirror of the reference's interface for this path (BFG508/TORCH,
python/cfdSolver): ``Simulation`` (torchSession), ``Execution`` and its
device drop-in ``DeviceExecution`` (torchExecution), ``hllc_flux_grid`` /
``hllc_flux`` (torchFlux), ``muscl_hancock_step_3d`` / ``reaction_step_3d``
(torchSolver3D) and ``muscl_hancock_step_2d`` / ``reaction_step_2d``
(torchSolver2D), the legacy drivers ``run_simulation`` /
``run_simulation_wqr`` (torchSolver, with time histories),
``run_torch_solver_2d`` / ``run_torch_solver_3d``, and ``read_checkpoint`` / ``write_checkpoint`` (torchIO,
schema-2 layout). The arithmetic runs in libtorchhip.so (HIP kernels for
gfx950, built by ``python -m torch_amd.build``) behind the C ABI in
include/torch_hip.h; there is no CPU fallback.
"""

from .torchExecution import DeviceExecution, Execution, validate_workers  # noqa: F401
from .torchFlux import hllc_flux, hllc_flux_grid, rusanov_flux, rusanov_flux_grid  # noqa: F401
from .torchIO import read_checkpoint, write_checkpoint  # noqa: F401
from .torchAdaptive import AdaptiveSimulation  # noqa: F401
from .torchSession import Simulation  # noqa: F401
from .torchSolver import run_simulation, run_simulation_wqr, run_simulation_wQR  # noqa: F401
from .torchSolver2D import muscl_hancock_step_2d, reaction_step_2d, run_torch_solver_2d  # noqa: F401
from .torchSolver3D import (  # noqa: F401
    LIMITERS,
    muscl_hancock_step_3d,
    reaction_step_3d,
    run_torch_solver_3d,
    uconsv_to_uprim_3d,
    uprim_to_uconsv_3d,
)
