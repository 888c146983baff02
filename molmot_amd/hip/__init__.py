"""MI355X batched API (counterpart of the reference's molmot_cuda package,
molmot_cuda/__init__.py:72-119).  Every function runs the gfx950 kernels of
libmolmot_hip.so; there is no CPU fallback."""
from ._lib import MolmotHIPError, available as library_available


def has_hip() -> bool:
    """True if the kernel library is built and a HIP device is visible
    (the role of has_cuda(), molmot_cuda/__init__.py:44-59)."""
    import torch
    return library_available() and torch.cuda.is_available()


has_cuda = has_hip

from .plan import HIPMolData, CUDAMolData  # noqa: E402
from .rate_equations import (  # noqa: E402
    solve_rate_equations_batch, rate_eq_map_device, rate_eq_points_device,
    force_vs_z_hip, force_vs_v_hip, force_map_2d_hip, spring_constant_hip,
    damping_coefficient_hip, optimize_parameters_hip,
    force_vs_z_cuda, force_vs_v_cuda, force_map_2d_cuda, spring_constant_cuda,
    damping_coefficient_cuda, optimize_parameters_cuda)
from .trajectories import (  # noqa: E402
    traj_euler_device, simulate_trajectories_hip, simulate_trajectory_hip,
    simulate_trajectories_cuda, simulate_trajectory_cuda, recoil_kick_scale, traj_euler_rf_device,
    traj_simulator_device, capture_velocity_hip, capture_velocity_cuda, traj_euler_schedule)
