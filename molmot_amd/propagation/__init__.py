"""Trajectory propagation (counterpart of molmot/propagation/__init__.py:8-32)."""
from .trajectories import (simulate_trajectory, simulate_trajectories, simulate_trajectories_3d,
                           monte_carlo_temperature, sample_direction, sample_maxwell_boltzmann,
                           sample_maxwell_boltzmann_speed, sample_gaussian_position,
                           sample_uniform_sphere, make_position_sampler, make_velocity_sampler)

# the reference's JIT names
simulate_trajectory_jit = simulate_trajectory
monte_carlo_temperature_jit = monte_carlo_temperature
