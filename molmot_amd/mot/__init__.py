from .simulator import DCMOTSimulator, RFMOTSimulator
from .force_scan import (force_vs_z, force_vs_v, force_map_2d, spring_constant,
                         damping_coefficient, capture_velocity, optimize_parameters)
from .analysis import survived, capture_fraction, capture_fraction_batch, capture_fraction_ensemble  # noqa: E402,F401
