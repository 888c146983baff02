from .simulator import DCMOTSimulator, RFMOTSimulator
from .force_scan import (force_vs_z, force_vs_v, force_map_2d, spring_constant,
                         damping_coefficient, capture_velocity, optimize_parameters)
from .analysis import survived, capture_fraction, capture_fraction_batch, capture_fraction_ensemble  # noqa: E402,F401
from .analysis import (sigma_vs_time, temperature_vs_time, density_vs_time, ensemble_time_statistics,  # noqa: E402,F401
                       sigma_vs_time_ensemble, temperature_vs_time_ensemble, density_vs_time_ensemble)
