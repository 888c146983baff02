from .fields import (LaserBeam, make_dc_mot_beams, make_rf_mot_beams, make_hallas_beams,
                     hallas_saturations, beam_table, BeamTable)


def __getattr__(name):
    # the solver imports the device plan, which imports .fields: load lazily
    if name in ("solve_rate_equations", "solve_rate_equations_auto", "solve_rate_equations_jit"):
        from . import rate_equations
        return getattr(rate_equations, name)
    raise AttributeError(name)
