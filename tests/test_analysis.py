"""survived / capture_fraction (molmot/mot/analysis.py:290-355) on
hand-built trajectory lists, and the batched form against the list form."""
import numpy as np

from molmot_amd.mot.analysis import capture_fraction, capture_fraction_batch, survived


def _res(t, z):
    return (np.asarray(t, dtype=float), np.asarray(z, dtype=float), np.zeros(len(t)))


def test_survived():
    assert survived(_res([0, 1], [1e-3, -2.9e-3]))
    assert not survived(_res([0, 1], [0.0, 3e-3]))          # strict inequality
    assert not survived(_res([0, 1], [0.0, -5e-3]), r_max=4e-3)


def test_capture_fraction_rules():
    t = np.arange(5) * 1e-3                                   # 0 .. 4 ms
    res = [_res(t, [0, 0, 1e-3, 9e-3, 9e-3]),                # trapped at t_min = 2 ms
           _res(t, [0, 0, 5e-3, 0, 0]),                      # outside at the sample read
           _res(t[:2], [0, 0]),                              # ran 1 ms < t_min: not counted
           _res(t, [0, 0, -2e-3, 0, 0])]
    assert capture_fraction(res, r_max=3e-3, t_min=2e-3) == 2 / 3
    assert capture_fraction(res, r_max=3e-3, t_min=10.0) == 0.0      # nothing ran that long
    assert capture_fraction([], 3e-3, 1e-3) == 0.0
    # t_min between samples: the first sample at or after it
    assert capture_fraction(res[:2], r_max=3e-3, t_min=2.5e-3) == 0.5


def test_batch_equals_list_form():
    rng = np.random.default_rng(0)
    t = np.arange(200) * 1e-5
    Z = rng.normal(0, 3e-3, size=(50, 200))
    for t_min in (0.0, 5e-4, 1.234e-3, 1.99e-3, 5.0):
        lst = capture_fraction([_res(t, z) for z in Z], 3e-3, t_min)
        assert capture_fraction_batch(t, Z, 3e-3, t_min) == lst


def _ref_results():
    import os
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_analysis.npz"),
                allow_pickle=False)
    res = [(d["t"][i, :m], d["z"][i, :m], d["v"][i, :m]) for i, m in enumerate(d["lens"])]
    return d, res


def test_sigma_temperature_density_vs_reference():
    """sigma_vs_time / temperature_vs_time / density_vs_time against the
    reference's own outputs (analysis.py:128-288) on 24 random-walk
    trajectories of uneven length, five of them escape-filled."""
    from molmot_amd.mot import analysis as an
    d, res = _ref_results()
    nb = int(d["time_bins"])
    tc, sz, sv, na = an.sigma_vs_time(res, nb)
    assert np.array_equal(tc, d["t_centres"]) and np.array_equal(na, d["n_alive"])
    assert np.allclose(sz, d["sigma_z"], rtol=1e-13, atol=0, equal_nan=True)
    assert np.allclose(sv, d["sigma_v"], rtol=1e-13, atol=0, equal_nan=True)
    _, T = an.temperature_vs_time(res, float(d["mass"]), nb)
    assert np.allclose(T, d["T"], rtol=1e-13, atol=0, equal_nan=True)
    _, n0 = an.density_vs_time(res, int(d["N_total"]), nb)
    assert np.allclose(n0, d["n0"], rtol=1e-12, atol=0)
    assert d["n_alive"].min() < len(res)
