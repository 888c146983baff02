"""Launch-shape autotuner on the GPU: the chosen shape is one of the
candidates, results do not depend on it, and the JSON cache is reused."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda", 0)


def test_lindblad_lanes_tuned_and_cached(mol, dev, tmp_path):
    import torch
    from molmot_amd.autotune import Tuner, tune_lindblad_lanes
    from molmot_amd.obe.lindblad import LindbladClass, lindblad_rk4_device
    from molmot_amd.molecules import synthetic_molecule
    m = synthetic_molecule(36, 4, seed=11)
    Ee, G = np.mean(m.energies[36:]), m.Gamma
    pairs = [{"freq": Ee - np.mean(m.energies[:12]) - G / (2 * np.pi), "q_fwd": 2, "q_bwd": 0, "s0": 1.0},
             {"freq": Ee - np.mean(m.energies[12:36]) - G / (2 * np.pi), "q_fwd": 0, "q_bwd": 2, "s0": 1.0}]
    rng = np.random.default_rng(3)
    pol = rng.normal(size=(2, 2, 3)) + 1j * rng.normal(size=(2, 2, 3))
    pol /= np.linalg.norm(pol, axis=-1, keepdims=True)
    cls = LindbladClass(m, pairs, 1e-3, 0.2, 1.0, 1.0, pol)
    kv = torch.tensor(np.linspace(-2, 2, 256) * cls.kfac, device=dev)
    ph = torch.tensor(np.linspace(0, 6, 256), device=dev)
    F0, _, _ = lindblad_rk4_device(cls, kv, ph, 0.05, 20, 10, "fp32")
    path = str(tmp_path / "tune.json")
    best, ms = tune_lindblad_lanes(cls, "fp32", kv, ph, tuner=Tuner(path))
    print("lanes", best, ms, flush=True)
    assert best in (64, 128, 256) and set(ms) == {64, 128, 256}
    assert cls.lanes("fp32") == best
    F1, _, _ = lindblad_rk4_device(cls, kv, ph, 0.05, 20, 10, "fp32")
    assert torch.allclose(F1, F0, rtol=1e-6, atol=1e-6 * float(F0.abs().max()))
    table = json.load(open(path))
    assert any(k.endswith("lindblad_valu|36+4|fp32|n=256") and v["best"] == best for k, v in table.items())
    # a fresh tuner reads the cache: no timing, same choice applied
    cls.set_lanes("fp32", 0)
    best2, _ = tune_lindblad_lanes(cls, "fp32", kv, ph, tuner=Tuner(path))
    assert best2 == best and cls.lanes("fp32") == best


def test_rate_block_tuned_results_identical(mol, dev):
    import torch
    from molmot_amd.autotune import Tuner, tune_rate_block
    from molmot_amd.hip.plan import HIPMolData
    from molmot_amd.hip.rate_equations import rate_eq_map_device
    from molmot_amd.obe.fields import make_dc_mot_beams
    plan = HIPMolData(mol, make_dc_mot_beams(mol, -0.2, 0.7, 1.0))
    v = torch.linspace(-3, 3, 96, dtype=torch.float64, device=dev)
    z = torch.linspace(-4e-3, 4e-3, 80, dtype=torch.float64, device=dev)
    outs = {}
    for b in (64, 128, 256):
        plan.set_block(b)
        assert plan.block == b
        outs[b] = [x.clone() if x is not None else None for x in rate_eq_map_device(plan, v, z, 0.16)]
    for b in (128, 256):
        for x, y in zip(outs[b], outs[64]):
            if x is not None:
                assert torch.equal(x, y)
    best, ms = tune_rate_block(plan, v, z, 0.16, tuner=Tuner(None))
    assert best in (64, 128, 256) and plan.block == best
    plan.set_block(0)
    assert plan.block == 256
