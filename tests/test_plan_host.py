"""Host-side plan helpers (no GPU)."""
import dataclasses

import numpy as np
import pytest

from molmot_amd import configs
from molmot_amd.hip.plan import component_split_signs
from molmot_amd.molecules import load_sroh_from_julia
from molmot_amd.obe.fields import beam_table


def test_component_split_signs_red_mot():
    mol = load_sroh_from_julia()
    bt = beam_table(configs.red_mot_beams(mol))
    assert np.array_equal(component_split_signs(bt), [1.0, -1.0, 1.0, -1.0])


def test_component_split_signs_rejects_mixed_component():
    mol = load_sroh_from_julia()
    beams = configs.red_mot_beams(mol)
    # components 0 (+) and 1 (-) merged: the split gradient is undefined per component
    merged = [dataclasses.replace(b, comp=int(b.comp) // 2) for b in beams]
    with pytest.raises(ValueError, match="mixes split signs"):
        component_split_signs(beam_table(merged))
    # merging components of equal sign is fine
    ok = [dataclasses.replace(b, comp=int(b.comp) % 2) for b in beams]
    assert np.array_equal(component_split_signs(beam_table(ok)), [1.0, -1.0])
