"""Host logic of the device material assembly (no GPU): axis masks as
index ranges, the choice of assembly, and its refusals."""
import numpy as np
import pytest

from rfx_amd import geometry
from rfx_amd.assembly import mask_range
from tests import cases


def test_mask_range():
    assert mask_range(np.array([0, 0, 1, 1, 1, 0], bool)) == (2, 5)
    assert mask_range(np.zeros(4, bool)) == (0, 0)
    assert mask_range(np.ones(3, bool)) == (0, 3)
    with pytest.raises(ValueError):
        mask_range(np.array([1, 0, 1], bool))


def test_box_masks_are_ranges():
    """Every axis mask the Box rule produces (interval and thin-sheet
    branches, uniform and graded coordinates) is one index range."""
    rng = np.random.default_rng(3)
    coords = [np.arange(40, dtype=np.float32) * np.float32(1e-3),
              np.cumsum(rng.uniform(0.1e-3, 1e-3, 40)).astype(np.float32)]
    for c in coords:
        for _ in range(200):
            lo, hi = sorted(rng.uniform(-0.005, 0.045, 2))
            if rng.random() < 0.2:
                hi = lo + float(rng.uniform(0, 2e-4))      # thin sheet
            m = geometry.axis_mask(c, lo, hi)
            a, b = mask_range(m)
            assert m.sum() == b - a


def test_assembly_choice_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("checks the CPU-only behaviour")
    from rfx_amd._lib import RfxHipError
    sim = cases.to_sim(cases.mixed_case("run"))
    sim.material_assembly = "auto"
    grid, prob, _ = sim._build_problem(8, "run")      # host assembly
    assert prob.mat.ids.shape == grid.shape
    sim.material_assembly = "device"
    with pytest.raises(RfxHipError):
        sim._build_problem(8, "run")
    sim.material_assembly = "sideways"
    with pytest.raises(ValueError):
        sim._build_problem(8, "run")


def test_device_assembly_refuses_host_edits():
    """Material sweeps, dispersion and overrides edit ids on the host."""
    spec = cases.mixed_case("forward")
    sim = cases.to_sim(spec)
    sim.material_assembly = "device"
    with pytest.raises(NotImplementedError):
        sim._build_problem(8, "forward", material_sweep=(None, "sigma", 0.1))
