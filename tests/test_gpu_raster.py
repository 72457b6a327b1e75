"""GPU parity of the compositing kernels against the FP64 oracle, through the
drop-in Rasterizer (forward and autograd backward)."""

import numpy as np
import pytest

from tests import harness

pytestmark = pytest.mark.gpu

ABS_TOL = 1e-4      # RGB, opacity, depth, normals: max abs error vs the FP64 oracle
GRAD_REL_TOL = 1e-3  # gradients: ||gpu - oracle|| / ||oracle|| per tensor, over the
#                      sites the oracle does not flag as ill-conditioned (rays
#                      tangent to the sphere or intervals at the rounding floor,
#                      where one-ulp differences flip whole contributions);
#                      at most MAX_FRAGILE of the sites may be flagged
MAX_FRAGILE = 0.02


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available()
    return torch.device("cuda:0")


CASES = [
    dict(seed=3, n=400, width=48, height=40),
    dict(seed=5, n=1500, width=97, height=61, fov=70.0),   # ragged: out-of-bounds lanes
    dict(seed=7, n=800, width=64, height=48, height_std=0.05, T=5),
    dict(seed=9, n=600, width=40, height=33, quantiles=[0.9, 0.5, 0.1]),
    # ray-map cameras (equidistant fisheye), binned through the cone tree
    dict(seed=11, n=1500, width=56, height=56, eye=(0.0, 0.0, 1.3), fisheye_fov=150.0),
    dict(seed=12, n=1500, width=64, height=64, eye=(0.2, 0.1, 0.6), fisheye_fov=180.0, height_std=0.05),
]


@pytest.mark.parametrize("cfg", CASES, ids=[f"case{i}" for i in range(len(CASES))])
def test_forward_matches_fp64_oracle(dev, cfg):
    case = harness.small_case(**cfg)
    print(f"forward case {cfg}", flush=True)
    out, _ = harness.run_gpu(case, dev)
    ref, _ = harness.run_oracle(case)
    for k in ("color", "opacity", "normal_distance", "normal"):
        err = np.abs(out[k] - ref[k]).max()
        assert err < ABS_TOL, (k, err)
    if case.depth_quantiles is not None:
        # unfilled entries stay -1 in both
        assert np.array_equal(out["quantile_depths"] == -1, ref["quantile_depths"] == -1)
        err = np.abs(out["quantile_depths"] - ref["quantile_depths"]).max()
        assert err < ABS_TOL, err
    for k in ("contrib", "point_err"):
        err = np.abs(out[k] - ref[k]).max()
        assert err < ABS_TOL * max(1.0, np.abs(ref[k]).max()), (k, err)
    # invariant: sum of per-site weights = opacity summed over pixels / (H W)
    assert abs(out["contrib"].sum() - out["opacity"].mean()) < 1e-4


@pytest.mark.parametrize("cfg", CASES, ids=[f"case{i}" for i in range(len(CASES))])
def test_backward_matches_fp64_oracle(dev, cfg):
    case = harness.small_case(**cfg)
    print(f"backward case {cfg}", flush=True)
    gin = harness.random_grads(case, seed=cfg["seed"])
    _, grads = harness.run_gpu(case, dev, gin)
    _, ref = harness.run_oracle(case, gin)
    keep = harness.well_conditioned(ref)
    for k in harness.GRAD_KEYS:
        err = harness.rel_err(grads[k][keep], ref[k][keep])
        assert err < GRAD_REL_TOL, (k, err)


@pytest.mark.parametrize("cfg,attr", [(CASES[0], "float"), (CASES[1], "float"), (CASES[2], "half"),
                                      (CASES[4], "float")],
                         ids=["case0", "case1", "case2-half", "case4-fisheye"])
def test_benchmark_matches_fp64_oracle(dev, cfg, attr):
    """Rasterizer.benchmark (colour only, per-ray stop, fp16 power faces)."""
    case = harness.small_case(**cfg)
    print(f"benchmark case {cfg} {attr}", flush=True)
    col, diff = harness.run_gpu_benchmark(case, dev, attr_dtype=attr)
    ref = harness.run_oracle_benchmark(case, diff, attr_dtype=attr)
    err = np.abs(col - ref).max()
    assert err < ABS_TOL, err
    assert ref.max() > 0.1
