"""Batched ensemble members.

A batch of several cubes must advance every member exactly as a run of that
member alone would: every kernel treats tiles independently and the halo
update applies its six-tile map per member.  The GPU test checks this bit
for bit for AcousticDynamics; the CPU test checks the per-member halo map
semantics on the host restatement.
"""
import numpy as np
import pytest
import torch

from fv3core_amd._config import DynamicalCoreConfig
from fv3core_amd.halo import CENTER, apply_map_numpy, build_map
from fv3core_amd.workloads import cube_state

N, NZ = 12, 8
FIELDS = ("delp", "pt", "u", "v", "w", "delz", "uc", "vc", "ua", "va", "omga", "pe", "pk",
          "peln", "pkz", "zh", "mfxd", "mfyd", "cxd", "cyd", "q_con")


def test_member_offset_map_equals_separate_updates():
    """Applying the six-tile map to each member of a stacked batch (what the
    halo kernel does with its member grid axis) equals separate updates."""
    n, h = 6, 3
    mp = build_map(n, (CENTER,), 3, h)
    rng = np.random.default_rng(0)
    members = 3
    f = rng.normal(size=(6 * members, n + 7, n + 7, 2))
    stacked = f.copy()
    for m in range(members):
        off = mp.copy()
        off[:, 0] += 6 * m
        off[:, 4] += 6 * m
        stacked = apply_map_numpy([stacked], off)[0]
    for m in range(members):
        alone = apply_map_numpy([f[6 * m:6 * m + 6]], mp)[0]
        np.testing.assert_array_equal(stacked[6 * m:6 * m + 6], alone)


@pytest.mark.gpu
@pytest.mark.parametrize("use_graph", [False, True])
def test_gpu_batched_members_equal_single_runs(use_graph):
    from fv3core_amd.ensemble import ensemble_acoustic, member_slice

    members = 3
    init = cube_state(N, NZ, seed=5)
    cfg = DynamicalCoreConfig(npz=NZ, n_split=2, k_split=1, nord=3, dt_atmos=225)
    acoustic, state = ensemble_acoustic(init, members, cfg, seed=10, use_graph=use_graph)
    singles = [member_slice(state, m) for m in range(members)]
    calls = 3 if use_graph else 1  # eager, capture, replay
    for _ in range(calls):
        acoustic(state)
    for m in range(members):
        one, _ = ensemble_acoustic(init, 1, cfg, seed=0)
        for _ in range(calls):
            one(singles[m])
        torch.cuda.synchronize()
        for k in FIELDS:
            a = getattr(state, k)[6 * m:6 * m + 6]
            b = getattr(singles[m], k)
            assert torch.equal(a, b), f"member {m} field {k} differs from its single run"
    # members were perturbed differently, so they must differ from each other
    assert not torch.equal(state.pt[0:6], state.pt[6:12])
