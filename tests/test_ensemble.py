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


def _perturbed_init(init, members, seed, pt_sd=0.5, uv_sd=1.0):
    """The initial state of each member as ensemble_acoustic perturbs it
    (pt * (1 + N(0, pt_sd) / 288), u, v + N(0, uv_sd), torch generator seeded
    with seed + m), with halos refreshed by the host halo map."""
    from fv3core_amd.halo import X_FACE, Y_FACE
    out = []
    st0 = init["state"]
    n = st0["delp"].shape[1] - 7
    for m in range(members):
        g = torch.Generator(device="cpu").manual_seed(seed + m)
        shp = st0["pt"].shape
        dpt = torch.randn(shp, generator=g, dtype=torch.float64).numpy()
        du = torch.randn(shp, generator=g, dtype=torch.float64).numpy()
        dv = torch.randn(shp, generator=g, dtype=torch.float64).numpy()
        st = {k: v.copy() for k, v in st0.items()}
        st["pt"] = st["pt"] * (1.0 + pt_sd * dpt / 288.0)
        st["u"] = st["u"] + uv_sd * du
        st["v"] = st["v"] + uv_sd * dv
        st["pt"] = apply_map_numpy([st["pt"]], build_map(n, (CENTER,), 3, 3))[0]
        st["u"], st["v"] = apply_map_numpy([st["u"], st["v"]],
                                           build_map(n, (Y_FACE, X_FACE), 3, 3))
        one = dict(init)
        one["state"] = st
        out.append(one)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("mixed", [False, True])
def test_gpu_uneven_ensemble_members_match_oracle(mixed):
    """An uneven ensemble (two perturbed C12 members in one batch and one C16
    member) advanced by UnevenEnsemble, each member against the FP64 oracle's
    AcousticDynamics call on that member alone: FP64 at 1e-9 (w and omega,
    the implicit solve's small residuals, at 1e-7); the mixed form within the
    1e-5 gate in ps, pt and delp."""
    from fv3core_amd.ensemble import UnevenEnsemble, ensemble_acoustic
    from fv3core_amd.stencils.fv_dynamics import init_pfull
    from tests.dyn_helpers import oracle_acoustic_call
    from tests.helpers import assert_close

    nz = NZ
    cfg = DynamicalCoreConfig(npz=nz, n_split=2, k_split=1, nord=3, dt_atmos=225)
    plan = [(12, 2, 0.5, 1.0, 120), (16, 1, 0.0, 0.0, 160)]
    parts, refs = [], []
    for n, members, pt_sd, uv_sd, seed in plan:
        init = cube_state(n, nz, seed=4)
        parts.append(ensemble_acoustic(init, members, cfg, seed=seed, pt_sd=pt_sd, uv_sd=uv_sd,
                                       mixed_precision=mixed))
        pfull = init_pfull(init["ak"], init["bk"], 1.0e5)
        for one in _perturbed_init(init, members, seed, pt_sd, uv_sd):
            refs.append((n, len(refs), oracle_acoustic_call(one, cfg, n, nz, pfull)))
    UnevenEnsemble(parts)()
    torch.cuda.synchronize()
    member = 0
    for (acoustic, state), (n, members, _, _, _) in zip(parts, plan):
        for m in range(members):
            _, _, S = refs[member]
            member += 1
            t = slice(6 * m, 6 * m + 6)
            c = (slice(None), slice(3, n + 3), slice(3, n + 3), slice(0, nz))
            ci = (slice(None), slice(3, n + 3), slice(3, n + 3), slice(0, nz + 1))
            for name in ("delp", "pt", "w", "delz", "pkz", "pe"):
                box = ci if name == "pe" else c
                got = getattr(state, name)[t].double().cpu().numpy()[box]
                if mixed:
                    tol = {"delp": 1e-5, "pt": 1e-5, "pe": 1e-5}.get(name)
                    if tol is None:
                        continue
                else:
                    tol = 1e-7 if name == "w" else 1e-9
                # w passes through zero: its error is read against 1e-4 of its
                # largest value (round-off differences of 1e-13 m/s at points
                # where |w| ~ 1e-7 are not parity failures)
                assert_close(got, S[name][box], tol, near_zero=1e-12,
                             what=f"C{n} member {m} {name}",
                             scale_floor=1e-4 if name == "w" else 1e-6)
            ps = state.pe[t][..., nz].double().cpu().numpy()[:, 3:n + 3, 3:n + 3]
            assert_close(ps, S["pe"][:, 3:n + 3, 3:n + 3, nz], 1e-5 if mixed else 1e-12,
                         what=f"C{n} member {m} ps")


@pytest.mark.gpu
@pytest.mark.parametrize("concurrent", [False, True])
def test_gpu_uneven_ensemble_one_graph_equals_eager(concurrent):
    """UnevenEnsemble(use_graph=True): one HIP graph for the calls of every
    resolution; each replay equals the sequential eager ensemble bit for bit,
    with unrelated eager work and host synchronisations between the replays.
    With concurrent=True the second part runs on its own stream (eagerly and
    as a parallel branch of the graph) and the results are the same bits."""
    from fv3core_amd.ensemble import UnevenEnsemble, ensemble_acoustic

    cfg = DynamicalCoreConfig(npz=NZ, n_split=2, k_split=1, nord=3, dt_atmos=225)
    runs = []
    for form in ("eager", "eager_c", "graph"):
        parts = []
        for n, members in ((12, 2), (16, 1)):
            init = cube_state(n, NZ, seed=6)
            parts.append(ensemble_acoustic(init, members, cfg, seed=50 + n))
        runs.append((UnevenEnsemble(parts, use_graph=form == "graph",
                                    concurrent=concurrent and form != "eager"), parts))
    scratch = torch.zeros(1 << 16, dtype=torch.float64, device="cuda")
    for call in range(4):
        for ens, _ in runs:
            ens()
        scratch.add_(1.0)
        torch.cuda.synchronize()
        for other in (1, 2):
            for (_, sa), (_, sb) in zip(runs[0][1], runs[other][1]):
                for k in FIELDS:
                    torch.testing.assert_close(
                        getattr(sb, k), getattr(sa, k), rtol=0, atol=0, equal_nan=True,
                        msg=f"call {call}: {k} differs between form {other} and eager")
    assert runs[2][0].graph is not None

