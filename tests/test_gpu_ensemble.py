"""Batched ensembles: members of one size share every pass launch of the fused
substep, and the ensemble step is replayed from a captured HIP graph.  Each
member must follow the CPU oracle's HD step exactly as a lone solver does
(relative L2 1e-12); graph replay, eager launches and autotuned tile widths
must give identical results."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def to_gpu(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def to_np(t):
    return t.detach().cpu().numpy()


# all four member sizes of the 96-member workload, 96^3 twice so that the
# batched 96-point kernels see a member index above zero
SIZES = [32, 48, 32, 64, 96, 48, 32, 64, 96]
NUS = [4e-3, 3e-3, 6e-3, 2e-3, 1.5e-3, 5e-3, 8e-3, 3e-3, 2.5e-3]


def members():
    from oracle import oracle as O
    rng = np.random.default_rng(7)
    out = []
    for m, n in enumerate(SIZES):
        og = O.Grid(n)
        k0 = rng.uniform(2.0, 4.0)
        v, _ = og.initialv_k4exp(k0, 1.0, 100 + m)
        out.append((og, v))
    return out


def make(ens_cls, init, graph=True):
    e = ens_cls(SIZES, NUS, dt=2e-3, ord=2)
    e.set_graph(graph)
    for m, (_, v) in enumerate(init):
        e.set_velocity(m, [to_gpu(x) for x in v])
    return e


def test_ensemble_members_match_oracle(gpu):
    from ghost_amd.solvers import Ensemble
    init = members()
    e = make(Ensemble, init)
    assert sorted(e.groups()) == [(32, 3), (48, 2), (64, 2), (96, 2)]
    e.step(6)
    for m, (og, v0) in enumerate(init):
        ref = [x.copy() for x in v0]
        for _ in range(6):
            og.hd_step(ref, 2e-3, NUS[m], ord=2)
        got = e.velocity(m)
        for k in range(3):
            assert rel(to_np(got[k]), ref[k]) < 1e-12, (m, k)


@pytest.mark.parametrize("branch", ["1", "0"])
def test_ensemble_graph_replay_equals_eager(gpu, branch, monkeypatch):
    """Several replays with the program's own work in between, each checked
    against eager launches of the same step; the graph captures the size
    groups as parallel branches (default) or one after another."""
    import torch
    from ghost_amd.solvers import Ensemble
    monkeypatch.setenv("GHX_ENS_BRANCH", branch)
    init = members()
    eg = make(Ensemble, init, graph=True)
    ee = make(Ensemble, init, graph=False)
    for rep in range(4):
        eg.step(1)
        ee.step(1)
        junk = torch.randn(1 << 20, device="cuda").sum()  # unrelated eager work between replays
        torch.cuda.synchronize()
        del junk
        for m in (0, 1, 3, 4, 8):
            a, b = eg.velocity(m), ee.velocity(m)
            for k in range(3):
                assert torch.equal(a[k], b[k]), (rep, m, k)


def test_ensemble_autotune_keeps_states_and_results(gpu):
    """Tile widths do not change the arithmetic (bit-identical steps); the
    radix schedules the autotuner may also pick do, so a group on another
    schedule is held to the oracle instead."""
    import torch
    from ghost_amd.solvers import Ensemble
    init = members()
    ea = make(Ensemble, init)
    tiles = ea.autotune(steps=2)
    assert len(tiles) == 4 and all(tz > 0 and ty > 0 for tz, ty in tiles)
    sched = ea.schedules()
    assert [(tz, ty) for _, tz, ty in sched] == tiles
    by_size = {n: sc for (n, _), (sc, _, _) in zip(ea.groups(), sched)}
    assert by_size[48] == 0 and by_size[96] == 0  # one schedule at these sizes
    eb = make(Ensemble, init)
    ea.step(3)
    eb.step(3)
    for m, (og, v0) in enumerate(init):
        a, b = ea.velocity(m), eb.velocity(m)
        if by_size[SIZES[m]] == 0:
            for k in range(3):
                assert torch.equal(a[k], b[k]), (m, k)
        else:
            ref = [x.copy() for x in v0]
            for _ in range(3):
                og.hd_step(ref, 2e-3, NUS[m], ord=2)
            for k in range(3):
                assert rel(to_np(a[k]), ref[k]) < 1e-12, (m, k)


@pytest.mark.parametrize("sched", [1, 2, 3])
def test_ensemble_radix_schedules_match_oracle(gpu, sched):
    """The z passes (bit 0) and the y passes (bit 1) on four points per lane
    at 32^3 and 64^3 (radix 2, 4, 4 and 4, 4, 4 instead of 4, 8 and 8, 8)."""
    from ghost_amd.solvers import Ensemble
    init = members()
    e = make(Ensemble, init)
    for gi, (n, _) in enumerate(e.groups()):
        e.set_schedule(gi, sched if n in (32, 64) else 0)
    assert {sc for (n, _), (sc, _, _) in zip(e.groups(), e.schedules()) if n in (32, 64)} == {sched}
    e.step(4)
    for m, (og, v0) in enumerate(init):
        if SIZES[m] not in (32, 64):
            continue
        ref = [x.copy() for x in v0]
        for _ in range(4):
            og.hd_step(ref, 2e-3, NUS[m], ord=2)
        got = e.velocity(m)
        for k in range(3):
            assert rel(to_np(got[k]), ref[k]) < 1e-12, (m, k)
