"""Batched ensembles (trisk_ensemble_* through the C ABI).

Every member of an ensemble must evolve bitwise as a single model of its level on the
cell-pair path (the kernels share the per-entity arithmetic).  That is checked for
several member counts, including odd ones, under every schedule the autotuner can
choose, and with graph replay.  The FP64 members are also checked against the oracle
to 1e-12 per field.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from dmplex_amd import (HIPBackend, ModelTables, TRiSKEnsemble, TRiSKShallowWater, create_hexagon_mesh, diagnostics,
                        rk4_steps_, tendency_)
from dmplex_amd import _device as D
from dmplex_amd.synthetic import earth_parameters, seeded_state
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

_MESHES = {}


def tables(n):
    if n not in _MESHES:
        dm = create_hexagon_mesh(n, centroidal=True)
        _MESHES[n] = ModelTables(dm, centers=dm.generators, **earth_parameters())
    return _MESHES[n]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def member_state(t, level_index, member):
    return seeded_state(t.grid, seed_coef=1000 + 16 * level_index + member, seed_noise=2000 + 16 * level_index + member)


def make_single(n, precision):
    return TRiSKShallowWater(tables(n).grid, path="pair", backend=HIPBackend(0, precision), **earth_parameters())


LEVELS = (2, 3, 4)
MEMBERS = (3, 1, 5)
DTS = (2400.0, 1200.0, 600.0)


def build(precision):
    ens = TRiSKEnsemble([tables(n) for n in LEVELS], MEMBERS, backend=HIPBackend(0, precision))
    y = ens.zeros()
    for li, (n, M) in enumerate(zip(LEVELS, MEMBERS)):
        for m in range(M):
            ens.set_member(y, li, m, member_state(tables(n), li, m))
    return ens, y


def single_results(precision, nsteps):
    out = {}
    for li, (n, M) in enumerate(zip(LEVELS, MEMBERS)):
        s = make_single(n, precision)
        for m in range(M):
            yd = s.upload_state(member_state(s, li, m))
            rk4_steps_(yd, DTS[li], nsteps, s)
            out[li, m] = s.download_state(yd)
    return out


@pytest.mark.parametrize("precision", ["fp64", "fp32_bulk"])
def test_members_equal_single_models(precision):
    say(f"ensemble {precision}: levels {LEVELS} x members {MEMBERS}, 6 steps vs single models")
    ens, y = build(precision)
    for li, (n, M) in enumerate(zip(LEVELS, MEMBERS)):
        for m in range(M):
            assert np.array_equal(ens.get_member(y, li, m), member_state(tables(n), li, m))
    ens.rk4_steps_(y, DTS, 6)
    ref = single_results(precision, 6)
    for (li, m), r in ref.items():
        got = ens.get_member(y, li, m)
        assert np.array_equal(got, r), (li, m, rel(got, r))


def test_tendency_equals_single_models():
    say("ensemble fp32_bulk tendency vs single models")
    ens, y = build("fp32_bulk")
    dy = ens.zeros()
    ens.tendency_(dy, y)
    for li, (n, M) in enumerate(zip(LEVELS, MEMBERS)):
        s = make_single(n, "fp32_bulk")
        for m in range(M):
            yd = s.upload_state(member_state(s, li, m))
            ref = s.zeros()
            tendency_(ref, yd, s)
            got = ens.member_device(dy, li, m)
            assert bool((got == ref).all()), (li, m)


EVEN = (4, 8, 4)


@pytest.mark.parametrize("precision", ["fp64", "fp32_bulk"])
@pytest.mark.parametrize("lanes", [1, 2, 4, 8])
@pytest.mark.parametrize("block_size", [128, 512])
@pytest.mark.parametrize("blocks_per_cu", [0, 4])
@pytest.mark.parametrize("split", ["chunked", "blocked", "interleaved"])
def test_every_schedule_gives_the_same_bits(precision, lanes, block_size, blocks_per_cu, split):
    """Members per lane, block size, persistence and XCD split change only the mapping of
    work to lanes: the states must be bitwise those of the default schedule and of a
    single model."""
    say(f"ensemble schedule {precision} lanes={lanes} {block_size} {blocks_per_cu} {split}")
    members = EVEN if lanes < 8 else (8, 8, 16)
    ens = TRiSKEnsemble([tables(n) for n in LEVELS], members, backend=HIPBackend(0, precision))
    y = ens.zeros()
    for li, (n, M) in enumerate(zip(LEVELS, members)):
        for m in range(M):
            ens.set_member(y, li, m, member_state(tables(n), li, m))
    ref = y.clone()
    ens.rk4_steps_(ref, DTS, 3)
    ens.set_schedule([dict(block_size=block_size, blocks_per_cu=blocks_per_cu, split=split,
                           members_per_lane=lanes)] * 3)
    ens.rk4_steps_(y, DTS, 3)
    assert bool((y == ref).all())
    s = make_single(3, precision)
    yd = s.upload_state(member_state(s, 1, 7))
    rk4_steps_(yd, DTS[1], 3, s)
    assert np.array_equal(ens.get_member(y, 1, 7), s.download_state(yd))


@pytest.mark.parametrize("precision", ["fp64", "fp32_bulk"])
def test_fused_stage_gives_the_same_bits(precision):
    """One cooperative launch per RK stage (the state, cell and edge phases with grid-wide
    barriers between them) gives the three-launch states bitwise: RK4 steps in calls of
    several lengths (each call's first stage also forms G) and a tendency."""
    say(f"ensemble fused stage {precision}")
    ens = TRiSKEnsemble([tables(n) for n in LEVELS], EVEN, backend=HIPBackend(0, precision))
    y = ens.zeros()
    for li, (n, M) in enumerate(zip(LEVELS, EVEN)):
        for m in range(M):
            ens.set_member(y, li, m, member_state(tables(n), li, m))
    ref = y.clone()
    ens.rk4_steps_(ref, DTS, 5)
    dref = ens.zeros()
    ens.tendency_(dref, y)
    ens.use_fused(True)
    for split in ((5,), (2, 3)):
        got = y.clone()
        for n in split:
            ens.rk4_steps_(got, DTS, n)
        assert bool((got == ref).all()), split
    d = ens.zeros()
    ens.tendency_(d, y)
    assert bool((d == dref).all())
    ens.set_timing(True)
    got = y.clone()
    ens.rk4_steps_(got, DTS, 5)
    ens.set_timing(False)
    assert bool((got == ref).all())
    stats = ens.kernel_stats()
    assert [k["name"] for k in stats] == ["ens_stage_kernel"] and stats[0]["launches"] == 20
    ens.use_fused(False)
    got = y.clone()
    ens.rk4_steps_(got, DTS, 5)
    assert bool((got == ref).all())
    with pytest.raises(ValueError, match="divisible"):
        TRiSKEnsemble([tables(n) for n in LEVELS], MEMBERS, backend=HIPBackend(0, precision)).use_fused(True)


def test_fp64_members_match_oracle():
    say("ensemble fp64 member vs oracle, level 4, 12 steps")
    ens = TRiSKEnsemble([tables(4)], [4], backend=HIPBackend(0, "fp64"))
    y = ens.zeros()
    states = [member_state(tables(4), 0, m) for m in range(4)]
    for m, s in enumerate(states):
        ens.set_member(y, 0, m, s)
    ens.rk4_steps_(y, [600.0], 12)
    o = Oracle(tables(4))
    ne = tables(4).grid.num_edges()
    for m in (0, 3):
        ref = o.rk4_steps(states[m].copy(), 600.0, 12)
        got = ens.get_member(y, 0, m)
        for sl in (slice(0, ne), slice(ne, None)):
            assert rel(got[sl], ref[sl]) < 1e-12


def test_graph_replay_and_autotune():
    import torch

    say("ensemble graph replay and autotune")
    ens, y = build("fp32_bulk")
    direct = y.clone()
    ens.rk4_steps_(direct, DTS, 19)
    before = y.clone()
    best, ms = ens.autotune(y, DTS, steps=2)
    assert bool((y == before).all()), "autotune must not change the state"
    assert len(best) == 3 and all(b["block_size"] in (128, 256, 512) for b in best) and all(t > 0 for t in ms)
    assert ens.schedule() == best
    ens.use_graphs(True)
    for _ in range(2):
        torch.cuda.synchronize()
        y2 = before.clone()
        ens.rk4_steps_(y2, DTS, 19)  # two graph replays of 8 steps and 3 direct steps
        torch.cuda.synchronize()
        assert bool((y2 == direct).all())


def test_member_diagnostics_equal_single_models():
    say("ensemble diagnostics vs single models")
    ens, y = build("fp64")
    d = ens.diagnostics(y)
    assert d.shape == (sum(MEMBERS), 3)
    k = 0
    for li, (n, M) in enumerate(zip(LEVELS, MEMBERS)):
        s = make_single(n, "fp64")
        for m in range(M):
            ref = diagnostics(s, s.upload_state(member_state(s, li, m)))
            assert tuple(d[k]) == ref
            k += 1


def test_ensemble_rejects_bad_arguments():
    say("ensemble error paths")
    with pytest.raises(ValueError, match="level 0"):
        dm = create_hexagon_mesh(2, centroidal=True)
        TRiSKEnsemble([ModelTables(dm, centers=dm.generators, pv_flux="gassmann", **earth_parameters())], [2])
    ens = TRiSKEnsemble([tables(2)], [2])
    y = ens.zeros()
    with pytest.raises(ValueError, match="member out of range"):
        ens.member_device(y, 0, 2)
    with pytest.raises(ValueError, match="schedule"):
        ens.set_schedule([dict(block_size=96, blocks_per_cu=0, split="chunked")] * 3)
    with pytest.raises(ValueError, match="members_per_lane"):
        TRiSKEnsemble([tables(2)], [3]).set_schedule([dict(block_size=256, blocks_per_cu=0, split="chunked",
                                                          members_per_lane=2)] * 3)
    ens.set_schedule([dict(block_size=256, blocks_per_cu=0, split="chunked", members_per_lane=2)] * 3)
    dt = np.array([600.0])
    with pytest.raises(ValueError, match="aligned"):  # refused on the host, before any launch
        D.check(D.lib().trisk_ensemble_rk4_steps(ens.handle, C.c_void_p(y.data_ptr() + 8),
                                                 dt.ctypes.data_as(C.POINTER(C.c_double)), 1, None))
    with pytest.raises(ValueError):
        ens.rk4_steps_(y, [1.0, 2.0], 1)
