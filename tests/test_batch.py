"""Batched sweeps: many instances stepped by one launch per kernel type.

CPU: the seeded draws and the device kinematics table (what
ibm_compute_force receives per step: float32 arguments, cos/sin of the
float32 angle rounded once from float64), and the partition of the
instances into cache groups.  GPU: every instance of a batch equals the same
instance run on its own, bit for bit (fixed-point spreading), with one cache
group or several; graph replays equal eager steps; tuning restores the
initial state; malformed batches are rejected.
"""

import ctypes as C
import math

import numpy as np
import pytest

from conftest import load_sub


def test_draws_are_seeded_and_in_range(pkg):
    bm = load_sub("batch")
    a = bm.draw_instances(12, seed=3)
    b = bm.draw_instances(12, seed=3)
    assert [c.name for c in a] == [c.name for c in b]
    for c in a:
        assert (c.Nx, c.Ny, c.Nz) in bm.GRIDS
        assert c.chord == pytest.approx(c.Ny / 2.5)
        assert 100 <= c.Re <= 400
        assert 0.1 * c.chord - 1e-9 <= c.A_heave <= 0.5 * c.chord + 1e-9
        assert abs(math.degrees(c.A_pitch)) <= 20 + 1e-9
    assert len({c.Re for c in a}) == 12


def test_kinematics_table_matches_the_host_arguments(pkg):
    bm = load_sub("batch")
    cfg = pkg.config.naca_case(64, 32, 4, 12.0)
    t = bm.kinematics_table(cfg, 5, 4)
    assert t.dtype == np.float32 and t.shape == (4, 7)
    for j in range(4):
        xc, yc, th, vx, vy, om = cfg.motion_state(5 + j)
        a = float(np.float32(th))
        assert t[j, 0] == np.float32(xc) and t[j, 1] == np.float32(yc)
        assert t[j, 2] == np.float32(math.cos(a)) and t[j, 3] == np.float32(math.sin(a))
        assert t[j, 6] == np.float32(om)


def test_cache_groups_partition_the_instances(pkg):
    bm = load_sub("batch")
    cfgs = bm.draw_instances(40, seed=1)
    size = [19 * 4 * c.Nx * c.Ny * c.Nz for c in cfgs]
    budget = 192 << 20
    groups = bm.cache_groups(cfgs, "f32", budget)
    assert sorted(k for g in groups for k in g) == list(range(40))
    for g in groups:
        assert len(g) == 1 or sum(size[k] for k in g) <= budget
    big = [k for k, s in enumerate(size) if s > budget]
    assert all([k] in groups for k in big)
    assert bm.cache_groups(cfgs, "f32", 0) == [list(range(40))]
    assert len(bm.cache_groups(cfgs, "f16", budget)) < len(groups)


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------
def _small(pkg):
    return [pkg.config.naca_case(48, 24, 4, 9.6, Re=150, name="a"),
            pkg.config.naca_case(40, 20, 6, 8.0, Re=300, k=0.9, aoa_mean_deg=3.0, name="b"),
            pkg.config.naca_case(64, 32, 4, 12.8, Re=220, heave_frac=0.4, name="c"),
            pkg.config.naca_case(48, 24, 4, 9.6, Re=380, phase_deg=10.0, name="d")]


def _assert_same(batch_objs, seq_objs, exact_fields=False):
    """Populations and marker forces bit for bit.  A batch with the clamp
    folded into its lattice step (the default) leaves fx, fy to be rebuilt
    from the markers when read; the stand-alone spread kernel evaluates the
    delta weights in its own order, so those agree within 1e-8 absolute
    (|F| <= f_max = 0.05).  With ``exact_fields`` (the clamp kernel wrote
    them) they agree bit for bit."""
    for (bl, bi), (sl, si) in zip(batch_objs, seq_objs):
        np.testing.assert_array_equal(bl.f.to_numpy(), sl.f.to_numpy())
        np.testing.assert_array_equal(bi.FX_k.to_numpy(), si.FX_k.to_numpy())
        if exact_fields:
            np.testing.assert_array_equal(bi.fx.to_numpy(), si.fx.to_numpy())
        else:
            np.testing.assert_allclose(bi.fx.to_numpy(), si.fx.to_numpy(), rtol=0, atol=1e-8)


@pytest.mark.gpu
@pytest.mark.parametrize("block,group,lds,fused", [(256, 8, 0, True), (128, 4, 0, True),
                                                   (256, 4, 40960, True), (256, 8, 0, False)])
def test_batch_equals_each_instance_alone(pkg, gpu, block, group, lds, fused):
    """Two launches per step (the clamp folded into the lattice step) or
    three (markers, clamp into the fields, lattice): populations, marker
    forces and force boxes equal the per-instance runs bit for bit (the
    two-launch step leaves the fields and box to be rebuilt when read)."""
    bm = load_sub("batch")
    cfgs = _small(pkg)
    sweep = bm.BatchSweep(cfgs, steps=16, block=block, group=group, lds=lds, fused_clamp=fused)
    assert sweep.fused_clamp == fused
    sweep.step(7)
    seq, _ = bm.sequential(cfgs, 7)
    for (bl, bi), (sl, si) in zip(zip(sweep.lbms, sweep.ibms), seq):
        assert bi._fields_stale == fused
        np.testing.assert_array_equal(bi.force_box(), si.force_box())   # rebuilds the fields
    _assert_same(list(zip(sweep.lbms, sweep.ibms)), seq, exact_fields=not fused)


@pytest.mark.gpu
def test_batched_bodies_stay_in_their_region_tables(pkg, gpu):
    """The fused lattice step reads each batched body's force inside the
    region its table predicts for that step: the predicted regions hold the
    stencils on a lattice larger than twice the body (no violation), and a
    table shrunk to one cell makes the run raise instead of dropping force."""
    bm = load_sub("batch")
    lib = load_sub("_lib")
    cfgs = [pkg.config.naca_case(96, 48, 6, 9.6, Re=200, name="wide"),
            pkg.config.naca_case(80, 40, 6, 8.0, Re=300, k=0.9, name="wide2")]
    sweep = bm.BatchSweep(cfgs, steps=12, solo_min_cells=1 << 30)
    assert all(not gr.solo for gr in sweep.groups)
    # at least one body's table is narrower than its lattice
    whole = [(0, 0, 0, c.Nx - 1, c.Ny - 1, c.Nz - 1) for c in cfgs]
    tabs = [r.cpu().numpy() for r in sweep._regions]
    assert any(tuple(t[0]) != w for t, w in zip(tabs, whole))
    sweep.step(5)
    assert all(i.region_violations() == 0 for i in sweep.ibms)
    seq, _ = bm.sequential(cfgs, 5)
    _assert_same(list(zip(sweep.lbms, sweep.ibms)), seq)
    sweep._regions[1][:] = 0                 # a region of one cell, (0, 0, 0)
    with pytest.raises(lib.ClbmError, match="instance 1"):
        sweep.step(2)


@pytest.mark.gpu
@pytest.mark.parametrize("storage,deviation", [("f16c", "inflow"), ("f16", "weights")])
def test_16bit_batch_equals_each_instance_alone(pkg, gpu, storage, deviation):
    bm = load_sub("batch")
    cfgs = _small(pkg)[:3]
    sweep = bm.BatchSweep(cfgs, storage=storage, deviation=deviation, steps=16)
    sweep.run(9, graph=True)
    seq, _ = bm.sequential(cfgs, 9, storage=storage, deviation=deviation)
    import torch
    for bl, (sl, _) in zip(sweep.lbms, seq):
        assert bl.deviation == deviation and bl.state_word == sl.state_word
        # the stored 16-bit codes themselves, not only their decoded values
        np.testing.assert_array_equal(bl.state.cpu().view(torch.uint8).numpy(),
                                      sl.state.cpu().view(torch.uint8).numpy())
    _assert_same(list(zip(sweep.lbms, sweep.ibms)), seq)


@pytest.mark.gpu
def test_hecht_harting_batch_equals_each_instance_alone(pkg, gpu):
    """The opt-in inlet reaches the batched kernels (the rule travels in each
    instance's argument record) and matches the per-instance path bit for bit."""
    bm = load_sub("batch")
    cfgs = _small(pkg)[:3]
    sweep = bm.BatchSweep(cfgs, steps=16, inlet="hecht-harting")
    sweep.run(7, graph=True)
    seq, _ = bm.sequential(cfgs, 7, inlet="hecht-harting")
    _assert_same(list(zip(sweep.lbms, sweep.ibms)), seq)
    ref, _ = bm.sequential(cfgs[:1], 7)           # the reference rule gives other populations
    assert not np.array_equal(ref[0][0].f.to_numpy(), seq[0][0].f.to_numpy())


@pytest.mark.gpu
@pytest.mark.parametrize("cache_bytes,ngroups", [(192 << 20, 1), (800_000, 2), (1, 3)])
def test_graph_replay_equals_eager(pkg, gpu, cache_bytes, ngroups):
    bm = load_sub("batch")
    cfgs = _small(pkg)[:3]
    a = bm.BatchSweep(cfgs, steps=32, cache_bytes=cache_bytes)
    b = bm.BatchSweep(cfgs, steps=32, cache_bytes=cache_bytes)
    assert len(a.groups) == ngroups
    a.run(11, graph=True)
    a.run(6, graph=True)       # resumes on the other parity: realigns, then replays
    b.step(17)
    _assert_same(list(zip(a.lbms, a.ibms)), list(zip(b.lbms, b.ibms)))
    seq, _ = bm.sequential(cfgs, 17)
    _assert_same(list(zip(a.lbms, a.ibms)), seq)


@pytest.mark.gpu
def test_solo_group_steps_its_own_objects_and_equals_each_instance_alone(pkg, gpu):
    """A lattice over solo_min_cells in a group of its own takes the split step
    (marker kernels on their own stream); the batch still equals the
    instances run one by one, bit for bit, across eager steps and replays."""
    bm = load_sub("batch")
    cfgs = _small(pkg)[:2] + [pkg.config.naca_case(96, 48, 4, 12.0, name="solo")]
    sweep = bm.BatchSweep(cfgs, steps=32, cache_bytes=1, solo_min_cells=10_000)
    assert [g.solo for g in sweep.groups] == [len(g.idx) == 1 and g.idx[0] == 2
                                            for g in sweep.groups]
    assert sum(g.solo for g in sweep.groups) == 1
    sweep.run(11, graph=True)
    assert sweep.lbms[2].last_step_kind in ("split", "fused")     # the split step ran
    sweep.step(2)
    sweep.run(4, graph=True)
    assert sweep.lbms[2].steps == 17 and sweep.lbms[0].steps == 17
    assert sweep.lbms[2].state_word == sweep.lbms[0].state_word == sweep.state
    seq, _ = bm.sequential(cfgs, 17)
    _assert_same(list(zip(sweep.lbms, sweep.ibms)), seq)


@pytest.mark.gpu
@pytest.mark.parametrize("lanes", [2, 3])
def test_lanes_step_groups_side_by_side_and_equal_each_instance_alone(pkg, gpu, lanes):
    """Groups advanced together on streams of their own (two solo lattices on
    the split step, batched groups replaying their graphs) give the same
    populations, marker forces and force fields as the instances run one by
    one, bit for bit; timing events come back for every group."""
    bm = load_sub("batch")
    cfgs = _small(pkg)[:3] + [pkg.config.naca_case(96, 48, 4, 12.0, name=f"solo{i}")
                              for i in range(2)]
    sweep = bm.BatchSweep(cfgs, steps=32, cache_bytes=lanes, solo_min_cells=10_000, lanes=lanes)
    assert sum(g.solo for g in sweep.groups) == 2 and len(sweep.groups) == 5
    times = []
    sweep.run(9, graph=True, times=times)
    sweep.step(1)
    sweep.run(6, graph=True, times=times)
    assert len(times) == 10
    import torch
    torch.cuda.synchronize()
    assert all(e0.elapsed_time(e1) > 0 for _, e0, e1 in times)
    assert all(l.steps == 16 and l.state_word == sweep.state for l in sweep.lbms)
    seq, _ = bm.sequential(cfgs, 16)
    _assert_same(list(zip(sweep.lbms, sweep.ibms)), seq)


@pytest.mark.gpu
def test_autotune_restores_the_initial_state(pkg, gpu, tmp_path):
    bm = load_sub("batch")
    cfgs = _small(pkg)[:2]
    sweep = bm.BatchSweep(cfgs, steps=64)
    times = sweep.autotune(steps=2, cache=str(tmp_path / "tune.json"))
    assert set(times) == {f"{b}/{g}/{l}" for b, g, l in bm.CANDIDATES}
    # one step after tuning (which took three per candidate): the forces of the
    # tuning's last step lie in another box and must be gone from fx, fy
    sweep.step(1)
    seq, _ = bm.sequential(cfgs, 1)
    _assert_same(list(zip(sweep.lbms, sweep.ibms)), seq)
    sweep.step(4)
    seq, _ = bm.sequential(cfgs, 5)
    _assert_same(list(zip(sweep.lbms, sweep.ibms)), seq)
    # the cached choice is reused
    again = bm.BatchSweep(cfgs, steps=8)
    assert again.autotune(cache=str(tmp_path / "tune.json")) == times


@pytest.mark.gpu
def test_batch_rejects_malformed_input(pkg, gpu):
    lib = load_sub("_lib")
    g = (lib.ClbmGrid * 2)(lib.ClbmGrid(8, 6, 4, lib.STORE_F32), lib.ClbmGrid(8, 6, 4, lib.STORE_F16))
    assert lib.load().clbm_batch_bytes(2, g, 256) == 0          # mixed storage
    assert b"storage" in lib.load().clbm_last_error()
    g1 = (lib.ClbmGrid * 1)(lib.ClbmGrid(8, 6, 4, lib.STORE_F32))
    assert lib.load().clbm_batch_bytes(1, g1, 96) == 0         # block size
    out = lib.ClbmBatch()
    with pytest.raises(lib.ClbmError, match="not built"):
        lib.call("clbm_batch_step", C.byref(out), 0, None)


@pytest.mark.parametrize("lds", [-1, 65537])
def test_batch_step_rejects_an_lds_reservation_out_of_range(pkg, lds):
    """Checked on the host before any launch (the tables are never read)."""
    lib = load_sub("_lib")
    b = lib.ClbmBatch(1, lib.STORE_F32, 256, 1, 16, 16, 16, lds)
    with pytest.raises(lib.ClbmError, match="lds must be"):
        lib.call("clbm_batch_step", C.byref(b), 0, None)


@pytest.mark.parametrize("case,match", [
    ("unbuilt", "not built"), ("sizes", "differ in size"), ("atomic", "fixed-point"),
    ("state", "unknown bits"), ("regions", "no region tables")])
def test_fused_batch_step_rejects_bad_arguments(pkg, case, match):
    """Checked on the host before any launch (the tables are never read)."""
    lib = load_sub("_lib")
    lb = lib.ClbmBatch(2, lib.STORE_F32, 256, 4, 16, 16, 16, 0)
    ib = lib.IbmBatch(2, lib.STORE_F32, lib.SPREAD_FIXED, 8, 4, 32, 16, 16, 16, 16, 16, 1)
    state = 0
    if case == "unbuilt":
        lb = lib.ClbmBatch()
    elif case == "sizes":
        ib.n = 3
    elif case == "atomic":
        ib.spread_mode = lib.SPREAD_ATOMIC_F32
    elif case == "regions":
        ib.regions = 0
    else:
        state = 8
    with pytest.raises(lib.ClbmError, match=match):
        lib.call("clbm_batch_step_acc", C.byref(lb), C.byref(ib), state, None)
    if case in ("atomic", "state", "regions"):
        with pytest.raises(lib.ClbmError, match=match):
            lib.call("ibm_batch_markers", C.byref(ib), state, None)
