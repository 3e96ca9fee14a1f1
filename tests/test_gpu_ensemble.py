"""Batched ensembles (bcs_ensemble): many veins of uneven size stepped by one
launch per frame kernel.  Every member must be bit-identical to the same
scene stepped alone (bcs_sim_step) -- under any launch tuning -- and, on a
subset, to the oracle."""
import numpy as np
import pytest

from conftest import assert_bits_equal

pytestmark = pytest.mark.gpu


def _scenes(bcs, n=6, fraction=0.1, seed=4):
    """Veins 0..n-1 of a 64-vein ensemble at a fraction of their cells (uneven
    radii and cell counts, two cell types each), plus a config-0 cylinder."""
    out = [bcs.Scene(params=bcs.ensemble_scene_params(64, k, seed=seed, cell_fraction=fraction)) for k in range(n)]
    out.append(bcs.Scene(0, 0.3, seed=seed))
    return out


def _sims(bcs, scenes):
    return [bcs.Simulation.from_scene(sc) for sc in scenes]


def _compare_members(sims, refs, what):
    for k, (a, b) in enumerate(zip(sims, refs)):
        for x, y, n in zip(a.get_particles(), b.get_particles(), ("pos", "vel", "force")):
            assert_bits_equal(x, y, f"{what}, member {k}: {n}")
        assert_bits_equal(a.get_vein()[0], b.get_vein()[0], f"{what}, member {k}: vein X")
        assert a.get_info().step_index == b.get_info().step_index


@pytest.mark.parametrize("tuning", [dict(blocks_per_cu=0, run=1, ray_mode=2, largest_first=1, contact_mode=1),
                                    dict(blocks_per_cu=2, run=4, ray_mode=1, largest_first=0, contact_mode=2),
                                    dict(blocks_per_cu=4, run=1, ray_mode=2, largest_first=1, contact_mode=2)])
def test_ensemble_members_equal_solo_runs(bcs, tuning):
    scenes = _scenes(bcs)
    solo = _sims(bcs, scenes)
    for s in solo:
        s.step(37)
    members = _sims(bcs, scenes)
    ens = bcs.Ensemble(members)
    ens.set_tuning(**tuning)
    ens.step(20)  # a 16-frame graph and single frames
    ens.step(17)
    _compare_members(members, solo, f"ensemble {tuning}")
    assert sum(m.np for m in members) > 50000


def test_ensemble_subset_matches_oracle(bcs, orc):
    scenes = _scenes(bcs, n=2, fraction=0.05)
    members = _sims(bcs, scenes)
    ens = bcs.Ensemble(members)
    ens.step(30)
    for k, (sc, m) in enumerate(zip(scenes, members)):
        o = orc.Oracle(sc.physics, sc.types, sc.mesh)
        o.set_particles(sc.positions, sc.velocities, sc.forces)
        o.step(30)
        for x, y, n in zip(m.get_particles(), o.get_particles(), ("pos", "vel", "force")):
            assert_bits_equal(x, y, f"member {k} vs oracle: {n}")
        assert_bits_equal(m.get_vein()[0], o.get_vein()[0], f"member {k} vs oracle: vein X")


def test_ensemble_autotune_reports_and_keeps_results(bcs):
    scenes = _scenes(bcs, n=3, fraction=0.05)
    solo = _sims(bcs, scenes)
    members = _sims(bcs, scenes)
    ens = bcs.Ensemble(members)
    rep = ens.autotune(frames_per_trial=2)
    frames = sum(1 + 2 for _ in rep["trials"])
    assert len(rep["trials"]) == 36
    assert rep["chosen"]["ms_per_frame"] == min(t["ms_per_frame"] for t in rep["trials"])
    t = ens.get_tuning()
    assert t["blocks_per_cu"] == rep["chosen"]["blocks_per_cu"] and t["run"] == rep["chosen"]["run"]
    ens.step(5)
    for s in solo:
        s.step(frames)
        s.step(5)
    _compare_members(members, solo, "after autotuning")


def test_ensemble_rejects_bad_members(bcs):
    sc = bcs.Scene(0, 0.05, seed=1)
    a = bcs.Simulation.from_scene(sc)
    with pytest.raises(bcs.BcsError):
        bcs.Ensemble([a, a])
    ph = sc.physics
    ph.dynamic_triangle_grid = 1
    b = bcs.Simulation.from_scene(sc, physics=ph)
    with pytest.raises(bcs.BcsError):
        bcs.Ensemble([b])
    ens = bcs.Ensemble([a])
    with pytest.raises(bcs.BcsError):
        ens.set_tuning(blocks_per_cu=-1)
    with pytest.raises(bcs.BcsError):
        ens.set_tuning(contact_mode=7)
