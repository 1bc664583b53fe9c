"""The batched ensemble's scene generator (exagoop_amd/ensemble.py): the
drawn parameters follow the stated distributions, scenes are reproducible
from their seed, and the particle records are well formed."""
import math

import numpy as np

from exagoop_amd import ensemble
from exagoop_amd.capi import I_MODEL, R_BULK, R_E, R_MASS, R_NU, R_VEL, R_VOLUME


def test_parameter_distributions_over_256_seeds():
    ps = [ensemble.draw(s) for s in range(256)]
    kinds = [p["kind"] for p in ps]
    assert 0.35 < kinds.count("disks") / 256 < 0.65
    t = np.log([p["target"] for p in ps])
    assert t.min() >= math.log(2048) - 1e-9 and t.max() <= math.log(65536) + 1e-9
    # log-uniform: the log's mean near the interval's midpoint
    assert abs(t.mean() - 0.5 * (math.log(2048) + math.log(65536))) < 0.3
    n = [p["n"] for p in ps]
    assert min(n) >= 16 and max(n) <= 48 and len(set(n)) > 20
    E = np.log10([p["E"] for p in ps])
    assert E.min() >= 3 and E.max() <= 5 and abs(E.mean() - 4) < 0.2
    nu = [p["nu"] for p in ps]
    assert min(nu) >= 0.2 and max(nu) <= 0.4
    sp = [p["speed"] for p in ps]
    assert min(sp) >= 0 and max(sp) <= 0.5
    for p in ps:
        assert abs(np.linalg.norm(p["direction"]) - 1) < 1e-12


def test_scenes_are_reproducible_and_well_formed():
    for seed in (0, 1, 2, 3, 11):
        s, pos, r, idata, prm = ensemble.scene(seed)
        s2, pos2, r2, idata2, _ = ensemble.scene(seed)
        assert np.array_equal(pos, pos2) and np.array_equal(r, r2)
        n = pos.shape[1]
        assert prm["particles"] == n
        # lattice discretisation of the target count
        assert 0.7 * prm["target"] <= n <= 1.3 * prm["target"]
        assert pos.min() > 0.0 and pos.max() < 1.0
        assert s.order_scheme == 2 and s.num_redist == 1 and list(s.ncells) == [prm["n"]] * 3
        assert np.all(r[R_MASS] > 0) and np.all(r[R_VOLUME] > 0)
        speed = np.sqrt((r[R_VEL:R_VEL + 3] ** 2).sum(axis=0))
        assert np.allclose(speed, prm["speed"])
        if prm["kind"] == "disks":
            assert np.all(idata[I_MODEL] == 0)
            assert np.all(r[R_E] == prm["E"]) and np.all(r[R_NU] == prm["nu"])
            # the two balls move in opposite directions
            c1 = np.asarray(ensemble.BALL_CENTRES[0])[:, None]
            first = ((pos - c1) ** 2).sum(axis=0) <= prm["radius"] ** 2
            vdir = np.asarray(prm["direction"]) * prm["speed"]
            v = r[R_VEL:R_VEL + 3]
            assert np.allclose(v[:, first], vdir[:, None])
            assert np.allclose(v[:, ~first], -vdir[:, None])
            assert 0 < first.sum() < n
        else:
            assert np.all(idata[I_MODEL] == 1) and np.all(r[R_BULK] == prm["E"])
            a, _, h = prm["block"]
            assert pos[0].max() <= a and pos[2].max() <= h


def test_full_ensemble_size():
    total = sum(ensemble.scene(s)[1].shape[1] for s in range(0, 256, 8))
    # 32 of 256 scenes; the mean of log-uniform [2k, 64k] is about 18.3k
    assert 0.3e6 < total < 1.0e6
