"""GPU tests of batched stepping (yc_batch_*) and the launch-shape autotuner.

A batch runs every cloth's step through the single-cloth kernels' own bodies, so each
cloth's trajectory must equal, bit for bit, the one ``Cloth.run`` gives with the
persistent CG; the single-cloth path is itself checked against the oracle in
test_gpu_cg.py, and one small batch is checked against the oracle directly here
(trajectories within 1e-6 relative).
"""
import sys

import numpy as np
import pytest

from oracle.oracle import OracleCloth

pytestmark = pytest.mark.gpu

KEYS = ("LPos", "LPosFix", "LVel", "EPos", "EPosBar", "EVel")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _cloths(ys, items, precond="jacobi"):
    out = []
    for spec, state in items:
        c = ys.Cloth(spec)
        c.set_cg_variant("persistent")
        c.set_preconditioner(precond)
        c.set_ClothState(state)
        out.append(c)
    return out


def _alone(ys, items, h, steps, tol, precond="jacobi"):
    """Each cloth stepped on its own (persistent CG, queued steps)."""
    res = []
    for spec, state in items:
        c = ys.Cloth(spec)
        c.set_cg_variant("persistent")
        c.set_preconditioner(precond)
        c.set_ClothState(state)
        it, rc = c.run(h, steps, tol=tol)
        assert rc == 0
        res.append((c.get_ClothState(), it))
        c.close()
    return res


def _small_items(cs):
    items = cs.batch_workload(n=7, seed=5, lo=5, hi=40)
    spec = cs.homo_spec(9, "twill")
    spec["wind_vel"] = [0.0, 4.0, 0.0]
    items.append((spec, cs.initial_state(spec)))
    spec = cs.heter_spec(10, "satin", 17)   # a non-square cloth: 10 x 17 crossings
    spec["length"] = 17
    items.append((spec, cs.initial_state(spec)))
    return items


@pytest.mark.parametrize("graph,order,nb,precond", [(1, 0, 64, "jacobi"), (0, 1, 128, "jacobi"), (1, 1, 256, "jacobi"),
                                                    (1, 0, 64, "block2"), (0, 1, 128, "block2")])
def test_batch_equals_cloths_alone(ys, cs, graph, order, nb, precond):
    items = _small_items(cs)
    h, steps, tol = items[0][0]["h"], 6, 1e-8
    say("batch of %s, graph %d order %d node_block %d precond %s" % ([s["width"] for s, _ in items], graph, order, nb,
                                                                    precond))
    ref = _alone(ys, items, h, steps, tol, precond)
    cl = _cloths(ys, items, precond)
    b = ys.ClothBatch(cl)
    b.set_opts(graph=graph, group_order=order, node_block=nb)
    it, st, rc = b.run(h, steps, tol=tol)
    say("  info %s iters %s" % (b.info(), list(it)))
    assert rc == 0 and not st.any()
    for i, c in enumerate(cl):
        s = c.get_ClothState()
        for k in KEYS:
            assert np.array_equal(s[k], ref[i][0][k]), (i, k)
        assert it[i] == ref[i][1]
    b.close()
    for c in cl:
        c.close()


@pytest.mark.parametrize("precond", ["jacobi", "block2"])
def test_batch_uneven_64_cloths_several_groups(ys, cs, precond):
    """The 64-cloth workload (sides 10..64, 1,743 tiles) needs more tiles than the
    device holds at once, so its solves run as several resident groups; two runs
    of one step each equal the cloths stepped alone."""
    items = cs.batch_workload()
    h, tol = items[0][0]["h"], 1e-8
    say("64-cloth batch, precond %s" % precond)
    ref = _alone(ys, items, h, 2, tol, precond)
    cl = _cloths(ys, items, precond)
    b = ys.ClothBatch(cl)
    info = b.info()
    say("  64-cloth batch info %s" % info)
    assert info["tiles"] == sum(((s["width"] + 7) // 8) * ((s["length"] + 7) // 8) for s, _ in items)
    it1, st1, rc1 = b.run(h, 1, tol=tol)
    it2, st2, rc2 = b.run(h, 1, tol=tol)
    assert rc1 == 0 and rc2 == 0
    for i, c in enumerate(cl):
        s = c.get_ClothState()
        for k in KEYS:
            assert np.array_equal(s[k], ref[i][0][k]), (i, k)
        assert it1[i] + it2[i] == ref[i][1]
    b.close()
    for c in cl:
        c.close()


def test_batch_matches_oracle(ys, cs):
    specs = [cs.homo_spec(5), cs.heter_spec(7, "satin", 7), cs.heter_spec(11, "twill", 11)]
    specs[2]["wind_vel"] = [0.0, 5.0, 0.0]
    items = [(s, cs.initial_state(s)) for s in specs]
    h, steps = items[0][0]["h"], 30
    cl = _cloths(ys, items)
    b = ys.ClothBatch(cl)
    it, st, rc = b.run(h, steps, tol=1e-12, maxit=20000)
    assert rc == 0
    for (spec, state), c in zip(items, cl):
        r = OracleCloth(spec)
        r.set_state(state)
        for _ in range(steps):
            st_, _, _ = r.step(h, None, tol=1e-12, maxit=20000)
            assert st_ == 0
        sg, sr = c.get_ClothState(), r.get_state()
        e = max(rel(sg[k], sr[k]) for k in ("LPos", "LVel", "EPos", "EVel", "EPosBar"))
        say("  %dx%d rel %.2e" % (spec["width"], spec["length"], e))
        assert e < 1e-6
    b.close()
    for c in cl:
        c.close()


def test_batch_rejects_bad_input(ys, cs):
    spec = cs.homo_spec(6)
    c = ys.Cloth(spec)
    with pytest.raises(ys.YcError):
        ys.ClothBatch([c, c])
    b = ys.ClothBatch([c])
    with pytest.raises(ys.YcError):
        b.set_opts(node_block=96)
    b.close()
    c2 = ys.Cloth(spec)
    c2.set_cg_variant("persistent")
    c2.set_preconditioner("block2")
    with pytest.raises(ys.YcError):   # one preconditioner per batch
        ys.ClothBatch([c, c2])
    c2.close()
    c.close()


def test_batch_reports_maxit_per_cloth(ys, cs):
    items = _small_items(cs)[:3]
    cl = _cloths(ys, items)
    b = ys.ClothBatch(cl)
    it, st, rc = b.run(items[0][0]["h"], 1, tol=1e-12, maxit=3)
    assert rc == ys.YC_ERR_NOT_CONVERGED
    assert (st == ys.YC_ERR_NOT_CONVERGED).all()
    assert (it == 3).all()
    b.close()
    for c in cl:
        c.close()


def test_autotune_batch_keeps_results(ys, cs, tmp_path):
    items = _small_items(cs)
    h = items[0][0]["h"]
    ref = _alone(ys, items, h, 3, 1e-8)
    cl = _cloths(ys, items)
    b = ys.ClothBatch(cl)
    cache = str(tmp_path / "tune.json")
    best, times = ys.autotune_batch(b, h, steps=2, cache=cache,
                                    candidates=[(1, 0, 64), (0, 0, 64), (1, 1, 128), (0, 1, 256)])
    say("  best %s times %s" % (best, {k: "%.3f ms" % (1e3 * v) for k, v in times.items()}))
    assert (best["graph"], best["group_order"], best["node_block"]) == min(times, key=times.get)
    assert b.opts() == best
    it, st, rc = b.run(h, 3, tol=1e-8)
    assert rc == 0
    for i, c in enumerate(cl):
        for k in KEYS:
            assert np.array_equal(c.get_ClothState()[k], ref[i][0][k]), (i, k)
    best2, times2 = ys.autotune_batch(b, h, cache=cache)   # from the cache
    assert best2 == best and times2 == {}
    b.close()
    for c in cl:
        c.close()


def test_autotune_cloth_picks_a_cg_form(ys, cs):
    spec = cs.heter_spec(24, "twill", 24)
    c = ys.Cloth(spec)
    c.set_ClothState(cs.initial_state(spec))
    best, times = ys.autotune_cloth(c, spec["h"], steps=2)
    say("  cloth best %s %s" % (best, times))
    assert best in times and set(times) <= set(ys.Cloth.CG_VARIANTS)
    c.close()


def test_batch_external_force_changes(ys, cs):
    """An external force set on one cloth between runs is picked up by the next run
    (the replayed step is re-captured), as stepping the cloths alone does."""
    items = _small_items(cs)[:4]
    h = items[0][0]["h"]
    spec1, st1 = items[1]
    f = cs.pressure_force(spec1, st1, 1.5)
    ref = []
    for i, (spec, state) in enumerate(items):
        c = ys.Cloth(spec)
        c.set_cg_variant("persistent")
        c.set_ClothState(state)
        c.run(h, 2, tol=1e-8)
        if i == 1:
            c.set_ext_force(f)
        c.run(h, 2, tol=1e-8)
        if i == 1:
            c.set_ext_force(None)
        c.run(h, 2, tol=1e-8)
        ref.append(c.get_ClothState())
        c.close()
    cl = _cloths(ys, items)
    b = ys.ClothBatch(cl)
    b.run(h, 2, tol=1e-8)
    cl[1].set_ext_force(f)
    b.run(h, 2, tol=1e-8)
    cl[1].set_ext_force(None)
    b.run(h, 2, tol=1e-8)
    for i, c in enumerate(cl):
        for k in KEYS:
            assert np.array_equal(c.get_ClothState()[k], ref[i][k]), (i, k)
    b.close()
    for c in cl:
        c.close()
