"""GPU parity of the two CG forms and of the tiled system matrix.

The fused form (one kernel and one reduction per iteration) runs the reference
solver's Jacobi-PCG recurrences with the p.Ap inner product expanded; its iterates
equal the reference form's up to rounding.  Tolerances: solutions within 1e-9
relative of the oracle's at a 1e-12 residual, iteration counts within 1 of the
oracle's, trajectories within 1e-6 relative over 100 steps.
"""
import sys

import numpy as np
import pytest

from oracle.oracle import OracleCloth
from tests.golden_cases import excited_state

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _specs(cs):
    # grids that are not multiples of the 8 x 8 matrix tile, and one that is
    a = cs.heter_spec(21, "plain", 21)
    b = cs.heter_spec(13, "satin", 13)
    b["wind_vel"] = [0.5, 3.0, 0.0]
    c = cs.homo_spec(16, "twill")
    return [("heter21_plain", a), ("heter13_satin_wind", b), ("homo16_twill", c)]


@pytest.mark.parametrize("idx", range(3))
def test_tiled_assembly_matches_oracle(ys, cs, idx):
    name, spec = _specs(cs)[idx]
    say("case %s W=%d" % (name, spec["width"]))
    st = excited_state(cs, spec, 300 + idx)
    g = ys.Cloth(spec)
    g.set_ClothState(st)
    r = OracleCloth(spec)
    r.set_state(st)
    h = spec["h"]
    g.assemble(h)
    r.assemble(h)
    eA = np.linalg.norm((g.csr() - r.csr(0)).toarray()) / np.linalg.norm(r.csr(0).toarray())
    eb = rel(g.vec("b"), r.vec("b"))
    say("  |dA|=%.2e |db|=%.2e" % (eA, eb))
    assert eA < 1e-10 and eb < 1e-10
    g.close()


VARIANTS = ["reference", "fused", "persistent"]


@pytest.mark.parametrize("idx", range(3))
@pytest.mark.parametrize("variant", VARIANTS)
def test_cg_variants_match_oracle(ys, cs, idx, variant):
    name, spec = _specs(cs)[idx]
    say("case %s variant %s" % (name, variant))
    st = cs.initial_state(spec)
    g = ys.Cloth(spec)
    g.set_cg_variant(variant)
    g.set_ClothState(st)
    r = OracleCloth(spec)
    r.set_state(st)
    h = spec["h"]
    g.assemble(h)
    r.assemble(h)
    for tol in (1e-8, 1e-12):
        it_g, rr_g, ok = g.solve(tol=tol, maxit=20000)
        x_r, it_r, rr_r, _ = r.solve(tol=tol, maxit=20000)
        e = rel(g.vec("x"), x_r)
        say("  tol %g iters gpu %d oracle %d relres %.2e %.2e |dx| %.2e" % (tol, it_g, it_r, rr_g, rr_r, e))
        assert ok and rr_g < tol
        assert abs(it_g - it_r) <= (0 if variant == "reference" else 1)
        assert e < (1e-9 if tol == 1e-12 else 1e-5)
    g.close()


@pytest.mark.parametrize("variant", VARIANTS)
def test_cg_variants_maxit_and_warm_start(ys, cs, variant):
    """maxit stops with the reference's count and residual; a start at the solution
    returns at once with zero iterations."""
    name, spec = _specs(cs)[0]
    st = cs.initial_state(spec)
    g = ys.Cloth(spec)
    g.set_cg_variant(variant)
    g.set_ClothState(st)
    r = OracleCloth(spec)
    r.set_state(st)
    h = spec["h"]
    g.assemble(h)
    r.assemble(h)
    it_g, rr_g, ok = g.solve(tol=1e-14, maxit=7)
    x_r, it_r, rr_r, _ = r.solve(tol=1e-14, maxit=7)
    say("  %s maxit 7: iters %d %d relres %.3e %.3e" % (variant, it_g, it_r, rr_g, rr_r))
    assert not ok and it_g == 7 and it_r == 7
    assert rel(g.vec("x"), x_r) < 1e-10
    assert abs(rr_g - rr_r) <= 1e-8 * rr_r
    x_exact, _, _, _ = r.solve(tol=1e-14, maxit=20000)
    it_g, rr_g, ok = g.solve(tol=1e-6, maxit=100, x0=x_exact)
    assert ok and it_g == 0
    g.close()


@pytest.mark.parametrize("variant", ["fused", "persistent"])
def test_fused_trajectory_100_steps(ys, cs, variant):
    spec, state, ext, steps = cs.workload("homo5")
    g = ys.Cloth(spec)
    g.set_cg_variant(variant)
    g.set_ClothState(state)
    r = OracleCloth(spec)
    r.set_state(state)
    h = spec["h"]
    worst = 0.0
    for s in range(steps):
        it, rr, rc = g.step(h, tol=1e-12, maxit=20000)
        st_, it_r, rr_r = r.step(h, ext, tol=1e-12, maxit=20000)
        assert rc == 0 and st_ == 0
        if s % 20 == 19:
            sg, sr = g.get_ClothState(), r.get_state()
            e = max(rel(sg[k], sr[k]) for k in ("LPos", "LVel", "EPos", "EVel", "EPosBar"))
            worst = max(worst, e)
            say("  step %d rel %.2e iters %d/%d" % (s + 1, e, it, it_r))
    assert worst < 1e-6
    g.close()


def test_fused_heter64_steps_match_reference_variant(ys, cs):
    """Ten steps of the 64 x 64 twill blend with both CG forms at 1e-10: states agree
    to the solve tolerance and the iteration counts stay within 1 per step."""
    spec, state, ext, _ = cs.workload("heter64_twill")
    out = {}
    for variant in VARIANTS:
        g = ys.Cloth(spec)
        g.set_cg_variant(variant)
        g.set_ClothState(state)
        its = []
        for _ in range(10):
            it, rr, rc = g.step(spec["h"], tol=1e-10, maxit=20000)
            assert rc == 0
            its.append(it)
        out[variant] = (g.get_ClothState(), its)
        g.close()
    sa, ia = out["reference"]
    for v in VARIANTS[1:]:
        sb, ib = out[v]
        e = max(rel(sb[k], sa[k]) for k in ("LPos", "LVel", "EPos", "EVel"))
        say("  iters ref %s %s %s |dstate| %.2e" % (ia, v, ib, e))
        assert max(abs(a - b) for a, b in zip(ia, ib)) <= 1
        assert e < 1e-7


def test_persistent_solves_from_two_threads(ys, cs):
    """Two contexts stepping the 128 x 128 twill from two host threads queue persistent
    solves at the same time.  Each solve's 256 tile blocks take the whole device, and the
    blocks of a solve wait on one another, so the library runs the solves of different
    contexts one at a time.  Both threaded runs must return OK with the trajectory and
    iteration count of the same run alone, bit for bit."""
    import threading
    spec, state, ext, _ = cs.workload("twill128")
    h, steps = spec["h"], 12

    def make():
        c = ys.Cloth(spec)
        c.set_cg_variant("persistent")
        c.set_preconditioner("block2")
        c.set_ClothState(state)
        return c

    runs = {"zero": lambda c: c.run(h, steps, tol=1e-8), "quad": lambda c: c.run_guess("quad", h, steps, tol=1e-8)}
    alone = {}
    for k, fn in runs.items():
        c = make()
        alone[k] = (fn(c), c.get_ClothState())
        c.close()
    ctx = {k: make() for k in runs}
    got = {}

    def go(k):
        try:
            got[k] = runs[k](ctx[k])
        except Exception as e:   # reported below, in the main thread
            got[k] = e

    threads = [threading.Thread(target=go, args=(k,)) for k in runs]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=60)
    assert not any(t.is_alive() for t in threads)
    for k in runs:
        assert not isinstance(got[k], Exception), got[k]
        (it, rc), s = got[k], ctx[k].get_ClothState()
        (it0, rc0), s0 = alone[k]
        say("  %s: threaded iters %d rc %d, alone iters %d rc %d" % (k, it, rc, it0, rc0))
        assert rc == 0 and rc0 == 0 and it == it0
        for f in s0:
            assert np.array_equal(s[f], s0[f]), (k, f)
        ctx[k].close()
