"""The configured workloads end to end against the oracle.

- 64 x 64 two-yarn blends in the reference's plain, twill and satin patterns: the first
  step's A, F and b within 1e-10 relative, then 200 steps of the persistent CG with the
  block-Jacobi preconditioner, states within 1e-6 relative of the oracle's every 50
  steps (both solve to 1e-12: this cloth amplifies solve-tolerance differences).
- 256 x 256 plain with its seeded pressure: 50 steps of the fused CG against the
  oracle's 50 steps (states within 1e-6 every 10 steps, both to 1e-10), every solve's
  recursive residual below the tolerance, and the true residual |b - A x| / |b| of the
  first and last step's solution below 2e-10 with A and b exported from the device.

The oracle runs on the host's cores (its multi-threaded assembly is bit for bit the
single-thread one; its CG sums inner products over fixed chunks).
"""
import os
import sys

import numpy as np
import pytest

from oracle.oracle import OracleCloth

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _threads():
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def _oracle(spec, state):
    r = OracleCloth(spec)
    r.set_threads(_threads())
    r.set_state(state)
    return r


@pytest.mark.parametrize("pattern", ["plain", "satin"])
def test_heter64_first_step(ys, cs, pattern):
    spec, state, ext, _ = cs.workload("heter64_" + pattern)
    say("heter64_%s first step" % pattern)
    g = ys.Cloth(spec)
    g.set_ClothState(state)
    r = _oracle(spec, state)
    h = spec["h"]
    g.assemble(h)
    r.assemble(h)
    Ar = r.csr(0)
    D = (g.csr() - Ar).tocsr()
    eA = np.sqrt((D.data ** 2).sum()) / np.sqrt((Ar.data ** 2).sum())
    eF, eb = rel(g.vec("F"), r.vec("F")), rel(g.vec("b"), r.vec("b"))
    say("  |dA|=%.2e |dF|=%.2e |db|=%.2e" % (eA, eF, eb))
    assert eA < 1e-10 and eF < 1e-10 and eb < 1e-10
    g.close()


@pytest.mark.parametrize("pattern", ["plain", "twill", "satin"])
def test_heter64_200_steps(ys, cs, pattern):
    spec, state, ext, steps = cs.workload("heter64_" + pattern)
    assert steps == 200
    say("heter64_%s: %d steps" % (pattern, steps))
    g = ys.Cloth(spec)
    g.set_cg_variant("persistent")
    g.set_preconditioner("block2")
    g.set_ClothState(state)
    r = _oracle(spec, state)
    h = spec["h"]
    worst = 0.0
    for chunk in range(steps // 50):
        _, rc = g.run(h, 50, tol=1e-12, maxit=20000)
        assert rc == 0
        for _ in range(50):
            assert r.step(h, None, tol=1e-12, maxit=20000)[0] == 0
        sg, so = g.get_ClothState(), r.get_state()
        e = max(rel(sg[k], so[k]) for k in ("LPos", "LVel", "EPos", "EVel", "EPosBar"))
        worst = max(worst, e)
        say("  step %d: state rel %.2e" % (50 * (chunk + 1), e))
    g.close()
    assert worst < 1e-6


def test_homo256_50_steps(ys, cs):
    spec, state, ext, steps = cs.workload("homo256")
    assert steps == 50
    say("homo256: %d steps" % steps)
    g = ys.Cloth(spec)
    g.set_cg_variant("fused")
    g.set_ClothState(state)
    g.set_ext_force(ext)
    r = _oracle(spec, state)
    h, tol = spec["h"], 1e-10
    worst = 0.0
    for s in range(steps):
        check = s in (0, steps - 1)
        if check:
            g.assemble(h)
            it, rr, ok = g.solve(tol=tol, maxit=20000)
            A, b, x = g.csr(), g.vec("b"), g.vec("x")
            true = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
            say("  step %d: iterations %d, recursive relres %.2e, true relres %.2e" % (s + 1, it, rr, true))
            assert ok and rr < tol and true < 2 * tol
            g.update(h)
        else:
            it, rr, rc = g.step(h, tol=tol, maxit=20000)
            assert rc == 0 and rr < tol
        assert r.step(h, ext, tol=tol, maxit=20000)[0] == 0
        if s % 10 == 9:
            sg, so = g.get_ClothState(), r.get_state()
            e = max(rel(sg[k], so[k]) for k in ("LPos", "LVel", "EPos", "EVel", "EPosBar"))
            worst = max(worst, e)
            say("  step %d: state rel %.2e" % (s + 1, e))
    g.close()
    assert worst < 1e-6
