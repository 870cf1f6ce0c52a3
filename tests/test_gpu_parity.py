"""GPU parity: the HIP step against the CPU oracle on the same seeded inputs.

Tolerances: per-step F, A and b within 1e-10 relative (Frobenius / 2-norm over the
whole object); trajectories within 1e-6 relative over 100 steps.  Solves use a
1e-12 relative residual on both sides so the comparison measures the step, not the
solver tolerance.
"""
import sys

import numpy as np
import pytest

from oracle.oracle import OracleCloth
from tests.golden_cases import cases, excited_state

pytestmark = pytest.mark.gpu

TOL_STEP = 1e-10
TOL_TRAJ = 1e-6


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _pair(ys, spec, state):
    g = ys.Cloth(spec)
    g.set_ClothState(state)
    r = OracleCloth(spec)
    r.set_state(state)
    return g, r


@pytest.mark.parametrize("idx", range(4))
def test_assembly_matches_oracle(ys, cs, idx):
    name, spec = cases(cs)[idx]
    say("case %s W=%d" % (name, spec["width"]))
    state = excited_state(cs, spec, 100 + idx)
    g, r = _pair(ys, spec, state)
    h = spec["h"]
    g.assemble(h)
    r.assemble(h)
    np.testing.assert_allclose(g.normals(), r.normals(), atol=1e-12)
    A_g, A_r = g.csr().toarray(), r.csr(0).toarray()
    assert A_g.shape == A_r.shape
    # the oracle's stored entries lie inside the stencil pattern the GPU exports
    pat = g.csr()
    pat.data[:] = 1.0
    assert np.all(pat.toarray()[A_r != 0] == 1.0)
    eA = np.linalg.norm(A_g - A_r) / np.linalg.norm(A_r)
    eF = rel(g.vec("F"), r.vec("F"))
    eb = rel(g.vec("b"), r.vec("b"))
    sl_g, cp_g, fp_g = g.crossing()
    sl_r, cp_r, fp_r = r.crossing()
    say("  |dA|=%.2e |dF|=%.2e |db|=%.2e" % (eA, eF, eb))
    assert eA < TOL_STEP and eF < TOL_STEP and eb < TOL_STEP
    assert rel(cp_g, cp_r) < TOL_STEP
    assert rel(fp_g, fp_r) < TOL_STEP
    np.testing.assert_allclose(sl_g, sl_r, rtol=1e-8, atol=1e-12)


@pytest.mark.parametrize("name", ["homo5", "heter64_twill"])
def test_workload_first_step(ys, cs, name):
    say("workload %s" % name)
    spec, state, ext, _ = cs.workload(name)
    g, r = _pair(ys, spec, state)
    h = spec["h"]
    g.assemble(h)
    r.assemble(h, ext)
    eA = np.linalg.norm((g.csr() - r.csr(0)).toarray()) / np.linalg.norm(r.csr(0).toarray())
    eb = rel(g.vec("b"), r.vec("b"))
    say("  |dA|=%.2e |db|=%.2e" % (eA, eb))
    assert eA < TOL_STEP and eb < TOL_STEP


def test_solve_matches_oracle(ys, cs):
    spec = cases(cs)[2][1]
    state = cs.initial_state(spec)
    g, r = _pair(ys, spec, state)
    h = spec["h"]
    g.assemble(h)
    r.assemble(h)
    it_g, rr_g, ok = g.solve(tol=1e-12)
    x_r, it_r, rr_r, st = r.solve(tol=1e-12, maxit=20000)
    say("  cg iters gpu %d oracle %d relres %.2e %.2e" % (it_g, it_r, rr_g, rr_r))
    assert ok and rr_g < 1e-12
    assert rel(g.vec("x"), x_r) < 1e-9


def test_trajectory_100_steps(ys, cs):
    spec, state, ext, steps = cs.workload("homo5")
    g, r = _pair(ys, spec, state)
    h = spec["h"]
    worst = 0.0
    for s in range(steps):
        it, rr, rc = g.step(h, tol=1e-12, maxit=20000)
        st, it_r, rr_r = r.step(h, ext, tol=1e-12, maxit=20000)
        assert rc == 0 and st == 0
        if s % 10 == 9 or s == steps - 1:
            sg, sr = g.get_ClothState(), r.get_state()
            e = max(rel(sg[k], sr[k]) for k in ("LPos", "LVel", "EPos", "EVel", "EPosBar"))
            worst = max(worst, e)
            say("  step %d rel %.2e" % (s + 1, e))
    assert worst < TOL_TRAJ


def test_large_cloth_first_step(ys, cs):
    """The 256 x 256 cloth with its seeded pressure (327k DOF): A and b against the oracle
    within 1e-10 (sparse Frobenius norms), then a fused-form solve whose true residual
    |b - A x| / |b|, computed with the oracle's A, meets the tolerance."""
    spec, state, ext, _ = cs.workload("homo256")
    g, r = _pair(ys, spec, state)
    h = spec["h"]
    g.set_cg_variant("fused")
    g.set_ext_force(ext)
    g.assemble(h)
    r.assemble(h, ext)
    Ar = r.csr(0).tocsr()
    D = (g.csr() - Ar).tocsr()
    eA = np.sqrt((D.data ** 2).sum()) / np.sqrt((Ar.data ** 2).sum())
    br = r.vec("b")
    eb = rel(g.vec("b"), br)
    say("  256x256 |dA|=%.2e |db|=%.2e nnz %d" % (eA, eb, Ar.nnz))
    assert eA < TOL_STEP and eb < TOL_STEP
    it, rr, ok = g.solve(tol=1e-8)
    x = g.vec("x")
    res = np.linalg.norm(br - Ar @ x) / np.linalg.norm(br)
    say("  fused CG iters %d recursive relres %.2e true relres %.2e" % (it, rr, res))
    assert ok and rr < 1e-8 and res < 1.5e-8
    g.close()


@pytest.mark.parametrize("W,Lg,pattern", [(12, 7, "twill"), (7, 12, "satin")])
def test_non_square_cloth(ys, cs, W, Lg, pattern):
    """Cloths with width != length (the tiles of A are then not square in number):
    assembly against the oracle at 1e-10 on an excited state, and a 30-step trajectory
    of the persistent form within 1e-6."""
    spec = cs.heter_spec(W, pattern, W * 100 + Lg)
    spec["length"] = Lg
    spec["handles"] = [(0, 0), (W - 1, 0)]
    st = excited_state(cs, spec, 7)
    g, r = _pair(ys, spec, st)
    h = spec["h"]
    g.assemble(h)
    r.assemble(h)
    eA = np.linalg.norm((g.csr() - r.csr(0)).toarray()) / np.linalg.norm(r.csr(0).toarray())
    eb = rel(g.vec("b"), r.vec("b"))
    say("  %dx%d |dA|=%.2e |db|=%.2e" % (W, Lg, eA, eb))
    assert eA < TOL_STEP and eb < TOL_STEP
    g.close()
    st0 = cs.initial_state(spec)
    g, r = _pair(ys, spec, st0)
    g.set_cg_variant("persistent")
    for _ in range(30):
        _, _, rc = g.step(h, tol=1e-12, maxit=20000)
        st_, _, _ = r.step(h, None, tol=1e-12, maxit=20000)
        assert rc == 0 and st_ == 0
    sg, sr = g.get_ClothState(), r.get_state()
    e = max(rel(sg[k], sr[k]) for k in ("LPos", "LVel", "EPos", "EVel", "EPosBar"))
    say("  %dx%d 30-step rel %.2e" % (W, Lg, e))
    assert e < 1e-6
    g.close()


def test_device_state_round_trip(ys, cs):
    """yc_set_state_device / yc_get_state_device with torch tensors on the device: the
    same state as the host path, in the caller's stream order, and a step from a device
    state equals a step from the same host state."""
    import torch
    spec = cs.heter_spec(12, "twill", 3)
    spec["length"] = 9   # non-square: the state index is u + v * width
    c0 = ys.Cloth(spec)   # a moving state: three steps from rest
    c0.set_ClothState(cs.initial_state(spec))
    c0.run(spec["h"], 3, tol=1e-10)
    st = c0.get_ClothState()
    c0.close()
    a, b = ys.Cloth(spec), ys.Cloth(spec)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        dev = {k: torch.from_numpy(np.ascontiguousarray(v)).to("cuda", non_blocking=False) for k, v in st.items()}
        a.set_ClothState_device(dev, stream=s)
        for k in dev:
            dev[k].zero_()    # queued after the transfer on the same stream
        back = a.get_ClothState_device(stream=s)
    s.synchronize()
    b.set_ClothState(st)
    host = a.get_ClothState()
    for k in st:
        assert np.array_equal(host[k], np.asarray(st[k])), k
        assert np.array_equal(back[k].cpu().numpy(), np.asarray(st[k])), k
    for c in (a, b):
        it, rr, rc = c.step(spec["h"], tol=1e-10)
        assert rc == 0
    sa, sb = a.get_ClothState(), b.get_ClothState()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    a.close()
    b.close()
