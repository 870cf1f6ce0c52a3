"""GPU tests of the drop-in boundary beyond SimStep: the reference's get_Spa* exports
(src/PhysicsCloth.h:72-77), its inline step written against them (src/Simulate.h:86-96)
from Python and from C++ (tests/abi_driver.cpp over include/yarnstep.hpp), and the
trajectory and frame formats fed from GPU rollouts (src/Saver.cpp:184-426,
src/Render.cpp:3-131).

Tolerances: the exported matrices within 1e-10 relative of the oracle's (Frobenius);
A against M - h^2 K_x - h K_v formed from the exports within 1e-13 (the same products
summed in another order); trajectories within 1e-8 of the oracle's at a 1e-12 solve
tolerance.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle.oracle import OracleCloth
from tests.golden_cases import cases, excited_state

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def frel(A, B):
    D = (A - B).tocsr()
    nb = np.sqrt((B.tocsr().data ** 2).sum())
    return np.sqrt((D.data ** 2).sum()) / (nb if nb > 0 else 1.0)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _pair(ys, spec, state):
    g = ys.Cloth(spec)
    g.set_ClothState(state)
    r = OracleCloth(spec)
    r.set_state(state)
    return g, r


def _global(st, spec, which):
    """LPos / LVel with the crossings' EPos / EVel in the global layout."""
    W, Lg = spec["width"], spec["length"]
    L, E = ("LPos", "EPos") if which == "pos" else ("LVel", "EVel")
    inter = [u + v * W for v in range(1, Lg - 1) for u in range(1, W - 1)]
    return np.concatenate([np.asarray(st[L]).ravel(), np.asarray(st[E])[inter].ravel()])


@pytest.mark.parametrize("idx", range(4))
def test_system_matrices_match_oracle(ys, cs, idx):
    """get_SpaM, get_SpaFPos, get_SpaFVel against the oracle's M, K_x, K_v; A and b
    formed from the exports as src/TrainHelper.cpp:19-20 forms them equal the assembled
    ones; get_SpaPos / get_SpaVel are the state in the global layout."""
    name, spec = cases(cs)[idx]
    say("case %s W=%d" % (name, spec["width"]))
    state = excited_state(cs, spec, 300 + idx)
    g, r = _pair(ys, spec, state)
    h = spec["h"]
    phy = ys.PhysicsCloth(g)
    phy.FillMV(h)
    r.assemble(h)
    A0, b0 = phy.get_SpaA(), g.vec("b")
    M, Kx, Kv = phy.get_SpaM(), phy.get_SpaFPos(), phy.get_SpaFVel()
    eM, eKx, eKv = frel(M, r.csr(1)), frel(Kx, r.csr(2)), frel(Kv, r.csr(3))
    say("  |dM|=%.2e |dKx|=%.2e |dKv|=%.2e" % (eM, eKx, eKv))
    assert eM < 1e-10 and eKx < 1e-10 and eKv < 1e-10
    A = phy.get_SpaA()
    assert frel(M - (h * h) * Kx - h * Kv, A) < 1e-13
    v = phy.get_SpaVel()
    b = h * (phy.get_SpaF() - Kv.T @ v) + M @ v
    assert rel(b, g.vec("b")) < 1e-12
    np.testing.assert_array_equal(phy.get_SpaPos(), _global(state, spec, "pos"))
    np.testing.assert_array_equal(v, _global(state, spec, "vel"))
    # the exports leave the assembled system in place: A and b bit for bit
    assert (A0 != phy.get_SpaA()).nnz == 0
    assert np.array_equal(b0, g.vec("b"))
    g.close()
    # and a solve after them matches the oracle's (from the rest pose: the excited states
    # above are not positive definite)
    g, r = _pair(ys, spec, cs.initial_state(spec))
    phy = ys.PhysicsCloth(g)
    phy.FillMV(h)
    r.assemble(h)
    phy.get_SpaM(), phy.get_SpaFPos(), phy.get_SpaFVel()
    it, rr, ok = g.solve(tol=1e-12)
    x_r, it_r, _, _ = r.solve(tol=1e-12, maxit=20000)
    say("  cg iters gpu %d oracle %d relres %.2e" % (it, it_r, rr))
    assert ok and rel(g.vec("x"), x_r) < 1e-9
    g.close()


@pytest.mark.parametrize("name", ["homo5", "heter7_twill"])
def test_inline_step_from_python(ys, cs, name):
    """The reference's inline step (src/Simulate.h:86-94) through the Python mirror:
    A and b formed on the host from the get_Spa* exports, EigenSolver.apply, then
    Cloth.update with F_E_is_Slides[0]; 30 steps against the oracle's SimStep."""
    spec = cs.homo_spec(5) if name == "homo5" else cs.heter_spec(7, "twill", 11)
    state = cs.initial_state(spec)
    g, r = _pair(ys, spec, state)
    h = spec["h"]
    phy = ys.PhysicsCloth(g)
    solver = ys.EigenSolver(g, tol=1e-12, maxit=20000)
    worst = 0.0
    for s in range(30):
        phy.FillMV(h)
        A = phy.get_SpaM() - phy.get_SpaFPos() * (h * h) - phy.get_SpaFVel() * h
        b = h * (phy.get_SpaF() - phy.get_SpaFVel().T @ phy.get_SpaVel()) + phy.get_SpaM() @ phy.get_SpaVel()
        new_vel = solver.apply(A, b)
        g.update_vel(new_vel, phy.F_E_is_Slides[0], h)
        phy.ResetMV()
        st, _, _ = r.step(h, None, tol=1e-12, maxit=20000)
        assert st == 0
        if s % 10 == 9:
            sg, sr = g.get_ClothState(), r.get_state()
            e = max(rel(sg[k], sr[k]) for k in ("LPos", "LVel", "EPos", "EVel", "EPosBar"))
            worst = max(worst, e)
            say("  step %d rel %.2e" % (s + 1, e))
    assert worst < 1e-8
    g.close()


def test_cpp_driver_inline_step_and_simstep(ys, cs, tmp_path):
    """tests/abi_driver.cpp, a reference-style C++ driver over include/yarnstep.hpp: its
    inline steps (src/Simulate.h:86-94) and its SimSteps both follow the oracle."""
    exe = os.path.join(ROOT, "tests", "abi_driver")
    if not os.path.exists(exe):
        pytest.fail("tests/abi_driver is not built (run __graft_entry__.build())")
    spec = cs.heter_spec(7, "satin", 5)
    state = cs.initial_state(spec)
    steps, h, tol = 25, spec["h"], 1e-12
    d = ys._abi.make_desc(spec)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(bytes(d))
        f.write(np.array([h, tol], dtype=np.float64).tobytes())
        f.write(np.array([steps], dtype=np.int32).tobytes())
        for k in ("LPos", "LPosFix", "LVel", "EPos", "EPosBar", "EVel"):
            f.write(np.ascontiguousarray(state[k], dtype=np.float64).tobytes())
    say("  abi_driver W=%d steps=%d" % (spec["width"], steps))
    subprocess.run([exe, str(inp), str(out)], check=True, timeout=120)
    N = spec["width"] * spec["length"]
    raw = np.fromfile(out, dtype=np.float64, count=2 * 15 * N)
    r = OracleCloth(spec)
    r.set_state(state)
    for _ in range(steps):
        assert r.step(h, None, tol=tol, maxit=20000)[0] == 0
    sr = r.get_state()
    for run in range(2):
        off = run * 15 * N
        got = {}
        for k, dim in (("LPos", 3), ("LPosFix", 3), ("LVel", 3), ("EPos", 2), ("EPosBar", 2), ("EVel", 2)):
            got[k] = raw[off:off + dim * N].reshape(N, dim)
            off += dim * N
        e = max(rel(got[k], sr[k]) for k in ("LPos", "LVel", "EPos", "EVel", "EPosBar"))
        say("  %s rel %.2e" % (("inline", "SimStep")[run], e))
        assert e < 1e-8


def test_load_system_rejects_entries_outside_the_stencil(ys, cs):
    import scipy.sparse as sp
    spec = cs.homo_spec(6)
    g = ys.Cloth(spec)
    n = g.n
    A = sp.identity(n, format="lil")
    A[0, 3 * 30] = 1.0     # node 0 and node 30 are far apart
    with pytest.raises(ys.YcError) as e:
        g.load_system(A.tocsr(), np.ones(n))
    assert e.value.code == ys.YC_ERR_ARG
    # a diagonal system solves to b / diag
    D = sp.diags(np.linspace(1.0, 2.0, n)).tocsr()
    x = ys.EigenSolver(g, tol=1e-14).apply(D, np.ones(n))
    np.testing.assert_allclose(x, 1.0 / np.linspace(1.0, 2.0, n), rtol=1e-13)
    g.close()


def _rollout(stepper, get, steps):
    fr = {k: [] for k in ("LPos", "LPosFix", "LVel", "EPos", "EPosBar", "EVel")}
    for s in range(steps):
        st = get()
        for k in fr:
            fr[k].append(np.asarray(st[k]).copy())
        stepper()
    return {k: np.stack(v) for k, v in fr.items()}


def test_h5_trajectory_from_gpu_rollout(ys, cs, tmp_path):
    """A GPU trajectory written in the reference's HDF5 layout reads back bit for bit,
    and matches the oracle's trajectory written the same way (1e-8)."""
    try:
        ys.h5io.lib()
    except OSError as e:
        pytest.skip("HDF5 C library not available: %s" % e)
    spec = cs.heter_spec(6, "twill", 21)
    state = cs.initial_state(spec)
    g, r = _pair(ys, spec, state)
    h = spec["h"]
    fg = _rollout(lambda: g.step(h, tol=1e-12, maxit=20000), g.get_ClothState, 20)
    fo = _rollout(lambda: r.step(h, None, tol=1e-12, maxit=20000), r.get_state, 20)
    pg, po = tmp_path / "gpu.h5", tmp_path / "oracle.h5"
    ys.h5io.save_trajectory(pg, spec, fg)
    ys.h5io.save_trajectory(po, spec, fo)
    size, props, bg = ys.h5io.load_trajectory(pg)
    _, _, bo = ys.h5io.load_trajectory(po)
    assert size == (6, 6)
    for k in fg:
        assert np.array_equal(bg[k], fg[k]), k
        assert rel(bg[k], bo[k]) < 1e-8, k
    g.close()


def test_obj_frames_from_gpu_rollout(ys, cs, tmp_path):
    """.obj frames (triangle and yarn forms) of a GPU rollout against the oracle's:
    the same lines, vertex coordinates within the writer's 5-decimal rounding."""
    spec = cs.heter_spec(7, "plain", 22)
    state = cs.initial_state(spec)
    g, r = _pair(ys, spec, state)
    h = spec["h"]
    for _ in range(15):
        g.step(h, tol=1e-12, maxit=20000)
        r.step(h, None, tol=1e-12, maxit=20000)
    for form in ("triangle", "yarn"):
        a, b = tmp_path / ("g_%s.obj" % form), tmp_path / ("o_%s.obj" % form)
        ys.render.render_cloth(a, spec, g.get_ClothState()["LPos"], form=form, normals=g.normals())
        ys.render.render_cloth(b, spec, r.get_state()["LPos"], form=form, normals=r.normals())
        la, lb = open(a).read().split("\n"), open(b).read().split("\n")
        assert len(la) == len(lb)
        for x, y in zip(la, lb):
            if x.startswith(("v ", "vn ")):
                va = np.array(x.split()[1:], dtype=float)
                vb = np.array(y.split()[1:], dtype=float)
                assert np.abs(va - vb).max() <= 1.01e-5, (x, y)
            else:
                assert x == y
    g.close()
