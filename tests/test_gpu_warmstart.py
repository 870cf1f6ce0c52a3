"""Learned warm start on the GPU: the MFMA forward and backward against the CPU
checker (oracle/mlp_ref.py), training reduces the loss, and on a cloth the
warm-started CG reaches the same tolerance in fewer iterations with the same
trajectory as zero-guess CG."""
import sys

import numpy as np
import pytest

from oracle import mlp_ref

pytestmark = pytest.mark.gpu


def say(m):
    print(m, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ws(ys):
    import yarnstep.warmstart as w
    return w


def _rand_setup(ws, seed, rows):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(rows, 32)).astype(np.float32) * 3 + 1
    T = rng.normal(size=(rows, 5)).astype(np.float32)
    M = (rng.uniform(size=(rows, 5)) > 0.3).astype(np.float32)
    mu = X.mean(0)
    sd = X.std(0)
    p = ws.init_params(seed) + np.random.default_rng(seed + 1).normal(0, 0.02, 37893).astype(np.float32)
    return X, T, M, mu, sd, p


def test_forward_matches_checker(ws):
    X, T, M, mu, sd, p = _rand_setup(ws, 1, 300)   # not a multiple of the 64-row tile
    say("mlp forward rows=300")
    net = ws.WarmStartMLP(params=p)
    net.set_norm(mu, sd, np.ones(5, np.float32))
    Y = net.forward(X)
    Yr, _, _ = mlp_ref.forward(p, X, mu, sd)
    err = np.linalg.norm(Y - Yr) / np.linalg.norm(Yr)
    say("  rel err %.2e" % err)
    assert err < 2e-3


def test_gradient_matches_checker(ws):
    X, T, M, mu, sd, p = _rand_setup(ws, 2, 1000)
    say("mlp gradient rows=1000")
    net = ws.WarmStartMLP(params=p)
    net.set_norm(mu, sd, np.ones(5, np.float32))
    net.set_data(X, T, M)
    idx = np.random.default_rng(3).permutation(1000)[:700].astype(np.int32)
    g, loss = net.grad(idx)
    lr, gr = mlp_ref.loss_grad(p, X[idx], T[idx], M[idx], mu, sd)
    err = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    say("  loss %.6f vs %.6f, grad rel err %.2e" % (loss, lr, err))
    assert abs(loss - lr) < 1e-4 * abs(lr)
    assert err < 5e-3
    # and per layer, so a wrong block cannot hide behind a large one
    o = 0
    for fo, fi in mlp_ref.SHAPES:
        for sz in (fo * fi, fo):
            e = np.linalg.norm(g[o:o + sz] - gr[o:o + sz]) / max(np.linalg.norm(gr[o:o + sz]), 1e-30)
            assert e < 2e-2, (o, e)
            o += sz


def test_training_reduces_loss(ws):
    rng = np.random.default_rng(5)
    X = rng.normal(size=(8192, 32)).astype(np.float32)
    teacher = ws.init_params(77)
    Yt, _, _ = mlp_ref.forward(teacher, X, np.zeros(32, np.float32), np.ones(32, np.float32), quant=False)
    T = (Yt / Yt.std(0)).astype(np.float32)
    M = np.ones_like(T)
    say("mlp training on a teacher network")
    net = ws.WarmStartMLP(seed=0)
    net.set_norm(np.zeros(32, np.float32), np.ones(32, np.float32), np.ones(5, np.float32))
    net.set_data(X, T, M)
    losses = [net.train_epoch(rng.permutation(8192).astype(np.int32), 512, 3e-3) for _ in range(15)]
    say("  losses %s" % ["%.3e" % l for l in losses[::3]])
    assert losses[-1] < 0.1 * losses[0]


def _global(spec, L, E):
    """A per-node field (L part N x 3, E part N x 2) in the solver's vector layout: the
    L DOFs node by node, then the E DOFs of the interior crossings."""
    W, Lg = spec["width"], spec["length"]
    inter = [u + v * W for v in range(1, Lg - 1) for u in range(1, W - 1)]
    return np.concatenate([np.asarray(L).ravel(), np.asarray(E)[inter].ravel()])


def test_fused_guess_equals_features_and_forward(ys, cs, ws):
    """The one-launch warm start (features, network, correction on the device) gives the
    same bits as the features read back, the forward kernel on them, and
    x0 = x_q + out_std * y formed on the host, x_q = 3 v_n - 3 v_{n-1} + v_{n-2} from three
    states (as the device forms it: v_n + (2 dv_n - dv_{n-1}))."""
    spec = cs.heter_spec(21, "twill", 41)
    spec["wind_vel"] = [0.0, 4.0, 0.0]
    c = ys.Cloth(spec)
    c.set_cg_variant("persistent")
    c.set_ClothState(cs.initial_state(spec))
    h = spec["h"]
    for _ in range(3):
        c.step(h, tol=1e-10)
    vs = []
    for _ in range(3):
        st = c.get_ClothState()
        vs.append(_global(spec, st["LVel"], st["EVel"]))
        if len(vs) < 3:
            c.step(h, tol=1e-10)
    v2, v1, v0 = vs
    x_q = v0 + (2.0 * (v0 - v1) - (v1 - v2))
    c.assemble(h)
    X = c.features()
    rng = np.random.default_rng(9)
    net = ws.WarmStartMLP(params=ws.init_params(3) + rng.normal(0, 0.05, 37893).astype(np.float32))
    net.set_norm(X.mean(0), X.std(0) + 1e-12, (np.abs(rng.normal(size=5)) * 1e-3 + 1e-4).astype(np.float32))
    Y = net.forward(X).astype(np.float64) * net.out_std.astype(np.float64)
    want = x_q + _global(spec, Y[:, :3], Y[:, 3:])
    c.warm_guess(net)
    got = c.vec("x")
    say("fused guess: max |diff| %.3e (|x0 - x_q| %.3e)" % (np.abs(got - want).max(), np.abs(want - x_q).max()))
    assert np.abs(want - x_q).max() > 0
    np.testing.assert_array_equal(got, want)
    net.close()
    c.close()


def test_warm_start_on_cloth(ys, cs, ws):
    """Train on a few seeded cloths, then compare on an unseen cloth, all solves to the
    same tolerance: the warm start needs fewer total CG iterations than the zero guess
    and than the linear and quadratic extrapolations (the latter is what it corrects),
    and every guess follows the zero guess's trajectory."""
    specs = []
    for i, W in enumerate((24, 32, 28)):
        s = cs.heter_spec(W, ("twill", "plain", "satin")[i], 300 + i)
        s["wind_vel"] = [0.0, 3.0 + 2 * i, 0.0]
        specs.append(s)
    say("warm start: generating data")
    X, T, M = ws.generate(specs, steps=40, skip=3)
    net, losses = ws.fit(X, T, M, epochs=12, batch=2048, lr=2e-3)
    say("  train loss %.3e -> %.3e" % (losses[0], losses[-1]))
    spec = cs.heter_spec(30, "satin", 399)
    spec["wind_vel"] = [0.0, 4.0, 0.0]
    st = cs.initial_state(spec)
    names = ("zero", "network", "extrap", "quad")
    cl = {}
    for k in names:
        cl[k] = ys.Cloth(spec)
        cl[k].set_cg_variant("persistent")
        cl[k].set_ClothState(st)
    h = spec["h"]
    tot = dict.fromkeys(names, 0)
    for s in range(30):
        it = {}
        i1, r1, rc = cl["zero"].step(h, tol=1e-10, maxit=20000)
        assert rc == 0
        it["zero"] = i1
        i2, r2, rc = cl["network"].step_warm(net, h, tol=1e-10, maxit=20000)
        assert rc == 0 and r2 < 1e-10
        it["network"] = i2
        for g in ("extrap", "quad"):
            it[g], rc = cl[g].run_guess(g, h, 1, tol=1e-10, maxit=20000)
            assert rc == 0
        if s >= 3:
            for k in names:
                tot[k] += it[k]
    sa = cl["zero"].get_ClothState()
    err = 0.0
    for g in ("network", "extrap", "quad"):   # every guess follows the zero guess's trajectory
        sb = cl[g].get_ClothState()
        err = max(err, max(np.linalg.norm(sa[k] - sb[k]) / max(np.linalg.norm(sa[k]), 1e-300) for k in ("LPos", "LVel")))
    say("  total iterations %s; state rel diff %.2e" % (tot, err))
    assert tot["network"] < tot["zero"]
    assert tot["network"] < tot["extrap"]
    assert tot["network"] < tot["quad"]
    assert err < 1e-6
    net.close()
    for c in cl.values():
        c.close()


def test_warm_step_in_one_launch_equals_separate_kernels(ys, cs, ws):
    """With the block-Jacobi preconditioner a warm step factors the rows in the network's
    launch (k_mlp_warm_bj: four groups per block, blocks grouped by XCD).  The step equals,
    bit for bit, the same step made of the separate kernels: the assembly with its rows from
    k_bj_factor, the guess of k_mlp_warm, the solve from that guess, the update.  The CG
    iterates depend on every bit of the rows, so this pins them; a cloth with partial tiles,
    odd sides and wind."""
    spec = cs.heter_spec(21, "satin", 5)
    spec["length"] = 19
    spec["wind_vel"] = [0.5, 3.0, 0.0]
    c0 = ys.Cloth(spec)
    c0.set_ClothState(cs.initial_state(spec))
    c0.run(spec["h"], 3, tol=1e-10)
    st = c0.get_ClothState()
    c0.close()
    h = spec["h"]
    rng = np.random.default_rng(4)
    net = ws.WarmStartMLP(params=ws.init_params(3) + rng.normal(0, 0.05, 37893).astype(np.float32))
    net.set_norm(np.zeros(32, np.float32), np.ones(32, np.float32), np.full(5, 1e-4, np.float32))
    a, b = ys.Cloth(spec), ys.Cloth(spec)
    for c in (a, b):
        c.set_cg_variant("persistent")
        c.set_preconditioner("block2")
        c.set_ClothState(st)
    ia, ra, rc = a.step_warm(net, h, tol=1e-10)
    assert rc == 0
    b.assemble(h)
    rows = b.precond_rows()
    b.warm_guess(net)
    guess = b.vec("x")
    ib, rb, ok = b.solve(tol=1e-10, x0=guess)
    b.update(h)
    say("one launch: %d iterations, separate kernels: %d; rows nonzero %d" % (ia, ib, np.count_nonzero(rows)))
    assert ok and ia == ib and np.count_nonzero(rows) > 0
    sa, sb = a.get_ClothState(), b.get_ClothState()
    for k in sa:
        np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)
    for c in (a, b):
        c.close()
    net.close()
