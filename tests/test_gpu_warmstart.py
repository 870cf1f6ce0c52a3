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


def test_warm_start_on_cloth(ys, cs, ws):
    """Train on a few seeded cloths, then compare zero-guess and warm-started CG on an
    unseen cloth: same tolerance, fewer total iterations, same trajectory."""
    specs = []
    for i, W in enumerate((24, 32)):
        s = cs.heter_spec(W, ("twill", "plain")[i], 300 + i)
        s["wind_vel"] = [0.0, 3.0 + 2 * i, 0.0]
        specs.append(s)
    say("warm start: generating data")
    X, T, M = ws.generate(specs, steps=40, skip=3)
    net, losses = ws.fit(X, T, M, epochs=10, batch=2048, lr=2e-3)
    say("  train loss %.3e -> %.3e" % (losses[0], losses[-1]))
    spec = cs.heter_spec(28, "satin", 399)
    spec["wind_vel"] = [0.0, 4.0, 0.0]
    a, b = ys.Cloth(spec), ys.Cloth(spec)
    st = cs.initial_state(spec)
    a.set_ClothState(st)
    b.set_ClothState(st)
    h = spec["h"]
    ita = itb = 0
    for s in range(30):
        i1, r1, rc1 = a.step(h, tol=1e-10, maxit=20000)
        i2, r2, rc2 = b.step_warm(net, h, tol=1e-10, maxit=20000)
        assert rc1 == 0 and rc2 == 0 and r2 < 1e-10
        if s >= 3:
            ita += i1
            itb += i2
    sa, sb = a.get_ClothState(), b.get_ClothState()
    e = max(np.linalg.norm(sa[k] - sb[k]) / max(np.linalg.norm(sa[k]), 1e-300) for k in ("LPos", "LVel"))
    say("  iterations zero-guess %d warm %d, state rel diff %.2e" % (ita, itb, e))
    assert itb < ita
    assert e < 1e-6
