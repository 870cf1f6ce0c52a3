"""yc_mlp_forward_noalloc against yc_mlp_forward through the C ABI: two calls in a row
(the second with fewer rows, so stale staging contents would show), bit for bit, and
the row limit rejected with YC_ERR_ARG.  Also checked against the CPU checker of the
network (oracle/mlp_ref.py) at the forward test's tolerance."""
import ctypes
import sys

import numpy as np
import pytest

from oracle import mlp_ref

pytestmark = pytest.mark.gpu

_fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))  # noqa: E731


def say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _net(ws, seed):
    rng = np.random.default_rng(seed)
    p = ws.init_params(seed) + rng.normal(0, 0.02, 37893).astype(np.float32)
    net = ws.WarmStartMLP(params=p)
    mu = rng.normal(size=32).astype(np.float32)
    sd = (1.0 + rng.uniform(size=32)).astype(np.float32)
    net.set_norm(mu, sd, np.ones(5, np.float32))
    return net, p, mu, sd


def _call(L, fn, net, X):
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.full((X.shape[0], 5), np.nan, dtype=np.float32)
    rc = getattr(L, fn)(net._h, _fp(X), X.shape[0], _fp(Y))
    return rc, Y


def test_noalloc_forward_equals_forward(ys):
    import yarnstep.warmstart as ws
    L = ys.load_library()
    net, p, mu, sd = _net(ws, 5)
    rng = np.random.default_rng(6)
    for rows in (5000, 777):   # the second call is smaller than the first
        say("noalloc forward rows=%d" % rows)
        X = (rng.normal(size=(rows, 32)) * 2).astype(np.float32)
        rc_a, Ya = _call(L, "yc_mlp_forward", net, X)
        rc_b, Yb = _call(L, "yc_mlp_forward_noalloc", net, X)
        assert rc_a == 0 and rc_b == 0
        assert np.array_equal(Ya, Yb)
        Yr, _, _ = mlp_ref.forward(p, X, mu, sd)
        assert np.linalg.norm(Yb - Yr) / np.linalg.norm(Yr) < 2e-3
    X = np.zeros((65537, 32), dtype=np.float32)
    rc, _ = _call(L, "yc_mlp_forward_noalloc", net, X)
    assert rc == ys.YC_ERR_ARG
    rc, _ = _call(L, "yc_mlp_forward_noalloc", net, X[:0])
    assert rc == ys.YC_ERR_ARG
    net.close()
