"""Reverse-mode gradients (yc_adjoint_*) against the forward-mode tangents.

The tangents are pinned elsewhere: the exact mode against central differences of the
oracle's step (tests/test_gpu_grad.py), the reference mode against the oracle-built
reference gradient.  The adjoint of a rollout is the transpose of the same tangent map, so
for a loss L = sum_t <G_t, s_t> (random weights G_t on every state array after every
step) the adjoint's dL/dtheta, dL/ds_0 . d and dL/dext . e must equal the tangents'
directional derivatives along the same directions.  Both sides solve to a 1e-14 relative
residual; the tolerance, 1e-8 relative to each derivative plus 1e-10 of the largest, is
the north star's gradient bar with room for the solves' rounding.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("LPos", "LPosFix", "LVel", "EPos", "EPosBar", "EVel")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _spec(cs, W=9, Lg=8, pattern="twill", seed=71):
    s = cs.heter_spec(W, pattern, seed)
    s["length"] = Lg
    s["wind_vel"] = [0.4, 3.0, -0.5]
    s["handles"] = [(0, 0), (W - 1, 0), (W // 2, 2)]
    return s


def _natural(ys, spec, steps=5):
    c = ys.Cloth(spec)
    c.set_ClothState(ys.cloth_spec.initial_state(spec))
    for _ in range(steps):
        c.step(spec["h"], tol=1e-14)
    st = c.get_ClothState()
    c.close()
    return st


def _dot(G, d):
    return sum(float(np.sum(np.asarray(G[k]) * np.asarray(d[k]))) for k in KEYS if k in G and k in d)


def _weights(rng, N, steps, scale):
    out = []
    for _ in range(steps):
        g = {k: rng.normal(0, 1.0, (N, 3)) for k in ("LPos", "LPosFix", "LVel")}
        g.update({k: rng.normal(0, 1.0, (N, 2)) for k in ("EPos", "EPosBar", "EVel")})
        for k in KEYS:
            g[k] = g[k] * scale[k]
        out.append(g)
    return out


def _state_dir(rng, spec, N):
    L = spec["L"]
    d = {"LPos": rng.normal(0, 1e-2 * L, (N, 3)), "LVel": rng.normal(0, 1e-3, (N, 3)),
         "EPos": rng.normal(0, 1e-2 * L, (N, 2)), "EVel": rng.normal(0, 1e-3, (N, 2))}
    d["LPosFix"] = d["LPos"] * 0.5
    d["EPosBar"] = d["EPos"] + rng.normal(0, 1e-3 * L, (N, 2))
    return d


@pytest.mark.parametrize("mode", ["ref", "exact"])
def test_adjoint_matches_tangents(ys, cs, mode):
    spec = _spec(cs)
    h = spec["h"]
    state = _natural(ys, spec)
    rng = np.random.default_rng(3)
    g = ys.Cloth(spec)
    N, n = g.N, g.n
    nparam = len(g.PARAMS)
    steps = 6
    ext = rng.normal(0, 2e-4, n)                 # a constant external force (N)
    edir = rng.normal(0, 1e-4, n)
    sdir = _state_dir(rng, spec, N)
    scale = {"LPos": 1.0 / spec["L"], "LPosFix": 1.0 / spec["L"], "LVel": 1.0, "EPos": 1.0 / spec["L"],
             "EPosBar": 1.0 / spec["L"], "EVel": 1.0}
    G = _weights(rng, N, steps, scale)
    psize = np.array([abs(p) if p != 0 else 1.0 for p in _param_values(g, spec)])

    # tangents: 12 parameter directions (scaled to each size), the state direction, the
    # external-force direction
    ndir = nparam + 2
    dparams = np.zeros((ndir, nparam))
    dparams[:nparam] = np.diag(psize)
    zero = {k: np.zeros_like(np.asarray(state[k])) for k in KEYS}
    g.set_ClothState(state)
    g.set_ext_force(ext)
    g.tangent_begin(dparams, [zero] * nparam + [sdir, zero])
    dext = np.zeros((ndir, n))
    dext[nparam + 1] = edir
    g.tangent_set_ext(dext)
    tan = np.zeros(ndir)
    for t in range(steps):
        g.step_tangent(h, mode, tol=1e-14, maxit=20000)
        for d in range(ndir):
            tan[d] += _dot(G[t], g.tangent_get(d))
    g.tangent_end()

    # adjoint of the same rollout
    g.set_ClothState(state)
    g.set_ext_force(ext)
    g.adjoint_begin(steps)
    for _ in range(steps):
        g.adjoint_step(h, tol=1e-14, maxit=20000)
    res = g.adjoint_backward(h, mode, gstates=G, tol=1e-14, maxit=20000, want_state=True, want_ext=True)
    g.adjoint_end()
    g.close()
    adj = np.zeros(ndir)
    adj[:nparam] = res["param"] * psize
    adj[nparam] = _dot(res["state0"], sdir)
    adj[nparam + 1] = float(np.sum(res["ext"] @ edir))
    big = np.max(np.abs(tan))
    names = list(g.PARAMS) + ["state", "ext"]
    for d in range(ndir):
        err = abs(adj[d] - tan[d])
        say("  %s d/d%-8s tangent % .10e adjoint % .10e rel %.2e" % (mode, names[d], tan[d], adj[d], err / max(abs(tan[d]), 1e-300)))
    say("  backward CG iterations %d" % res["cg_iters"])
    for d in range(ndir):
        assert abs(adj[d] - tan[d]) <= 1e-8 * abs(tan[d]) + 1e-10 * big, names[d]


def _param_values(g, spec):
    out = []
    for name in g.PARAMS:
        if name[:-1] in ("rho", "Y", "B") and name[-1] in "01":
            y = int(name[-1])
            out.append(spec["yarns"][min(y, len(spec["yarns"]) - 1)][name[:-1]])
        else:
            out.append(spec[name])
    return out


def test_adjoint_checkpoints_recompute_the_same_gradient(ys, cs):
    """Storing every state, every third and only the first give the same gradient: the
    recomputed steps repeat the recorded ones bit for bit."""
    spec = _spec(cs, 10, 10, "satin", 5)
    h = spec["h"]
    state = _natural(ys, spec)
    rng = np.random.default_rng(8)
    steps = 7
    g = ys.Cloth(spec)
    G = _weights(rng, g.N, steps, {k: 1.0 for k in KEYS})
    out = []
    for K in (1, 3, steps):
        g.set_ClothState(state)
        g.adjoint_begin(steps, K)
        for _ in range(steps):
            g.adjoint_step(h, tol=1e-12)
        r = g.adjoint_backward(h, "exact", gstates=G, tol=1e-12, want_state=True)
        g.adjoint_end()
        out.append(np.concatenate([r["param"], np.concatenate([np.ravel(r["state0"][k]) for k in KEYS])]))
        say("  K=%d |grad| %.6e" % (K, np.linalg.norm(out[-1])))
    g.close()
    for o in out[1:]:
        assert np.array_equal(o, out[0])


def test_adjoint_target_loss(ys, cs):
    """The device targets' loss and gradient equal the host-formed ones: the loss of step t
    is w (sum |x - X|^2 + sum |uv - U|^2), its state gradient 2 w (x - X), 2 w (uv - U)."""
    spec = _spec(cs, 8, 8, "plain", 12)
    h = spec["h"]
    state = _natural(ys, spec)
    steps = 5
    # targets: the rollout of the cloth with a heavier first yarn
    tspec = dict(spec)
    tspec["yarns"] = [dict(spec["yarns"][0], rho=spec["yarns"][0]["rho"] * 1.3), spec["yarns"][1]]
    t = ys.Cloth(tspec)
    t.set_ClothState(state)
    X, U = [], []
    for _ in range(steps):
        t.step(h, tol=1e-12)
        s = t.get_ClothState()
        X.append(s["LPos"])
        U.append(s["EPos"])
    t.close()
    X, U = np.array(X), np.array(U)
    g = ys.Cloth(spec)
    N = g.N
    w = 1.0 / N
    g.set_ClothState(state)
    g.adjoint_begin(steps)
    traj = []
    for _ in range(steps):
        g.adjoint_step(h, tol=1e-12)
        traj.append(g.get_ClothState())
    g.adjoint_set_targets(X, U, w)
    r1 = g.adjoint_backward(h, "ref", tol=1e-12)
    loss = sum(w * (np.sum((s["LPos"] - x) ** 2) + np.sum((s["EPos"] - u) ** 2)) for s, x, u in zip(traj, X, U))
    G = [{"LPos": 2 * w * (s["LPos"] - x), "EPos": 2 * w * (s["EPos"] - u)} for s, x, u in zip(traj, X, U)]
    r2 = g.adjoint_backward(h, "ref", gstates=G, tol=1e-12)
    after = g.get_ClothState()
    g.adjoint_end()
    g.close()
    say("  loss device %.12e host %.12e; |grad| %.6e vs %.6e" % (r1["loss"], loss, np.linalg.norm(r1["param"]),
                                                               np.linalg.norm(r2["param"])))
    assert abs(r1["loss"] - loss) <= 1e-12 * loss
    assert np.allclose(r1["param"], r2["param"], rtol=1e-12, atol=0)
    # the backward pass leaves the context at the last recorded state
    for k in KEYS:
        assert np.array_equal(after[k], traj[-1][k])


@pytest.mark.parametrize("mode,precond", [("ref", "jacobi"), ("exact", "block2")])
def test_batch_adjoint_equals_single_cloths(ys, cs, mode, precond):
    """Cloths of uneven size recorded and differentiated together (yc_batch_adjoint_*):
    each cloth's loss and parameter gradient equal its own single-cloth adjoint's."""
    from yarnstep.batch import ClothBatch
    sizes = [(8, 8, "plain", 21), (12, 10, "twill", 22), (6, 9, "satin", 23), (16, 16, "twill", 24)]
    steps = 5
    specs, states, targets = [], [], []
    for W, Lg, pat, seed in sizes:
        sp = _spec(cs, W, Lg, pat, seed)
        st = _natural(ys, sp, 3)
        tsp = dict(sp)
        tsp["yarns"] = [dict(sp["yarns"][0], rho=sp["yarns"][0]["rho"] * 1.2), sp["yarns"][1]]
        t = ys.Cloth(tsp)
        t.set_ClothState(st)
        X, U = [], []
        for _ in range(steps):
            t.step(sp["h"], tol=1e-12)
            s = t.get_ClothState()
            X.append(s["LPos"])
            U.append(s["EPos"])
        t.close()
        specs.append(sp)
        states.append(st)
        targets.append((np.array(X), np.array(U)))
    h = specs[0]["h"]
    single = []
    for sp, st, (X, U) in zip(specs, states, targets):
        c = ys.Cloth(sp)
        c.set_cg_variant("persistent")
        c.set_preconditioner(precond)
        c.set_ClothState(st)
        c.adjoint_begin(steps)
        c.adjoint_run(h, steps, tol=1e-12)
        c.adjoint_set_targets(X, U, 1.0 / c.N)
        r = c.adjoint_backward(h, mode, tol=1e-12)
        single.append((r["loss"], r["param"]))
        c.adjoint_end()
        c.close()
    cloths = []
    for sp, st, (X, U) in zip(specs, states, targets):
        c = ys.Cloth(sp)
        c.set_cg_variant("persistent")
        c.set_preconditioner(precond)
        c.set_ClothState(st)
        c.adjoint_begin(steps)
        cloths.append(c)
    b = ClothBatch(cloths)
    b.adjoint_run(h, steps, tol=1e-12)
    for c, (X, U) in zip(cloths, targets):
        c.adjoint_set_targets(X, U, 1.0 / c.N)
    dparam, loss, iters = b.adjoint_backward(h, mode, tol=1e-12)
    b.close()
    for c in cloths:
        c.adjoint_end()
        c.close()
    for i, (l1, g1) in enumerate(single):
        rel = np.max(np.abs(dparam[i] - g1)) / max(np.max(np.abs(g1)), 1e-300)
        say("  cloth %d (%dx%d) loss %.6e / %.6e, max rel gradient difference %.2e, backward iterations %d" %
            (i, sizes[i][0], sizes[i][1], loss[i], l1, rel, iters[i]))
        assert loss[i] == pytest.approx(l1, rel=1e-12)
        assert rel <= 1e-12
