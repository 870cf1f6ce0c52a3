"""Forward-mode tangents of SimStep against central finite differences of the oracle.

The exact-mode tangent of a step is the derivative of the step as computed, so the
oracle's step, differenced along the same parameter or state direction, pins it.  The
start is a natural state (5 steps from rest under gravity and wind): from a randomly
excited state the reference's system is indefinite and the step itself is unstable.
Tolerance: |tangent - fd| <= 1e-6 |fd| + 1e-9 max|fd| over the whole new state (central
differences at 1e-5 of each parameter's size, Richardson-extrapolated central differences
for the state direction, CG at 1e-14 on both sides: about 1e-8 of relative noise; the
absolute floor covers parameters whose effect over one step is below the CG noise, such
as mu and kc).
"""
import copy
import sys

import numpy as np
import pytest

from oracle.oracle import OracleCloth

pytestmark = pytest.mark.gpu

KEYS = ("LPos", "LPosFix", "LVel", "EPos", "EPosBar", "EVel")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def flat(st):
    return np.concatenate([np.asarray(st[k]).ravel() for k in KEYS])


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / max(nb, 1e-300)


def spec_with(spec, k, delta):
    """Copy of spec with parameter k (Cloth.PARAMS order) moved by delta."""
    s = copy.deepcopy(spec)
    name = ("rho0", "rho1", "Y0", "Y1", "B0", "B1", "S", "mu", "kc", "kf", "df", "handle_k")[k]
    if name[:-1] in ("rho", "Y", "B") and name[-1] in "01":
        s["yarns"][int(name[-1])][name[:-1]] += delta
    else:
        s[name] += delta
    return s


def param_value(spec, k):
    name = ("rho0", "rho1", "Y0", "Y1", "B0", "B1", "S", "mu", "kc", "kf", "df", "handle_k")[k]
    if name[:-1] in ("rho", "Y", "B") and name[-1] in "01":
        return spec["yarns"][int(name[-1])][name[:-1]]
    return spec[name]


def oracle_rollout(spec, state, steps, h):
    r = OracleCloth(spec)
    r.set_state(state)
    for _ in range(steps):
        st, it, rr = r.step(h, None, tol=1e-14, maxit=20000)
        assert st == 0
    return flat(r.get_state())


def random_state_dir(spec, N, seed):
    rng = np.random.default_rng(seed)
    L = spec["L"]
    d = {"LPos": rng.normal(0, 1e-2 * L, (N, 3)), "LVel": rng.normal(0, 1e-3, (N, 3)),
         "EPos": rng.normal(0, 1e-2 * L, (N, 2)), "EVel": rng.normal(0, 1e-3, (N, 2))}
    d["LPosFix"] = d["LPos"].copy()
    d["EPosBar"] = d["EPos"] + rng.normal(0, 1e-3 * L, (N, 2))
    return d


def add_dir(state, d, eps):
    return {k: np.asarray(state[k]) + eps * np.asarray(d[k]) for k in KEYS}


def _spec(cs):
    s = cs.heter_spec(7, "twill", 71)
    s["wind_vel"] = [0.4, 3.0, -0.5]
    s["handles"] = [(0, 0), (6, 0), (3, 2)]
    return s


def natural_state(cs, spec, steps=5):
    r = OracleCloth(spec)
    r.set_state(cs.initial_state(spec))
    for _ in range(steps):
        st, _, _ = r.step(spec["h"], None, tol=1e-14, maxit=20000)
        assert st == 0
    return r.get_state()


@pytest.mark.parametrize("steps", [1, 6])
def test_exact_tangent_matches_fd(ys, cs, steps):
    spec = _spec(cs)
    h = spec["h"]
    state = natural_state(cs, spec)
    g = ys.Cloth(spec)
    N = g.N
    nparam = len(g.PARAMS)
    scale = np.array([abs(param_value(spec, k)) for k in range(nparam)])
    # 12 parameter directions (one parameter each, scaled to its size) + 1 state direction
    dparams = np.zeros((nparam + 1, nparam))
    dparams[:nparam] = np.diag(scale)
    sdir = random_state_dir(spec, N, 5)
    zero = {k: np.zeros_like(np.asarray(state[k])) for k in KEYS}
    g.set_ClothState(state)
    g.tangent_begin(dparams, [zero] * nparam + [sdir])
    for _ in range(steps):
        it, rr, rc = g.step_tangent(h, "exact", tol=1e-14, maxit=20000)
    # Over several steps the solves' rounding (about 5e-12 m in the positions) divided by
    # the perturbation leaves about 2e-7 of absolute noise in every difference quotient,
    # so the multi-step check differences at 2e-5 and 4e-5 with Richardson extrapolation,
    # allows 2e-5 relative and an absolute floor of 1e-5 of the largest derivative.
    eps = 1e-5 if steps == 1 else 2e-5
    tol = 1e-6 if steps == 1 else 2e-5
    floor = 1e-9 if steps == 1 else 1e-5
    res = []
    for d in range(nparam + 1):
        tan = flat(g.tangent_get(d))
        if d < nparam:
            def cdp(e):
                up = oracle_rollout(spec_with(spec, d, e * scale[d]), state, steps, h)
                dn = oracle_rollout(spec_with(spec, d, -e * scale[d]), state, steps, h)
                return (up - dn) / (2 * e)
            fd = cdp(eps) if steps == 1 else (4.0 * cdp(eps) - cdp(2 * eps)) / 3.0
            name = g.PARAMS[d]
        else:
            # the state direction is curved enough for the second-order error of a plain
            # central difference to show (about 2e-5 at 1e-3): Richardson-extrapolate it
            def cd(es):
                up = oracle_rollout(spec, add_dir(state, sdir, es), steps, h)
                dn = oracle_rollout(spec, add_dir(state, sdir, -es), steps, h)
                return (up - dn) / (2 * es)
            fd = (4.0 * cd(5e-4) - cd(1e-3)) / 3.0
            name = "state"
        res.append((name, tan, fd))
    g.tangent_end()
    g.close()
    big = max(np.linalg.norm(fd) for _, _, fd in res)
    for name, tan, fd in res:
        err = np.linalg.norm(tan - fd)
        say("  %d steps, d/d%-8s |fd| %.3e |tan-fd| %.2e rel %.2e" % (steps, name, np.linalg.norm(fd), err,
                                                                  err / max(np.linalg.norm(fd), 1e-300)))
        assert err <= tol * np.linalg.norm(fd) + floor * big, name


def _glob_vel(spec, st):
    """Global velocity vector [L: 3k+c] ++ [E: 3N + 2e + c] of a state dict."""
    W, Lg = spec["width"], spec["length"]
    N = W * Lg
    Nc = (W - 2) * (Lg - 2)
    out = np.zeros(3 * N + 2 * Nc)
    for v in range(Lg):
        for u in range(W):
            k = u + v * W
            s = k   # state index u + v*width
            out[3 * k:3 * k + 3] = st["LVel"][s]
            if 0 < u < W - 1 and 0 < v < Lg - 1:
                e = (u - 1) + (v - 1) * (W - 2)
                out[3 * N + 2 * e:3 * N + 2 * e + 2] = st["EVel"][s]
    return out


def _ref_tangent(cs, spec, state, dparam, dstate, h, eps=1e-6):
    """The reference's one-step tangent built from the oracle: b with M, K_v and the
    non-jamming shear angle held at the base state (differences of F), the solve with
    A^T, and the update rule differentiated by hand."""
    W, Lg = spec["width"], spec["length"]
    N = W * Lg
    base = OracleCloth(spec)
    base.set_state(state)
    base.assemble(h)
    A = base.dense(0)
    M = base.csr(1).toarray()
    Kv = base.csr(3).toarray()
    base.freeze_shear_angle(True)
    pb = base.frozen_shear_angle()
    is_slide = base.crossing()[0].reshape(-1, 2)
    xsol, _, _, st = base.solve(tol=1e-14, maxit=20000)
    assert st == 0

    def probe(sgn):
        s = spec
        for k in range(len(dparam)):
            if dparam[k] != 0.0:
                s = spec_with(s, k, sgn * eps * dparam[k])
        r = OracleCloth(s)
        r.set_state(add_dir(state, dstate, sgn * eps))
        r.set_frozen_shear_angle(pb)
        r.assemble(h)
        return r.vec("F"), r.crossing()[0].reshape(-1, 2)

    Fp, sp = probe(1.0)
    Fm, sm = probe(-1.0)
    dF = (Fp - Fm) / (2 * eps)
    dis = (sp - sm) / (2 * eps)
    dv = _glob_vel(spec, dstate)
    db = h * (dF - Kv.T @ dv) + M @ dv
    dsol = np.linalg.solve(A.T, db)
    out = {k: np.array(dstate[k], dtype=float, copy=True) for k in KEYS}
    handles = set(tuple(x) for x in spec["handles"])
    for v in range(Lg):
        for u in range(W):
            k = u + v * W
            s = u + v * Lg
            dl = dsol[3 * k:3 * k + 3]
            out["LVel"][s] = dl
            out["LPos"][s] = dstate["LPos"][s] + h * dl
            if (u, v) not in handles:
                out["LPosFix"][s] = out["LPos"][s]
            if 0 < u < W - 1 and 0 < v < Lg - 1:
                e = (u - 1) + (v - 1) * (W - 2)
                de = dsol[3 * N + 2 * e:3 * N + 2 * e + 2]
                xe = xsol[3 * N + 2 * e:3 * N + 2 * e + 2]
                out["EVel"][s] = de
                out["EPos"][s] = dstate["EPos"][s] + h * de
                out["EPosBar"][s] = dstate["EPosBar"][s] + h * (dis[e] * xe + is_slide[e] * de)
    return flat(out)


def test_ref_tangent_matches_oracle_construction(ys, cs):
    """Reference-mode tangent: the reference's detached gradient (A, M, K_v constant,
    non-jamming shear angle constant, backward solve with A) as a one-step tangent map,
    against its construction from the oracle's F differences, M, K_v and a dense A^T
    solve."""
    spec = _spec(cs)
    h = spec["h"]
    state = natural_state(cs, spec)
    g = ys.Cloth(spec)
    N = g.N
    nparam = len(g.PARAMS)
    scale = np.array([abs(param_value(spec, k)) for k in range(nparam)])
    picks = [0, 3, 4, 6, 7, 11]
    dparams = np.zeros((len(picks) + 1, nparam))
    for i, k in enumerate(picks):
        dparams[i, k] = scale[k]
    zero = {k: np.zeros_like(np.asarray(state[k])) for k in KEYS}
    sdir = random_state_dir(spec, N, 9)
    dstates = [zero] * len(picks) + [sdir]
    g.set_ClothState(state)
    g.tangent_begin(dparams, dstates)
    g.step_tangent(h, "ref", tol=1e-14, maxit=20000)
    res = []
    for d in range(len(dparams)):
        tan = flat(g.tangent_get(d))
        # differences of F: 1e-6 along a parameter, 1e-4 along the state direction (where
        # rounding noise, about 1e-11 / eps relative, dominates below 1e-4)
        ref = _ref_tangent(cs, spec, state, dparams[d], dstates[d], h, eps=1e-6 if d < len(picks) else 1e-4)
        res.append((g.PARAMS[picks[d]] if d < len(picks) else "state", tan, ref))
    g.tangent_end()
    g.close()
    big = max(np.linalg.norm(r) for _, _, r in res)
    for name, tan, ref in res:
        err = np.linalg.norm(tan - ref)
        say("  ref mode d/d%-8s |ref| %.3e |tan-ref| %.2e" % (name, np.linalg.norm(ref), err))
        assert err <= 1e-6 * np.linalg.norm(ref) + 1e-9 * big, name


def test_ref_and_exact_modes_differ(ys, cs):
    """The reference's gradient is not the derivative of the step: over one step the two
    modes differ by far more than the tolerance (here in the density direction)."""
    spec = _spec(cs)
    state = natural_state(cs, spec)
    out = []
    for mode in ("exact", "ref"):
        g = ys.Cloth(spec)
        g.set_ClothState(state)
        dp = np.zeros((1, len(g.PARAMS)))
        dp[0, 0] = spec["yarns"][0]["rho"]
        g.tangent_begin(dp)
        g.step_tangent(spec["h"], mode, tol=1e-14, maxit=20000)
        out.append(flat(g.tangent_get(0)))
        g.close()
    d = rel(out[1], out[0])
    say("  |ref - exact| / |exact| = %.3e" % d)
    assert d > 1e-3


@pytest.mark.parametrize("form", ["adjoint", "tangent"])
def test_autograd_rollout_gradient_matches_fd(ys, cs, form):
    """The torch.autograd.Function over an 8-step rollout: the gradient of the position
    loss against a target trajectory (seeded ground-truth parameters, as the parameter-
    learning drivers use) matches the oracle's central difference of the same loss along
    the gradient direction, in exact mode."""
    import torch
    from importlib import import_module
    grad = import_module(ys.__name__ + ".grad")
    spec = _spec(cs)
    state = natural_state(cs, spec)
    names = ["rho0", "Y1", "B0", "S"]
    truth = np.array([grad.get_param(spec, n) for n in names])
    steps = 8
    target, _ = grad.rollout_jacobian(spec, state, steps, names, mode="exact", tol=1e-14, maxit=20000)
    theta0 = truth * np.array([1.3, 0.8, 1.2, 0.7])
    fn = grad.make_autograd_function(form)
    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    pos = fn.apply(theta, grad.with_params(spec, names, theta0), state, steps, names, "exact")
    loss = torch.mean((pos - torch.from_numpy(target)) ** 2)
    loss.backward()
    g = theta.grad.numpy()
    # central difference of the oracle's loss along g, in units of the parameters' sizes
    dirn = g * theta0 ** 2
    dirn = dirn / np.linalg.norm(dirn / theta0)

    def oracle_loss(vals):
        r = OracleCloth(grad.with_params(spec, names, vals))
        r.set_state(state)
        out = []
        for _ in range(steps):
            st, _, _ = r.step(spec["h"], None, tol=1e-14, maxit=20000)
            assert st == 0
            out.append(r.get_state()["LPos"])
        return np.mean((np.array(out) - target) ** 2)

    e = 1e-4
    fd = (oracle_loss(theta0 + e * dirn) - oracle_loss(theta0 - e * dirn)) / (2 * e)
    an = float(g @ dirn)
    say("  %s: loss %.4e  directional derivative: autograd %.8e  oracle fd %.8e  rel %.2e" %
        (form, float(loss.detach()), an, fd, abs(an - fd) / abs(fd)))
    assert abs(an - fd) <= 1e-5 * abs(fd)


def test_autograd_function_keeps_theta_device(ys, cs):
    """theta on the GPU in float32: the positions come back on its device and the
    gradient with its device and dtype, equal (to float32 rounding) to the gradient of a
    float64 CPU theta."""
    import torch
    from importlib import import_module
    grad = import_module(ys.__name__ + ".grad")
    spec = _spec(cs)
    state = natural_state(cs, spec)
    names = ["rho0", "S"]
    steps = 3
    vals = np.array([grad.get_param(spec, n) for n in names])
    target, _ = grad.rollout_jacobian(grad.with_params(spec, names, vals * 1.1), state, steps, names, mode="ref")
    fn = grad.make_autograd_function()
    out = {}
    for dev, dt in (("cpu", torch.float64), ("cuda", torch.float32)):
        theta = torch.tensor(vals, dtype=dt, device=dev, requires_grad=True)
        pos = fn.apply(theta, spec, state, steps, names, "ref")
        assert pos.device.type == dev
        loss = torch.mean((pos - torch.from_numpy(target).to(dev)) ** 2)
        loss.backward()
        assert theta.grad.device.type == dev and theta.grad.dtype == dt
        out[dev] = theta.grad.double().cpu().numpy()
    say("  gradient cpu %s cuda %s" % (out["cpu"], out["cuda"]))
    np.testing.assert_allclose(out["cuda"], out["cpu"], rtol=1e-5)


@pytest.mark.parametrize("steps", [1, 3])
def test_ext_force_tangent_matches_fd(ys, cs, steps):
    """The tangent of the trajectory along a direction of the external force (the path of
    the reference's control parameters, which enter SimStep through ext_force) against
    central finite differences of oracle rollouts under the perturbed force."""
    spec = _spec(cs)
    h = spec["h"]
    state = natural_state(cs, spec)
    g = ys.Cloth(spec)
    n = g.n
    f0 = cs.pressure_force(spec, state, 1.0)
    dext = np.random.default_rng(9).normal(0.0, 3e-7, n)
    g.set_ClothState(state)
    g.set_ext_force(f0)
    g.tangent_begin(np.zeros((1, len(g.PARAMS))))
    g.tangent_set_ext(dext[None, :])
    for _ in range(steps):
        g.step_tangent(h, "exact", tol=1e-14, maxit=20000)
    tan = flat(g.tangent_get(0))
    g.tangent_end()
    g.close()

    def roll(e):
        r = OracleCloth(spec)
        r.set_state(state)
        for _ in range(steps):
            st, _, _ = r.step(h, f0 + e * dext, tol=1e-14, maxit=20000)
            assert st == 0
        return flat(r.get_state())

    def cd(e):
        return (roll(e) - roll(-e)) / (2 * e)
    fd = (4.0 * cd(0.25) - cd(0.5)) / 3.0
    err = np.linalg.norm(tan - fd)
    say("  %d steps, d/dext |fd| %.3e |tan| %.3e |tan-fd| %.2e rel %.2e" % (steps, np.linalg.norm(fd), np.linalg.norm(tan), err,
                                                                       err / np.linalg.norm(fd)))
    assert err <= (1e-6 if steps == 1 else 2e-5) * np.linalg.norm(fd)


@pytest.mark.parametrize("mode", ["exact", "ref"])
def test_tangent_from_rest_pose_is_finite(ys, cs, mode):
    """From the constructor's rest pose every yarn is straight (bend angle 0, where acos
    has no derivative).  The tangents stay finite there, and over two steps the exact
    ones match central differences of the oracle.  The step is not smooth at straight
    yarns (the bend Jacobians are zero at angle 0 and tend to a finite limit, computed
    with 1/sin cancellations, as the angle goes to 0), so the difference quotients
    themselves scatter: between steps of 1e-5 and 2e-5 of each parameter they move by
    up to 6e-5 of their size, and the check allows 3e-4."""
    spec = cs.homo_spec(7)
    h = spec["h"]
    state = cs.initial_state(spec)
    g = ys.Cloth(spec)
    names = ("rho0", "Y0", "B0", "S")
    ks = [g.PARAMS.index(n) for n in names]
    scale = np.array([abs(param_value(spec, k)) for k in ks])
    dp = np.zeros((len(ks), len(g.PARAMS)))
    for i, k in enumerate(ks):
        dp[i, k] = scale[i]
    g.set_ClothState(state)
    g.tangent_begin(dp)
    for _ in range(2):
        g.step_tangent(h, mode, tol=1e-14, maxit=20000)
    tans = [flat(g.tangent_get(i)) for i in range(len(ks))]
    g.tangent_end()
    g.close()
    for i, k in enumerate(ks):
        assert np.isfinite(tans[i]).all() and np.linalg.norm(tans[i]) > 0, names[i]
        if mode != "exact":
            continue
        e = 1e-5
        fd = (oracle_rollout(spec_with(spec, k, e * scale[i]), state, 2, h)
              - oracle_rollout(spec_with(spec, k, -e * scale[i]), state, 2, h)) / (2 * e)
        err = np.linalg.norm(tans[i] - fd)
        say("  rest pose, 2 steps, d/d%-4s |fd| %.3e |tan-fd| %.2e rel %.2e" % (names[i], np.linalg.norm(fd), err,
                                                                          err / np.linalg.norm(fd)))
        assert err <= 3e-4 * np.linalg.norm(fd), names[i]
