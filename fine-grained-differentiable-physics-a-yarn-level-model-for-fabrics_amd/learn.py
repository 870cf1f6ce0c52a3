"""Material-parameter learning from trajectories (TrainHeterFull-style).

The reference learns rho, Y, B of two yarns and the cloth's S and mu by SGD on
unconstrained parameters theta, each mapped into a box by
p = (upper - lower) sigmoid(theta) + lower (src/TrainHelper.cpp:3-13), with the boxes
and per-parameter learning rates of src/TrainHeter.h:82-95, 125-143.  A sample is a
rollout from the first frame of a stored trajectory; its loss adds, per step, the mean
over nodes of |LPos - LPos_gt|^2 and of |EPos - EPos_gt|^2, and the batch loss divides
the sum by (successful steps - 1) (src/TrainHeter.h:196-226).  A step with a NaN/Inf
gradient is skipped (src/TrainHeter.h:246-256).

Here the rollouts run on the GPU.  The parameter gradient comes from the reverse mode
by default (``Cloth.adjoint_*``: the rollout is recorded, and one backward pass with one
backward solve per step gives every parameter's derivative), or from the step's
forward-mode tangents (``Cloth.step_tangent``: one direction per learned parameter,
propagated with the rollout, nothing stored).  ``mode="ref"`` gives the reference's gradient (A, M,
K_v and the non-jamming shear angle held constant, backward solve with A),
``mode="exact"`` the derivative of the step as computed.
"""
import math

import numpy as np

from . import _abi
from .cloth import Cloth, configure_solver
from .grad import get_param, with_params

# reference parameter names -> this build's (yarn 1/2 of the reference are yarns 0/1)
NAMES = ("rho0", "rho1", "Y0", "Y1", "B0", "B1", "S", "mu")
REF_NAMES = {"rho0": "rho1", "rho1": "rho2", "Y0": "Y1", "Y1": "Y2", "B0": "B1", "B1": "B2", "S": "S", "mu": "mu"}
# (upper, lower) boxes, src/TrainHeter.h:82-95
RANGES = {"rho0": (0.003, 0.001), "rho1": (0.003, 0.001), "Y0": (800000.0, 0.0), "Y1": (300000.0, 0.0),
          "B0": (0.00018, 0.00005), "B1": (0.00018, 0.00005), "S": (1200.0, 0.0), "mu": (1.0, 0.0)}
# SGD learning rates per parameter, src/TrainHeter.h:136-143
LRS = {"rho0": 0.1, "rho1": 1e-4, "Y0": 12e16, "Y1": 7e16, "B0": 0.219, "B1": 0.145, "S": 19e10, "mu": 14e9}


def constrain(theta, rng):
    """src/TrainHelper.cpp:3-7: (upper - lower) sigmoid(theta) + lower."""
    up, lo = rng
    return (up - lo) / (1.0 + math.exp(-theta)) + lo


def unconstrain(p, rng):
    """src/TrainHelper.cpp:9-13, the inverse of ``constrain``."""
    up, lo = rng
    return -math.log((up - lo) / (p - lo) - 1.0)


def dconstrain(theta, rng):
    """d constrain / d theta = (upper - lower) sigmoid(theta) (1 - sigmoid(theta))."""
    up, lo = rng
    s = 1.0 / (1.0 + math.exp(-theta))
    return (up - lo) * s * (1.0 - s)


def trajectory(spec, state, steps, tol=1e-10, cg="persistent"):
    """Ground-truth frames: the state and `steps` SimSteps from it, as arrays
    (steps + 1, N, 3|2) under the get_ClothState keys."""
    c = Cloth(spec)
    try:
        try:
            c.set_cg_variant(cg)
        except _abi.YcError:   # the persistent form does not fit this cloth
            c.set_cg_variant("fused")
        c.set_ClothState(state)
        frames = [c.get_ClothState()]
        for _ in range(steps):
            _, _, rc = c.step(spec["h"], tol=tol)
            if rc != 0:
                raise RuntimeError("ground-truth rollout failed with status %d" % rc)
            frames.append(c.get_ClothState())
    finally:
        c.close()
    return {k: np.stack([f[k] for f in frames]) for k in frames[0]}


def sample_loss_terms(spec, frames, names, mode="ref", tol=1e-10, cg="persistent", precond="block2"):
    """Rollout from frames[0] along the stored trajectory.  Returns (loss_sum, grad_sum,
    steps_ok): the summed per-step position losses, their derivatives with respect to
    the named (constrained) parameters, and the number of steps that succeeded; the
    rollout stops at the first non-finite step as the reference's does."""
    c = Cloth(spec)
    nsteps = frames["LPos"].shape[0] - 1
    loss = 0.0
    grad = np.zeros(len(names))
    ok = 0
    try:
        configure_solver(c, cg, precond)
        c.set_ClothState({k: frames[k][0] for k in frames})
        dp = np.zeros((len(names), len(Cloth.PARAMS)))
        for i, n in enumerate(names):
            dp[i, Cloth.PARAMS.index(n)] = 1.0
        c.tangent_begin(dp)
        N = c.N
        for t in range(nsteps):
            try:
                c.step_tangent(spec["h"], mode, tol=tol)
            except (_abi.NonFinite, _abi.NotConverged):
                break
            s = c.get_ClothState()
            dL = s["LPos"] - frames["LPos"][t + 1]
            dE = s["EPos"] - frames["EPos"][t + 1]
            if not (np.isfinite(dL).all() and np.isfinite(dE).all()):
                break
            ok += 1
            loss += float(np.sum(dL * dL)) / N + float(np.sum(dE * dE)) / N
            for i in range(len(names)):
                tg = c.tangent_get(i)
                grad[i] += 2.0 * (float(np.sum(dL * tg["LPos"])) + float(np.sum(dE * tg["EPos"]))) / N
        c.tangent_end()
    finally:
        c.close()
    return loss, grad, ok


def sample_loss_terms_adjoint(spec, frames, names, mode="ref", tol=1e-10, cg="persistent", precond="block2",
                              checkpoint_every=1):
    """``sample_loss_terms`` by the reverse mode: the rollout is recorded (a state every
    checkpoint_every steps), the position loss's targets sit on the device, and one
    backward pass with one backward solve per step returns the loss and the derivatives of
    all parameters at once, whatever their number."""
    c = Cloth(spec)
    nsteps = frames["LPos"].shape[0] - 1
    try:
        configure_solver(c, cg, precond)
        c.set_ClothState({k: frames[k][0] for k in frames})
        c.adjoint_begin(nsteps, checkpoint_every)
        ok = 0
        try:   # every step queued at once; on a failure, again step by step to find the last good one
            c.adjoint_run(spec["h"], nsteps, tol=tol)
            s = c.get_ClothState()
            if np.isfinite(s["LPos"]).all() and np.isfinite(s["EPos"]).all():
                ok = nsteps
        except (_abi.NonFinite, _abi.NotConverged):
            pass
        if ok < nsteps:
            c.set_ClothState({k: frames[k][0] for k in frames})
            c.adjoint_begin(nsteps, checkpoint_every)
        for _ in range(nsteps - ok):
            try:
                c.adjoint_step(spec["h"], tol=tol)
            except (_abi.NonFinite, _abi.NotConverged):
                break
            s = c.get_ClothState()
            if not (np.isfinite(s["LPos"]).all() and np.isfinite(s["EPos"]).all()):
                break
            ok += 1
        if ok == 0:
            c.adjoint_end()
            return 0.0, np.zeros(len(names)), 0
        c.adjoint_set_targets(frames["LPos"][1:], frames["EPos"][1:], 1.0 / c.N)
        res = c.adjoint_backward(spec["h"], mode, tol=tol)
        c.adjoint_end()
    finally:
        c.close()
    grad = np.array([res["param"][Cloth.PARAMS.index(n)] for n in names])
    return res["loss"], grad, ok


def batch_loss_terms_adjoint(specs, frames_list, names, mode="ref", tol=1e-10, precond="jacobi"):
    """``sample_loss_terms_adjoint`` for many cloths at once: the rollouts are recorded
    together and their backward passes run together (``ClothBatch.adjoint_run`` /
    ``adjoint_backward``: one launch per phase over all cloths, the solves in one
    launch).  Returns a list of (loss_sum, grad_sum, steps_ok) per cloth.  If
    a step of the batch fails (NaN/Inf, no convergence), every cloth is redone alone, so
    each stops at its own first failed step as the reference's rollouts do."""
    from .batch import ClothBatch
    nsteps = frames_list[0]["LPos"].shape[0] - 1
    if any(f["LPos"].shape[0] - 1 != nsteps for f in frames_list):
        raise ValueError("the batched rollouts need trajectories of one length")
    cloths, b = [], None
    try:
        for spec, f in zip(specs, frames_list):
            c = Cloth(spec)
            cloths.append(c)
            c.set_cg_variant("persistent")
            c.set_preconditioner(precond)
            c.set_ClothState({k: f[k][0] for k in f})
            c.adjoint_begin(nsteps)
        b = ClothBatch(cloths)
        try:
            b.adjoint_run(specs[0]["h"], nsteps, tol=tol)
        except (_abi.NonFinite, _abi.NotConverged):
            b.close()
            b = None
            return [sample_loss_terms_adjoint(sp, f, names, mode, tol, precond=precond) for sp, f in zip(specs, frames_list)]
        for c, f in zip(cloths, frames_list):
            c.adjoint_set_targets(f["LPos"][1:], f["EPos"][1:], 1.0 / c.N)
        dparam, loss, _ = b.adjoint_backward(specs[0]["h"], mode, tol=tol)
    finally:
        if b is not None:
            b.close()
        for c in cloths:
            c.close()
    idx = [Cloth.PARAMS.index(n) for n in names]
    return [(float(loss[i]), dparam[i, idx], nsteps) for i in range(len(cloths))]


class ParameterLearner:
    """SGD on the boxed parameters of one cloth, as the reference's training loop.

    spec: the cloth (its parameter values are the starting point); names: the learned
    parameters (default: all eight of the reference); lrs / ranges: per-parameter
    learning rates and (upper, lower) boxes, the reference's by default; form: "adjoint"
    (reverse mode, one backward solve per step for all parameters) or "tangent" (one
    forward tangent per learned parameter)."""

    def __init__(self, spec, names=NAMES, lrs=None, ranges=None, mode="ref", tol=1e-10, form="adjoint"):
        if form not in ("adjoint", "tangent"):
            raise ValueError("form must be 'adjoint' or 'tangent'")
        self.form = form
        if len(spec["yarns"]) == 1:
            names = tuple(n for n in names if n not in ("rho1", "Y1", "B1"))
        self.spec = spec
        self.names = tuple(names)
        self.lrs = dict(LRS, **(lrs or {}))
        self.ranges = dict(RANGES, **(ranges or {}))
        self.mode = mode
        self.tol = tol
        for n in self.names:
            up, lo = self.ranges[n]
            if not lo < get_param(spec, n) < up:
                raise ValueError("%s = %g is outside its box (%g, %g)" % (n, get_param(spec, n), lo, up))
        self.theta = np.array([unconstrain(get_param(spec, n), self.ranges[n]) for n in self.names])
        self.history = []

    def params(self):
        return {n: constrain(t, self.ranges[n]) for n, t in zip(self.names, self.theta)}

    def current_spec(self):
        p = self.params()
        return with_params(self.spec, self.names, [p[n] for n in self.names])

    def loss_and_grad(self, samples):
        """Batch loss (src/TrainHeter.h:161-226) over `samples` (trajectory dicts) and
        its gradient with respect to theta."""
        spec = self.current_spec()
        tot, g, ok = 0.0, np.zeros(len(self.names)), 0
        for frames in samples:
            fn = sample_loss_terms_adjoint if self.form == "adjoint" else sample_loss_terms
            l_, g_, ok_ = fn(spec, frames, self.names, self.mode, self.tol)
            tot += l_
            g += g_
            ok += ok_
        denom = (ok - 1) if ok - 1 > 0 else 1
        dtheta = np.array([dconstrain(t, self.ranges[n]) for n, t in zip(self.names, self.theta)])
        return tot / denom, g / denom * dtheta

    def step(self, samples):
        """One SGD update; a non-finite gradient skips it.  Returns (loss, applied)."""
        loss, g = self.loss_and_grad(samples)
        applied = bool(np.isfinite(g).all())
        if applied:
            self.theta = self.theta - np.array([self.lrs[n] for n in self.names]) * g
        self.history.append({"loss": loss, "params": self.params(), "applied": applied})
        return loss, applied


def batch_trajectories(items, steps, tol=1e-8):
    """Ground-truth frames of many cloths at once: the cloths (spec, state) are stepped
    together by a ``ClothBatch`` one step per run, and each cloth's state is read after
    every step.  Returns one trajectory dict per cloth (as ``trajectory``)."""
    from .batch import ClothBatch
    cloths = []
    try:
        for spec, state in items:
            c = Cloth(spec)
            c.set_cg_variant("persistent")
            c.set_ClothState(state)
            cloths.append(c)
        frames = [[c.get_ClothState()] for c in cloths]
        b = ClothBatch(cloths)
        try:
            for _ in range(steps):
                _, st, rc = b.run(items[0][0]["h"], 1, tol=tol)
                if rc != 0:
                    raise RuntimeError("ground-truth batch step failed: statuses %s" % list(st))
                for f, c in zip(frames, cloths):
                    f.append(c.get_ClothState())
        finally:
            b.close()
    finally:
        for c in cloths:
            c.close()
    return [{k: np.stack([x[k] for x in f]) for k in f[0]} for f in frames]


def perturbed_spec(spec, names, seed, lo=0.5, hi=2.0):
    """The spec with each named parameter scaled by U(lo, hi) (seeded), clipped into its
    box: the starting point of a learning run."""
    rng = np.random.default_rng(seed)
    vals = []
    for n in names:
        up, low = RANGES[n]
        v = get_param(spec, n) * rng.uniform(lo, hi)
        eps = 1e-6 * (up - low)
        vals.append(min(max(v, low + eps), up - eps))
    return with_params(spec, names, vals)
