"""Control of a cloth through forces at its handle nodes (YarnSimNetControl).

The reference's control module (src/YarnSimNet.h:87-246) holds twelve scalars: a lift
and a drag strength for each of six phases.  Phase p's external force is
lift_p * Lpos + drag_p * Dpos, where Lpos / Dpos are n-vectors (the step's global
layout) with a one at the z / y component of every handle node; the seventh force
index is the zero force.  Its driver (src/TrainControl.h:11-265) rolls the cloth
`sample - 1` steps per sample from the first frame of a stored trajectory, step smp
under force index smp while smp <= force_stop_frame (5) and under the zero force after,
chains `batch` samples (each sample restarts the phases from the state the previous one
ended in), and takes as loss the squared distance between the mass centres (mean node
positions) of the rollout's last state and the trajectory's last frame.  SGD with
learning rate 1e-4 updates the strengths.

Here the gradient comes from forward-mode tangents of the external force
(``Cloth.tangent_set_ext``): one direction per strength, whose force tangent is Lpos or
Dpos during the steps of its phase and zero otherwise, propagated with the rollout.
``mode`` selects the reference's gradient ("ref") or the derivative of the step as
computed ("exact"), as for the material parameters.
"""
import numpy as np

from . import _abi
from .cloth import Cloth, configure_solver

PHASES = 6
ZERO_FORCE = PHASES                    # force index of the zero external force
FORCE_STOP_FRAME = 5                   # src/TrainControl.h:70
LR = 1e-4                              # src/TrainControl.h:104
# (lift, drag) initial strengths of the six phases, src/TrainControl.h:52-68
REF_STRENGTHS = ((0.00161396, 0.00093096), (0.00166182, 0.00096396), (0.00104908, 0.00043926),
                 (0.00158723, 0.00178333), (0.00081228, 0.00132253), (0.00053757, 0.00172603))
REF_HANDLES = ((0, 0), (4, 0), (0, 4), (4, 4))   # src/TrainControl.h:45-48
# strength names in parameter order: lift_force_one, drag_force_one, ..., drag_force_six
NAMES = tuple("%s_force_%s" % (kind, num) for num in ("one", "two", "three", "four", "five", "six")
              for kind in ("lift", "drag"))


def control_spec(spec):
    """The control module's cloth (src/YarnSimNet.h:178-180): plain weave, upright pose,
    height 5 and no fixed nodes (the module never calls set_handles; its handle list
    only places the forces)."""
    s = dict(spec)
    s.update({"pattern": "plain", "pose": "upright", "hight": 5.0, "handles": []})
    return s


def force_positions(spec, handles=REF_HANDLES):
    """(Lpos, Dpos): ones at global index (first * width + second) * 3 + 2 (lift, z) and
    + 1 (drag, y) of every handle pair (src/YarnSimNet.h:139-146)."""
    W, Lg = int(spec["width"]), int(spec["length"])
    n = 3 * W * Lg + 2 * (W - 2) * (Lg - 2)
    lift, drag = np.zeros(n), np.zeros(n)
    for a, b in handles:
        k = a * W + b
        if not 0 <= k < W * Lg:
            raise ValueError("handle (%d, %d) is outside the %dx%d cloth" % (a, b, W, Lg))
        lift[3 * k + 2] = 1.0
        drag[3 * k + 1] = 1.0
    return lift, drag


def force_index(smp, force_stop_frame=FORCE_STOP_FRAME):
    """Force index of sample step smp (src/TrainControl.h:173-178)."""
    return smp if smp <= force_stop_frame else ZERO_FORCE


class ControlNet:
    """The reference's YarnSimNetControl over one GPU cloth.

    spec: the cloth (``control_spec`` is applied); strengths: (6, 2) lift / drag per
    phase; handles: the (first, second) pairs that place the forces; h: the step."""

    def __init__(self, spec, strengths=REF_STRENGTHS, handles=REF_HANDLES, h=1e-3):
        self.spec = control_spec(spec)
        self.strengths = np.array(strengths, dtype=np.float64).reshape(PHASES, 2)
        self.handles = tuple(tuple(p) for p in handles)
        self.h = float(h)
        self.lift, self.drag = force_positions(self.spec, self.handles)
        self.cloth = None

    @property
    def n(self):
        return self.lift.size

    def ext_force(self, idx, strengths=None):
        """External force of force index idx (0-5: the phases, 6: zero)."""
        s = self.strengths if strengths is None else np.asarray(strengths, dtype=np.float64).reshape(PHASES, 2)
        if idx == ZERO_FORCE:
            return np.zeros(self.n)
        if not 0 <= idx < PHASES:
            raise ValueError("invalid force index %d" % idx)
        return self.lift * s[idx, 0] + self.drag * s[idx, 1]

    def force_tangents(self, idx):
        """(12, n) tangents of force idx with respect to the twelve strengths."""
        d = np.zeros((2 * PHASES, self.n))
        if idx != ZERO_FORCE:
            d[2 * idx] = self.lift
            d[2 * idx + 1] = self.drag
        return d

    def reset_cloth(self):
        """A fresh cloth at the constructor's rest pose (src/YarnSimNet.h:165-184)."""
        self.close()
        self.cloth = Cloth(self.spec)

    def forward(self, idx, state=None, tol=1e-10, maxit=None):
        """One SimStep under force idx, from `state` when given (src/YarnSimNet.h:186-205).
        Returns (state, is_success); as SimStep, only a NaN/Inf fails the step
        (src/TrainHelper.cpp:65-90)."""
        if self.cloth is None:
            self.reset_cloth()
        if state is not None:
            self.cloth.set_ClothState(state)
        self.cloth.set_ext_force(self.ext_force(idx))
        _, _, rc = self.cloth.step(self.h, tol=tol, maxit=maxit)
        return self.cloth.get_ClothState(), rc != _abi.YC_ERR_NONFINITE

    def close(self):
        if self.cloth is not None:
            self.cloth.close()
            self.cloth = None


def mass_centre_loss(lpos, goal_lpos):
    """|mean(goal) - mean(lpos)|^2 (src/TrainControl.h:191-197)."""
    d = np.asarray(goal_lpos).reshape(-1, 3).mean(axis=0) - np.asarray(lpos).reshape(-1, 3).mean(axis=0)
    return float(d @ d)


def rollout_loss_and_grad(net, init_state, goal_lpos, batch=1, sample=8, mode="exact", tol=1e-12, maxit=None,
                          force_stop_frame=FORCE_STOP_FRAME, cg="persistent", precond="block2"):
    """The driver's loss for the current strengths and its gradient with respect to the
    twelve of them (src/TrainControl.h:120-201).  A failed step ends its sample, as the
    reference's break does, and the next sample goes on from the last good state.
    Returns (loss, grad (12,), successful steps)."""
    c = Cloth(net.spec)
    ok = 0
    try:
        configure_solver(c, cg, precond)
        c.set_ClothState(init_state)
        c.tangent_begin(np.zeros((2 * PHASES, len(Cloth.PARAMS))))
        for _ in range(batch):
            for smp in range(sample - 1):
                idx = force_index(smp, force_stop_frame)
                c.set_ext_force(net.ext_force(idx))
                c.tangent_set_ext(net.force_tangents(idx))
                try:
                    c.step_tangent(net.h, mode, tol=tol, maxit=maxit)
                except (_abi.NonFinite, _abi.NotConverged):
                    break
                ok += 1
        lpos = c.get_ClothState()["LPos"]
        d = np.asarray(goal_lpos).reshape(-1, 3).mean(axis=0) - lpos.mean(axis=0)
        grad = np.array([-2.0 * float(d @ c.tangent_get(j)["LPos"].mean(axis=0)) for j in range(2 * PHASES)])
        c.tangent_end()
    finally:
        c.close()
    return float(d @ d), grad, ok


class ControlLearner:
    """SGD on the twelve strengths, as the reference's control training.

    The reference calls the optimiser's step once per parameter whose value is finite
    (src/TrainControl.h:216-220), so with the gradient unchanged in between an update
    moves the strengths by (number of finite strengths) * lr * grad; `repeat_steps`
    keeps that (default) or applies one step."""

    def __init__(self, net, lr=LR, mode="exact", repeat_steps=True, tol=1e-12):
        self.net = net
        self.lr = float(lr)
        self.mode = mode
        self.repeat_steps = repeat_steps
        self.tol = tol
        self.history = []

    def update(self, grad):
        """Applies one optimiser update for `grad`; returns the number of SGD steps."""
        s = self.net.strengths.reshape(-1)
        nsteps = int(np.isfinite(s).sum()) if self.repeat_steps else 1
        self.net.strengths = (s - nsteps * self.lr * np.asarray(grad, dtype=np.float64)).reshape(PHASES, 2)
        return nsteps

    def step(self, init_state, goal_lpos, batch=1, sample=8):
        loss, g, ok = rollout_loss_and_grad(self.net, init_state, goal_lpos, batch, sample, self.mode, self.tol)
        self.update(g)
        self.history.append({"loss": loss, "grad": g, "steps_ok": ok, "strengths": self.net.strengths.copy()})
        return loss
