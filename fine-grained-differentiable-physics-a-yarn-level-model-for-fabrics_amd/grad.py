"""Gradients of yarn-level rollouts with respect to the material parameters.

The reference learns per-yarn density, stretch and bend moduli and the cloth's shear
stiffness and friction by backpropagating a position loss through SimStep
(src/TrainHeter.h:81-146, loss src/TrainHeter.h:211-226, autograd through
EigenSolver::apply, src/Solvers.h:35-52).  Here the step's tangents are propagated
forward on the GPU (Cloth.step_tangent, one direction per parameter), so the Jacobian of
the whole trajectory with respect to the parameters is available after the forward
rollout and the backward pass is a contraction with it.  The autograd Function
(make_autograd_function) uses the reverse mode by default: the forward pass records the
rollout, the backward pass gives every parameter's derivative with one backward solve per
step (Cloth.adjoint_*).

Two gradients are offered: "ref", the reference's own (M, K_v, A and the non-jamming
shear angle held constant, backward solve with A instead of A^T), and "exact", the
derivative of the step as computed.
"""
import copy

import numpy as np

from .cloth import Cloth, configure_solver

PARAMS = Cloth.PARAMS


def get_param(spec, name):
    if name[:-1] in ("rho", "Y", "B") and name[-1] in "01":
        return spec["yarns"][int(name[-1])][name[:-1]]
    return spec[name]


def with_params(spec, names, values):
    """Copy of spec with the named parameters set."""
    s = copy.deepcopy(spec)
    for n, v in zip(names, values):
        if n[:-1] in ("rho", "Y", "B") and n[-1] in "01":
            s["yarns"][int(n[-1])][n[:-1]] = float(v)
        else:
            s[n] = float(v)
    return s


def rollout_jacobian(spec, state, steps, names, mode="ref", tol=1e-10, maxit=None, ext_force=None, cg="persistent",
                     precond="block2"):
    """Runs `steps` SimSteps from `state` and returns (positions, jac): positions
    (steps, N, 3) after each step, jac (len(names), steps, N, 3) their derivatives with
    respect to the named parameters."""
    if len(spec["yarns"]) == 1 and any(n in ("rho1", "Y1", "B1") for n in names):
        raise ValueError("a one-yarn cloth has no second yarn parameters")
    c = Cloth(spec)
    try:
        configure_solver(c, cg, precond)
        c.set_ClothState(state)
        c.set_ext_force(ext_force)
        dp = np.zeros((len(names), len(PARAMS)))
        for i, n in enumerate(names):
            dp[i, PARAMS.index(n)] = 1.0
        c.tangent_begin(dp)
        N = c.N
        pos = np.zeros((steps, N, 3))
        jac = np.zeros((len(names), steps, N, 3))
        for t in range(steps):
            c.step_tangent(spec["h"], mode, tol=tol, maxit=maxit)
            pos[t] = c.get_ClothState()["LPos"]
            for i in range(len(names)):
                jac[i, t] = c.tangent_get(i)["LPos"]
        c.tangent_end()
    finally:
        c.close()
    return pos, jac


def position_loss(pos, target):
    """Mean squared position error over steps and nodes (src/TrainHeter.h:211-226)."""
    return float(np.mean((pos - target) ** 2))


def loss_and_grad(spec, state, steps, names, target, mode="ref", **kw):
    pos, jac = rollout_jacobian(spec, state, steps, names, mode=mode, **kw)
    r = pos - target
    loss = float(np.mean(r ** 2))
    g = np.array([2.0 * np.sum(r * jac[i]) / r.size for i in range(len(names))])
    return loss, g


def make_autograd_function(form="adjoint"):
    """torch.autograd.Function over the rollout: forward(theta, spec, state, steps, names,
    mode) -> positions (steps, N, 3) as float64 on theta's device; backward returns the
    gradient with theta's device and dtype (EigenSolver::apply's role in the reference,
    src/Solvers.h:35-52, for the whole rollout).
    form "adjoint": the forward pass records the rollout on the GPU and the backward pass
    runs the reverse mode (one backward solve per step, every parameter at once), with the
    upstream gradient as the loss gradient of each step's positions.
    form "tangent": the forward pass propagates one tangent per parameter and the backward
    pass contracts the upstream gradient with that Jacobian."""
    import torch

    class YarnRollout(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, spec, state, steps, names, mode):
            vals = theta.detach().cpu().numpy().astype(np.float64)
            ctx.theta_device, ctx.theta_dtype = theta.device, theta.dtype
            if form == "tangent":
                pos, jac = rollout_jacobian(with_params(spec, names, vals), state, steps, names, mode=mode)
                ctx.save_for_backward(torch.from_numpy(jac).to(theta.device))
                return torch.from_numpy(pos).to(theta.device)
            c = Cloth(with_params(spec, names, vals))
            configure_solver(c)
            c.set_ClothState(state)
            c.adjoint_begin(steps)
            pos = np.zeros((steps, c.N, 3))
            for t in range(steps):
                c.adjoint_step(spec["h"], tol=1e-10)
                pos[t] = c.get_ClothState()["LPos"]
            ctx.cloth, ctx.names, ctx.mode, ctx.h = c, list(names), mode, spec["h"]
            return torch.from_numpy(pos).to(theta.device)

        @staticmethod
        def backward(ctx, g):
            if form == "tangent":
                (jac,) = ctx.saved_tensors
                grad = (jac * g.to(jac.dtype).unsqueeze(0)).reshape(jac.shape[0], -1).sum(dim=1)
            else:
                c = ctx.cloth
                gn = g.detach().to(dtype=torch.float64).cpu().numpy()
                try:
                    res = c.adjoint_backward(ctx.h, ctx.mode, gstates=[{"LPos": gn[t]} for t in range(gn.shape[0])], tol=1e-10)
                finally:
                    c.adjoint_end()
                    c.close()
                grad = torch.from_numpy(np.array([res["param"][PARAMS.index(n)] for n in ctx.names]))
            return grad.to(device=ctx.theta_device, dtype=ctx.theta_dtype), None, None, None, None, None

    return YarnRollout
