"""The control path on the GPU: ControlNet steps against the oracle under the same
handle forces, and the gradient of the driver's mass-centre loss with respect to the
twelve lift / drag strengths against central finite differences of oracle rollouts."""
import sys

import numpy as np
import pytest

from oracle.oracle import OracleCloth

pytestmark = pytest.mark.gpu

KEYS = ("LPos", "LPosFix", "LVel", "EPos", "EPosBar", "EVel")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _net(ys, scale=0.02, W=7):
    # the reference's strengths accelerate the corner nodes of a 7x7 test cloth by
    # hundreds of m/s^2; scaled down the rollout stays in the regime of its data
    spec = ys.cloth_spec.homo_spec(W)
    return ys.control.ControlNet(spec, strengths=np.array(ys.control.REF_STRENGTHS) * scale, h=1e-3)


def force_index(smp):
    return 6 if smp > 5 else smp    # the driver's schedule, force_stop_frame = 5


def _oracle_roll(net, state, strengths, batch, sample):
    r = OracleCloth(net.spec)
    r.set_state(state)
    for _ in range(batch):
        for smp in range(sample - 1):
            st, _, _ = r.step(net.h, net.ext_force(force_index(smp), strengths), tol=1e-14, maxit=20000)
            assert st == 0
    return r.get_state()


def test_control_forward_matches_oracle(ys, cs):
    net = _net(ys)
    state = cs.initial_state(net.spec)
    net.reset_cloth()
    r = OracleCloth(net.spec)
    r.set_state(state)
    s = state
    worst = 0.0
    for smp in range(8):
        idx = ys.control.force_index(smp)
        s, ok = net.forward(idx, s if smp == 0 else None, tol=1e-14, maxit=20000)
        assert ok
        st, _, _ = r.step(net.h, net.ext_force(idx), tol=1e-14, maxit=20000)
        assert st == 0
        o = r.get_state()
        for k in KEYS:
            scale = max(np.abs(o[k]).max(), 1e-30)
            worst = max(worst, np.abs(s[k] - o[k]).max() / scale)
    net.close()
    say("  8 control steps, worst relative state difference %.2e" % worst)
    assert worst < 1e-9


@pytest.mark.parametrize("batch", [1, 2])
def test_control_gradient_matches_oracle_fd(ys, cs, batch):
    ctl = ys.control
    net = _net(ys)
    state = cs.initial_state(net.spec)
    sample = 8
    # goal: the end state of a rollout under other strengths
    goal = _oracle_roll(net, state, net.strengths * 1.7, batch, sample)["LPos"]
    loss, g, ok = ctl.rollout_loss_and_grad(net, state, goal, batch=batch, sample=sample, mode="exact", tol=1e-14,
                                            maxit=20000)
    assert ok == batch * (sample - 1)
    base = _oracle_roll(net, state, net.strengths, batch, sample)["LPos"]
    assert loss == pytest.approx(ctl.mass_centre_loss(base, goal), rel=1e-9)
    s0 = net.strengths.ravel()
    fd = np.zeros(12)
    for j in range(12):
        def cd(e):
            up, dn = s0.copy(), s0.copy()
            up[j] += e
            dn[j] -= e
            lu = ctl.mass_centre_loss(_oracle_roll(net, state, up, batch, sample)["LPos"], goal)
            ld = ctl.mass_centre_loss(_oracle_roll(net, state, dn, batch, sample)["LPos"], goal)
            return (lu - ld) / (2 * e)
        # the positions are about 5 m (the hanging height), so their rounding limits
        # the difference quotient: at 1e-2 of each strength, Richardson-extrapolated,
        # it is good to about 2e-7
        e = 1e-2 * abs(s0[j])
        fd[j] = (4.0 * cd(e) - cd(2 * e)) / 3.0
    err = np.abs(g - fd)
    for j in range(12):
        say("  batch %d  d/d%-17s fd % .6e  tangent % .6e  |diff| %.2e" % (batch, ctl.NAMES[j], fd[j], g[j], err[j]))
    assert np.all(err <= 1e-6 * np.abs(fd) + 1e-9 * np.abs(fd).max())


def test_control_ref_mode_gradient_is_close_to_exact(ys, cs):
    """The reference's gradient (A held constant, backward solve with A) and the exact
    one differ only through the Jacobians of the state path: over a short rollout they
    agree to a few percent."""
    ctl = ys.control
    net = _net(ys)
    state = cs.initial_state(net.spec)
    goal = _oracle_roll(net, state, net.strengths * 1.7, 1, 8)["LPos"]
    _, ge, _ = ctl.rollout_loss_and_grad(net, state, goal, sample=8, mode="exact")
    _, gr, _ = ctl.rollout_loss_and_grad(net, state, goal, sample=8, mode="ref")
    rel = np.linalg.norm(gr - ge) / np.linalg.norm(ge)
    say("  ref vs exact control gradient: relative difference %.3e" % rel)
    assert np.isfinite(gr).all() and rel < 0.1


def test_control_learner_step_lowers_the_loss(ys, cs):
    ctl = ys.control
    net = _net(ys)
    state = cs.initial_state(net.spec)
    goal = _oracle_roll(net, state, net.strengths * 1.7, 1, 8)["LPos"]
    loss0, g, _ = ctl.rollout_loss_and_grad(net, state, goal, sample=8)
    # a step of 2 % of the strengths' size along -g
    lr = 0.02 * np.linalg.norm(net.strengths) / np.linalg.norm(g) / 12
    L = ctl.ControlLearner(net, lr=lr)
    L.step(state, goal, sample=8)
    loss1, _, _ = ctl.rollout_loss_and_grad(net, state, goal, sample=8)
    say("  control loss %.6e -> %.6e" % (loss0, loss1))
    assert L.history[0]["loss"] == pytest.approx(loss0, rel=1e-12)
    assert loss1 < loss0
