"""GPU tests of the parameter learner (TrainHeterFull-style SGD on boxed parameters).

The loss gradient of the exact mode is checked against central finite differences of
the same loss computed with oracle rollouts on the CPU (the learner's tangents, the
box map and the loss assembly all enter), a gradient step must lower the loss, and the
batched ground-truth rollouts must equal single-cloth rollouts bit for bit.
"""
import sys

import numpy as np
import pytest

from oracle.oracle import OracleCloth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _problem(ys, cs, steps=4):
    truth = cs.heter_spec(6, "plain", 61)
    truth["wind_vel"] = [0.3, 4.0, -0.5]
    r = OracleCloth(truth)
    r.set_state(cs.initial_state(truth))
    for _ in range(3):
        st, _, _ = r.step(truth["h"], None, tol=1e-14, maxit=20000)
        assert st == 0
    state = r.get_state()
    frames = ys.learn.trajectory(truth, state, steps, tol=1e-12)
    start = ys.learn.perturbed_spec(truth, ys.learn.NAMES, seed=5)
    return truth, start, frames


def oracle_loss(spec, frames):
    """The reference's sample loss (src/TrainHeter.h:196-226) on an oracle rollout."""
    r = OracleCloth(spec)
    r.set_state({k: frames[k][0] for k in frames})
    N = frames["LPos"].shape[1]
    tot, ok = 0.0, 0
    for t in range(frames["LPos"].shape[0] - 1):
        st, _, _ = r.step(spec["h"], None, tol=1e-14, maxit=20000)
        assert st == 0
        s = r.get_state()
        dL = s["LPos"] - frames["LPos"][t + 1]
        dE = s["EPos"] - frames["EPos"][t + 1]
        tot += float(np.sum(dL * dL)) / N + float(np.sum(dE * dE)) / N
        ok += 1
    return tot / max(ok - 1, 1)


@pytest.mark.parametrize("form", ["tangent", "adjoint"])
def test_exact_loss_gradient_matches_oracle_fd(ys, cs, form):
    truth, start, frames = _problem(ys, cs)
    lrn = ys.learn.ParameterLearner(start, mode="exact", tol=1e-13, form=form)
    loss, g = lrn.loss_and_grad([frames])
    say("  loss %.6e (oracle %.6e)" % (loss, oracle_loss(lrn.current_spec(), frames)))
    assert loss == pytest.approx(oracle_loss(lrn.current_spec(), frames), rel=1e-6)
    eps = 1e-3
    gmax = np.abs(g).max()
    for i, n in enumerate(lrn.names):
        vals = []
        for sgn in (1, -1, 2, -2):
            th = lrn.theta.copy()
            th[i] += sgn * eps
            p = [ys.learn.constrain(t, lrn.ranges[m]) for t, m in zip(th, lrn.names)]
            vals.append(oracle_loss(ys.grad.with_params(start, lrn.names, p), frames))
        fd1 = (vals[0] - vals[1]) / (2 * eps)
        fd2 = (vals[2] - vals[3]) / (4 * eps)
        fd = (4 * fd1 - fd2) / 3   # Richardson
        say("  %s %-4s grad %+.6e fd %+.6e" % (form, n, g[i], fd))
        assert abs(g[i] - fd) <= 1e-4 * abs(fd) + 1e-6 * gmax


@pytest.mark.parametrize("mode", ["ref", "exact"])
def test_gradient_step_lowers_the_loss(ys, cs, mode):
    truth, start, frames = _problem(ys, cs)
    lrn = ys.learn.ParameterLearner(start, mode=mode, tol=1e-13)
    tg = ys.learn.ParameterLearner(start, mode=mode, tol=1e-13, form="tangent")
    # the reverse mode's gradient is the tangents' one
    la, ga = lrn.loss_and_grad([frames])
    lt, gt = tg.loss_and_grad([frames])
    say("  %s adjoint vs tangent: loss %.3e / %.3e, |dg| / |g| %.2e" % (mode, la, lt, np.linalg.norm(ga - gt) / np.linalg.norm(gt)))
    assert la == pytest.approx(lt, rel=1e-12)
    assert np.linalg.norm(ga - gt) <= 1e-8 * np.linalg.norm(gt)
    l0, g = lrn.loss_and_grad([frames])
    lr = 0.2 * l0 / float(g @ g)
    lrn.lrs = {n: lr for n in lrn.names}
    l0b, applied = lrn.step([frames])
    l1, _ = lrn.loss_and_grad([frames])
    say("  %s loss %.6e -> %.6e" % (mode, l0, l1))
    assert applied and l0b == l0
    assert l1 < l0


def test_batched_ground_truth_equals_single_rollouts(ys, cs):
    items = cs.batch_workload(n=3, seed=9, lo=6, hi=20)
    got = ys.learn.batch_trajectories(items, 3, tol=1e-9)
    for (spec, state), fr in zip(items, got):
        ref = ys.learn.trajectory(spec, state, 3, tol=1e-9)
        for k in ref:
            assert np.array_equal(fr[k], ref[k]), k
