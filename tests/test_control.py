"""Host logic of the control path (the reference's YarnSimNetControl and its driver):
force placement, the phase schedule, the force tangents and the optimiser update.
No GPU needed."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def ctl(ys):
    return ys.control


def test_force_positions_on_the_reference_handles(ys, ctl):
    spec = ctl.control_spec(ys.cloth_spec.homo_spec(5))
    lift, drag = ctl.force_positions(spec)
    assert lift.size == drag.size == 3 * 25 + 2 * 9
    # the four corners of a 5x5 cloth: nodes 0, 4, 20, 24, z for lift and y for drag
    assert sorted(np.flatnonzero(lift)) == [2, 14, 62, 74]
    assert sorted(np.flatnonzero(drag)) == [1, 13, 61, 73]
    assert lift.sum() == drag.sum() == 4.0


def test_force_positions_index_order_on_a_wide_cloth(ctl, ys):
    # (first, second) -> node first * width + second, as the module's index formula
    spec = ctl.control_spec(ys.cloth_spec.homo_spec(6))
    spec["width"], spec["length"] = 8, 6
    lift, drag = ctl.force_positions(spec, [(1, 3)])
    assert list(np.flatnonzero(lift)) == [3 * (1 * 8 + 3) + 2]
    assert list(np.flatnonzero(drag)) == [3 * (1 * 8 + 3) + 1]
    with pytest.raises(ValueError):
        ctl.force_positions(spec, [(6, 0)])


def test_phase_schedule(ctl):
    # steps 0..5 under the six phases, the zero force after (force_stop_frame = 5)
    assert [ctl.force_index(s) for s in range(9)] == [0, 1, 2, 3, 4, 5, 6, 6, 6]
    assert [ctl.force_index(s, 2) for s in range(5)] == [0, 1, 2, 6, 6]


def test_control_spec_has_no_fixed_nodes(ctl, ys):
    s = ctl.control_spec(ys.cloth_spec.homo_spec(5, pattern="satin"))
    assert s["handles"] == [] and s["pattern"] == "plain" and s["pose"] == "upright" and s["hight"] == 5.0


def test_force_tangents_are_the_derivatives_of_the_force(ctl, ys):
    net = ctl.ControlNet(ys.cloth_spec.homo_spec(5))
    rng = np.random.default_rng(1)
    s0 = rng.uniform(1e-4, 2e-3, (6, 2))
    for idx in range(7):
        D = net.force_tangents(idx)
        for j in range(12):
            e = np.zeros(12)
            e[j] = 1e-3
            diff = (net.ext_force(idx, s0.ravel() + e) - net.ext_force(idx, s0)) / 1e-3
            assert np.allclose(diff, D[j], rtol=0, atol=1e-12)
    assert not net.ext_force(6).any()
    with pytest.raises(ValueError):
        net.ext_force(7)


def test_mass_centre_loss(ctl):
    a = np.zeros((4, 3))
    b = np.zeros((4, 3))
    b[:, 2] = 2.0      # goal 2 higher: loss 4
    b[0, 0] = 4.0      # mean x of the goal 1: loss 5
    assert ctl.mass_centre_loss(a, b) == pytest.approx(5.0, rel=1e-15)


def test_learner_update_repeats_once_per_finite_strength(ctl, ys):
    net = ctl.ControlNet(ys.cloth_spec.homo_spec(5))
    s0 = net.strengths.copy()
    g = np.arange(1.0, 13.0)
    L = ctl.ControlLearner(net, lr=1e-4)
    assert L.update(g) == 12
    assert np.allclose(net.strengths.ravel(), s0.ravel() - 12 * 1e-4 * g, rtol=0, atol=1e-18)
    net.strengths[2, 1] = np.nan    # 11 finite strengths: 11 optimiser steps
    s1 = net.strengths.copy()
    assert L.update(g) == 11
    fin = np.isfinite(s1.ravel())
    assert np.allclose(net.strengths.ravel()[fin], s1.ravel()[fin] - 11 * 1e-4 * g[fin], rtol=0, atol=1e-18)
    L1 = ctl.ControlLearner(net, lr=1e-4, repeat_steps=False)
    net.strengths = s0.copy()
    assert L1.update(g) == 1
    assert np.allclose(net.strengths.ravel(), s0.ravel() - 1e-4 * g, rtol=0, atol=1e-18)
