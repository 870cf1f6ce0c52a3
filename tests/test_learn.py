"""Host logic of the parameter learner: the parameter boxes of the reference's training
(src/TrainHelper.cpp:3-13) and the perturbed starting points.  No GPU needed."""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def learn(ys):
    return ys.learn


def test_box_map_closed_forms(learn):
    rng = (0.003, 0.001)
    # sigmoid(0) = 1/2: the middle of the box; the limits at +-infinity are the bounds
    assert learn.constrain(0.0, rng) == pytest.approx(0.002, rel=1e-15)
    assert learn.constrain(50.0, rng) == pytest.approx(0.003, rel=1e-12)
    assert learn.constrain(-50.0, rng) == pytest.approx(0.001, rel=1e-12)
    # sigmoid(log 3) = 3/4
    assert learn.constrain(math.log(3.0), (1200.0, 0.0)) == pytest.approx(900.0, rel=1e-14)


@pytest.mark.parametrize("name", ["rho0", "Y1", "B0", "S", "mu"])
def test_box_map_inverse_and_derivative(learn, name):
    rng = learn.RANGES[name]
    up, lo = rng
    for frac in (0.05, 0.3, 0.5, 0.77, 0.98):
        p = lo + frac * (up - lo)
        th = learn.unconstrain(p, rng)
        assert learn.constrain(th, rng) == pytest.approx(p, rel=1e-12)
        e = 1e-6
        fd = (learn.constrain(th + e, rng) - learn.constrain(th - e, rng)) / (2 * e)
        assert learn.dconstrain(th, rng) == pytest.approx(fd, rel=1e-8)
        # d/dtheta of the box map equals (p - lo)(up - p)/(up - lo)
        assert learn.dconstrain(th, rng) == pytest.approx((p - lo) * (up - p) / (up - lo), rel=1e-10)


def test_reference_boxes_hold_the_base_yarns(ys, learn):
    cs = ys.cloth_spec
    for y, names in ((cs.COTTON, ("rho0", "Y0", "B0")), (cs.POLYESTER, ("rho1", "Y1", "B1"))):
        for key, n in zip(("rho", "Y", "B"), names):
            up, lo = learn.RANGES[n]
            assert lo < y[key] < up


def test_perturbed_spec_is_seeded_and_boxed(ys, learn):
    cs = ys.cloth_spec
    spec = cs.heter_spec(12, "plain", 3)
    a = learn.perturbed_spec(spec, learn.NAMES, seed=7)
    b = learn.perturbed_spec(spec, learn.NAMES, seed=7)
    c = learn.perturbed_spec(spec, learn.NAMES, seed=8)
    for n in learn.NAMES:
        up, lo = learn.RANGES[n]
        va = ys.grad.get_param(a, n)
        assert va == ys.grad.get_param(b, n)
        assert lo < va < up
        r = va / ys.grad.get_param(spec, n)
        assert r <= 2.0 + 1e-12 and (r >= 0.5 - 1e-12 or va <= lo + 1e-5 * (up - lo))
    assert any(ys.grad.get_param(a, n) != ys.grad.get_param(c, n) for n in learn.NAMES)


def test_learner_rejects_values_outside_the_box(ys, learn):
    spec = ys.cloth_spec.homo_spec(6)
    spec["S"] = 5000.0
    with pytest.raises(ValueError):
        learn.ParameterLearner(spec, names=("S",))
