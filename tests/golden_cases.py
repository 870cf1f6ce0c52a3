"""Shared seeded cases for the parity tests (inputs only; expected values come from
the oracle or from the committed fixtures)."""
import numpy as np


def excited_state(cs, spec, seed, pos_jit=2e-5, vel=0.02, uv_jit=2e-5, evel=0.01):
    """Rest state plus seeded perturbations of every state array, so that the
    velocity-dependent (inertia, M-dot) and sliding-friction terms are non-zero."""
    rng = np.random.default_rng(seed)
    s = cs.initial_state(spec)
    L = spec["L"]
    N = s["LPos"].shape[0]
    s["LPos"] = s["LPos"] + rng.normal(0, pos_jit * L / 1e-3, (N, 3))
    s["LPosFix"] = s["LPos"] + rng.normal(0, 1e-6, (N, 3))
    s["LVel"] = rng.normal(0, vel, (N, 3))
    s["EPos"] = s["EPos"] + rng.normal(0, uv_jit * L / 1e-3, (N, 2))
    s["EPosBar"] = s["EPos"] + rng.normal(0, 1e-4 * L, (N, 2))
    s["EVel"] = rng.normal(0, evel, (N, 2))
    return s


def cases(cs):
    """(name, spec) pairs covering plain/twill/satin, one and two yarns, wind on/off."""
    out = []
    out.append(("homo5_plain", cs.homo_spec(5)))
    s = cs.heter_spec(7, "twill", 11)
    out.append(("heter7_twill", s))
    s = cs.heter_spec(9, "satin", 12)
    s["wind_vel"] = [0.3, 4.0, -1.0]
    out.append(("heter9_satin_wind", s))
    s = cs.homo_spec(6, "plain")
    s["wind_density"] = 0.0
    s["wind_drag"] = 0.0
    s["handles"] = [(0, 0), (5, 0), (2, 3)]
    out.append(("homo6_nowind_3handles", s))
    return out
