"""Cloth descriptions for the implicit yarn-level step.

A cloth spec is a plain dict with the reference's Cloth constructor arguments
(src/Cloth.h:55-60), its Yarn list (src/Yarn.h:6-17), the Environment gravity
and wind (src/Environment.h, src/Wind.h) and the handle list
(src/Cloth.cpp:386-394).  The workloads below are the seeded synthetic cloths
the tests and the benchmark run; every random draw is seeded.
"""
import math

import numpy as np

PI_REF = 3.1415926535897  # src/Constants.h:24


def yarn_from_stiffness(R, rho, ks, kb):
    """Yarn with stretching stiffness ks [N] and bending stiffness kb [N m^2]:
    the reference stores Y and B with ks = Y*PI*R^2, kb = B*PI*R^2 (src/Yarn.h:14-15)."""
    area = PI_REF * R ** 2
    return {"rho": float(rho), "Y": float(ks) / area, "B": float(kb) / area}


# Material and environment values of the reference's own drivers
# (SimHomo src/Simulate.h:14-32, SimHeter src/Simulate.h:335-386): crossing spacing
# L = 1 mm, yarn radius R = 0.5 mm, cotton / polyester yarns, wind (0, 5, 0) m/s.
COTTON = {"rho": 0.002, "Y": 500000.0, "B": 0.00014}
POLYESTER = {"rho": 0.0025, "Y": 170000.0, "B": 0.00011}


def base_spec(W, pattern="plain"):
    W = int(W)
    return {
        "width": W, "length": W, "hight": 5.0, "L": 0.001, "R": 0.0005,
        "yarns": [dict(COTTON)],
        "kc": 1.0, "kf": 1000.0, "df": 1000.0, "mu": 0.5, "S": 1000.0, "handle_k": 1e4,
        "pattern": pattern, "pose": "upright",
        "G": [0.0, 0.0, 9.8],
        "wind_vel": [0.0, 5.0, 0.0], "wind_density": 2.0, "wind_drag": 0.5,
        "handles": [(0, 0), (W - 1, 0)],
        "h": 1e-3,
    }


def homo_spec(W, pattern="plain"):
    """Homogeneous cotton cloth of W x W crossings (SimHomo-style)."""
    return base_spec(W, pattern=pattern)


def heter_spec(W, pattern, seed):
    """Two-yarn blend (SimHeter-style): warp row / weft column i uses yarn i%2
    (src/Cloth.cpp:199-202).  rho, Y and B of each yarn are drawn log-uniform within
    +-0.3 decade of the cotton / polyester values, redrawn while Y/rho exceeds
    cotton's (the reference's step becomes unstable for stiffer-per-mass yarns); S
    and mu are drawn once per cloth (they are cloth-level in the reference,
    src/Cloth.h:34)."""
    rng = np.random.default_rng(seed)
    spec = base_spec(W, pattern=pattern)
    cap = COTTON["Y"] / COTTON["rho"]

    def lu(center, dec=0.3):
        return float(center * 10.0 ** rng.uniform(-dec, dec))

    yarns = []
    for base in (COTTON, POLYESTER):
        while True:
            y = {k: lu(base[k]) for k in ("rho", "Y", "B")}
            if y["Y"] / y["rho"] <= cap:
                break
        yarns.append(y)
    spec["yarns"] = yarns
    spec["S"] = lu(1000.0)
    spec["mu"] = lu(0.5)
    return spec


def initial_state(spec):
    """The constructor's rest state (src/Cloth.cpp:52-99): node (i, j) at
    (i L, 0, hight - j L) upright or (i L, j L, hight) flat, EPos = (i L, j L)."""
    W, Lg, L, hgt = spec["width"], spec["length"], spec["L"], spec["hight"]
    N = W * Lg
    s = {k: np.zeros((N, 3)) for k in ("LPos", "LPosFix", "LVel")}
    s.update({k: np.zeros((N, 2)) for k in ("EPos", "EPosBar", "EVel")})
    for j in range(Lg):
        for i in range(W):
            k = i + j * W  # state index u + v*width (= the reference's u + v*length when square)
            if spec["pose"] == "upright":
                s["LPos"][k] = (i * L, 0.0, hgt - j * L)
            else:
                s["LPos"][k] = (i * L, j * L, hgt)
            s["EPos"][k] = (i * L, j * L)
    s["LPosFix"][:] = s["LPos"]
    s["EPosBar"][:] = s["EPos"]
    return s


def jittered_state(spec, seed, sigma=1e-4):
    """Rest state with N(0, sigma) position jitter applied through set_ClothState,
    so LPosFix carries the same jitter (src/Cloth.cpp:421-438)."""
    s = initial_state(spec)
    rng = np.random.default_rng(seed)
    s["LPos"] = s["LPos"] + rng.normal(0.0, sigma, s["LPos"].shape)
    s["LPosFix"] = s["LPos"].copy()
    return s


def pressure_force(spec, state, p):
    """Uniform pressure p [N/m^2] along +z as an external generalized force on the
    L DOFs: each triangle of the cell split (src/Cloth.cpp:272-286) gives area*p/3 to
    each of its vertices.  Returns an n-vector in the global layout."""
    W, Lg = spec["width"], spec["length"]
    N = W * Lg
    n = 3 * N + 2 * (W - 2) * (Lg - 2)
    x = np.zeros((N, 3))
    for v in range(Lg):
        for u in range(W):
            x[u + v * W] = state["LPos"][u + v * W]
    f = np.zeros(n)
    for i in range(W - 1):
        for j in range(Lg - 1):
            tl, tr, bl, br = i + j * W, i + 1 + j * W, i + (j + 1) * W, i + 1 + (j + 1) * W
            tris = [(tl, tr, br), (tl, br, bl)] if (i + j) % 2 == 0 else [(tl, tr, bl), (tr, br, bl)]
            for a, b, c in tris:
                area = 0.5 * np.linalg.norm(np.cross(x[b] - x[a], x[c] - x[a]))
                for k in (a, b, c):
                    f[3 * k + 2] += area * p / 3.0
    return f


# The workloads.  Each returns (spec, state, ext_force or None, steps).  The
# initial state is the constructor's rest pose without position jitter: with a
# 1e-4 m jitter on a 1 cm grid and ks = 1e4 N the reference's system matrix is
# indefinite and its CG diverges already at 17 x 17 crossings, so the workloads use
# the reference drivers' own material values at the requested sizes instead.
def workload(name):
    if name == "homo5":          # 5x5 plain cotton, 100 steps
        spec = homo_spec(5)
        return spec, initial_state(spec), None, 100
    if name.startswith("heter64"):  # 64x64 blend in plain / twill / satin, 200 steps
        pattern = name.split("_")[1] if "_" in name else "plain"
        spec = heter_spec(64, pattern, 1)
        return spec, initial_state(spec), None, 200
    if name == "homo256":        # 256x256 plain with a seeded +z pressure, 50 steps
        spec = homo_spec(256)
        st = initial_state(spec)
        p = float(np.random.default_rng(2).uniform(0.0, 2.0))
        return spec, st, pressure_force(spec, st, p), 50
    if name == "twill128":       # 128x128 blend twill with seeded wind (6.38 m/s), 140 steps
        # The reference's stretch Jacobian (src/PhysicsCloth.cpp:473) gives a compressed
        # segment the transverse stiffness ks (1/l - 1/D) with no clamp; at h = 1 ms it
        # outweighs the segment's mass after a few percent of compression, so once the
        # hanging cloth folds A stops being definite and the step diverges: near step 154
        # with this wind, in the oracle and on the GPU alike (near step 100 without wind).
        # The length returned is the span that stays stable from rest.
        spec = heter_spec(128, "twill", 4)
        rng = np.random.default_rng(40)
        spec["wind_vel"] = [0.0, float(rng.uniform(2.0, 8.0)), 0.0]
        return spec, initial_state(spec), None, 140
    raise KeyError(name)


def batch_workload(n=64, seed=3, lo=10, hi=64):
    """Cloths of uneven size for batched stepping and parameter learning: square sides
    drawn uniformly from lo..hi, each a two-yarn blend with its own seed, weave
    patterns cycling plain / twill / satin.  Returns [(spec, state)]."""
    rng = np.random.default_rng(seed)
    sides = rng.integers(lo, hi + 1, size=n)
    out = []
    for i, W in enumerate(sides):
        spec = heter_spec(int(W), ("plain", "twill", "satin")[i % 3], 1000 * seed + i)
        out.append((spec, initial_state(spec)))
    return out


def ndof(spec):
    W, Lg = spec["width"], spec["length"]
    return 3 * W * Lg + 2 * (W - 2) * (Lg - 2)
