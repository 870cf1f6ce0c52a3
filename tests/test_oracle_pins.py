"""Pins of the oracle's hand-derived Jacobians and of the friction, crimp and jamming
terms, each against what the mathematics fixes (CPU only).

Every check isolates one element of the reference's lists (yo_set_element) and one
term (yo_set_terms) and compares the oracle's Jacobian entries with central
differences of the same element's force.  Where the reference's derivation deviates
from the derivative of its own force, the test states the deviation as written in the
reference source (file:line) and asserts that the oracle's entry equals the difference
quotient plus exactly that change, computed from the element's geometry and from
difference quotients of energies written from the reference's definitions.  A dropped
term, a wrong sign or index, or a transposed block in the oracle fails one of them.

- Bending_D (src/PhysicsCloth.cpp:643-822): the L-L blocks are the Hessian of the
  bending energy kb theta^2 / Delta_uv plus +cos(theta) P where the derivative has
  -cos(theta) P (:661-679); the E-E self entries carry theta^2 where the derivative has
  theta (:693); the E-L entries are 2 theta times the derivative (:683-684); the L-E
  entries are exact; placement Q8 (:693-822).
- Crimp_Bending_F/_D (:824-987) and the stretch terms feeding the compression: the
  compression accumulators equal signed gradients of the bending and stretch energies
  on the offset yarn copies.
- ParaColli_D (:1028-1031): the E-E entries are -kc L relu(d - Delta_uv), the force's
  value, not its derivative (Q6); the force itself is differentiated exactly.
- Inertia_D_Pos (:1188-1600): L-L lacks the Delta_uv factor (:1250); the mixed x-u
  blocks differentiate the w-part of dM/dx in u only through the Delta_uv prefactor
  (the sign at :1362); E-E is exact.
- MDotVDot_F/_D_Pos (:1903-2308): the force is -Mdot qdot; the x columns are the exact
  derivative of Mdot qdot (added with + into F_*_Pos); the u columns carry the written
  signs of w_uv and wTw_uv (:2046-2050), the factor 4 in w_dot_uv (:2052-2053) and the
  (rho/6) M_dot term (:2195-2200).
- ConSlideFriction (:2579-2641): limit cases of the tanh blends and the Q3 argument,
  and the assembly's wiring of the inputs.
- Shear (:2643-3089): the jamming force is minus the gradient of the lock-up energy
  plus the broadcast of (1 - con_func^2 * con_func_x) (Q4, :2699-2705); the jamming
  Hessian equals the force's derivative in the limit phi -> pi/2, where every Q4 term
  vanishes; the non-jamming cross blocks misplace a (phi - pi/2) factor and drop P1 P3
  (:3062-3066).
"""
import copy

import numpy as np
import pytest

from oracle.oracle import OracleCloth, friction_eval
from tests.golden_cases import excited_state

PI = 3.1415926535897   # src/Constants.h:24


def _one(spec, state, kind, idx, terms, h=1e-3):
    c = OracleCloth(spec)
    c.set_terms(terms)
    c.set_element(kind, idx)
    c.set_state(state)
    c.assemble(h)
    return c


def _gdof(spec, k, comp):
    """Global index of component comp (0-2 L, 3-4 E) of node k."""
    W, Lg = spec["width"], spec["length"]
    if comp < 3:
        return 3 * k + comp
    u, v = k % W, k // W
    return 3 * W * Lg + 2 * ((u - 1) + (v - 1) * (W - 2)) + comp - 3


def _fd(spec, state, kind, idx, terms, nodes, eps_x=1e-8, eps_u=1e-9, key_x="LPos", key_u="EPos"):
    """dF/dq for the L and (interior) E coordinates of `nodes`: {global col: dF}."""
    W, Lg = spec["width"], spec["length"]
    out = {}
    for k in nodes:
        u, v = k % W, k // W
        comps = [0, 1, 2] + ([3, 4] if 0 < u < W - 1 and 0 < v < Lg - 1 else [])
        for comp in comps:
            key, a, eps = (key_x, comp, eps_x) if comp < 3 else (key_u, comp - 3, eps_u)
            s1, s2 = copy.deepcopy(state), copy.deepcopy(state)
            s1[key][k, a] += eps
            s2[key][k, a] -= eps
            out[_gdof(spec, k, comp)] = (_one(spec, s1, kind, idx, terms).vec("F") -
                                         _one(spec, s2, kind, idx, terms).vec("F")) / (2 * eps)
    return out


def _llT(K, F, D):
    """The L-L block (force node F, DOF node D) as d F_F / d x_D (blocks enter A
    transposed, src/MatrixMp.cpp:65-76)."""
    return K[3 * F:3 * F + 3, 3 * D:3 * D + 3].T


def _fdblk(fd, F, D, spec):
    return np.stack([fd[_gdof(spec, D, b)][3 * F:3 * F + 3] for b in range(3)], axis=1)


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def _kb(spec, yarn):
    return spec["yarns"][yarn]["B"] * PI * spec["R"] ** 2


def _ks(spec, yarn):
    return spec["yarns"][yarn]["Y"] * PI * spec["R"] ** 2


# ---------------------------------------------------------------------------
# Bending
# ---------------------------------------------------------------------------
def _bend_nodes(spec, bi):
    """(n2, n0, n1, direction c, yarn) of bend bi in the reference's list order
    (src/Cloth.cpp:230-270): warp rows first, then weft columns."""
    W, Lg = spec["width"], spec["length"]
    nyarn = len(spec["yarns"])
    nw = Lg * (W - 2)
    if bi < nw:
        v, j = divmod(bi, W - 2)
        return j + v * W, j + 1 + v * W, j + 2 + v * W, 0, 0 if nyarn == 1 else v % 2
    u, j = divmod(bi - nw, Lg - 2)
    return u + j * W, u + (j + 1) * W, u + (j + 2) * W, 1, 0 if nyarn == 1 else u % 2


@pytest.mark.parametrize("pattern,seed,bend", [("plain", 3, "warp"), ("twill", 4, "weft"), ("satin", 6, "warp")])
def test_bending_jacobian_against_differences(cs, pattern, seed, bend):
    W = 6
    spec = cs.heter_spec(W, pattern, seed)
    spec.update(wind_density=0.0, wind_drag=0.0)
    st = excited_state(cs, spec, seed, vel=0.0, evel=0.0)
    bi = (2 * (W - 2) + 1) if bend == "warp" else (W * (W - 2) + 2 * (W - 2) + 1)
    n2, n0, n1, c, yarn = _bend_nodes(spec, bi)
    K = _one(spec, st, "bend", bi, ["bend"]).dense(2)
    fd = _fd(spec, st, "bend", bi, ["bend"], (n2, n0, n1))
    x = st["LPos"]
    a, b = x[n1] - x[n0], x[n2] - x[n0]
    l1, l2 = np.linalg.norm(a), np.linalg.norm(b)
    d1, d2 = a / l1, b / l2
    th = np.arccos(-d1 @ d2)
    D = st["EPos"][n1, c] - st["EPos"][n2, c]
    kb = _kb(spec, yarn)
    assert 0.01 < th < 0.5
    # L-L: Hessian of kb theta^2 / D plus the written +cos(theta) P (src/PhysicsCloth.cpp:661-679)
    P1, P2 = np.eye(3) - np.outer(d1, d1), np.eye(3) - np.outer(d2, d2)
    q1 = 4 * kb * th * np.cos(th) / (l1 ** 2 * D * np.sin(th)) * P1
    q2 = 4 * kb * th * np.cos(th) / (l2 ** 2 * D * np.sin(th)) * P2
    corr = {(n1, n1): q1, (n1, n0): -q1, (n0, n1): -q1, (n2, n2): q2, (n2, n0): -q2, (n0, n2): -q2,
            (n0, n0): q1 + q2, (n1, n2): 0 * q1, (n2, n1): 0 * q1}
    num = den = 0.0
    for (F, Dn), cc in corr.items():
        e = _llT(K, F, Dn) - _fdblk(fd, F, Dn, spec) - cc
        num += np.sum(e ** 2)
        den += np.sum(_fdblk(fd, F, Dn, spec) ** 2)
    assert np.sqrt(num / den) < 1e-6
    # without the documented change the blocks are far from the derivative
    assert _rel(_llT(K, n1, n1), _fdblk(fd, n1, n1, spec)) > 0.05
    E = {k: _gdof(spec, k, 3 + c) for k in (n2, n0, n1)}
    L = lambda k: slice(3 * k, 3 * k + 3)   # noqa: E731
    # L-E: exact where stored; the n1-n2 cross entries are not stored (src/PhysicsCloth.cpp:780-822)
    for (kl, ke) in ((n1, n1), (n2, n2), (n0, n1), (n0, n2)):
        assert _rel(K[L(kl), E[ke]], fd[E[ke]][L(kl)]) < 1e-6
    for (kl, ke) in ((n1, n2), (n2, n1)):
        assert np.all(K[L(kl), E[ke]] == 0) and np.linalg.norm(fd[E[ke]][L(kl)]) > 0
    # E-E: theta^2 where the derivative has theta (:693); no E1-E2 entries (Q8)
    for k in (n1, n2):
        assert K[E[k], E[k]] == pytest.approx(th * fd[E[k]][E[k]], rel=1e-6)
    assert K[E[n1], E[n2]] == 0 and K[E[n2], E[n1]] == 0 and abs(fd[E[n2]][E[n1]]) > 0
    # E-L: 2 theta times the derivative (:683-684); E2L0 lands on row E(n0), column L(n2)
    dEL = lambda ke, kl: np.array([fd[_gdof(spec, kl, a_)][E[ke]] for a_ in range(3)])   # noqa: E731
    for (ke, kl) in ((n1, n1), (n1, n0), (n2, n2)):
        assert _rel(K[E[ke], L(kl)], 2 * th * dEL(ke, kl)) < 1e-6
    assert _rel(K[E[n0], L(n2)], -2 * th * dEL(n1, n0)) < 1e-6
    for (ke, kl) in ((n1, n2), (n2, n1), (n2, n0)):
        assert np.all(K[E[ke], L(kl)] == 0) and np.linalg.norm(dEL(ke, kl)) > 0


# ---------------------------------------------------------------------------
# Compression: stretch and crimp bending (src/PhysicsCloth.cpp:427-577, 824-987, 2593-2596)
# ---------------------------------------------------------------------------
def _grad(f, X, eps):
    g = np.zeros_like(X)
    for idx in np.ndindex(X.shape):
        Xp, Xm = X.copy(), X.copy()
        Xp[idx] += eps
        Xm[idx] -= eps
        g[idx] = (f(Xp) - f(Xm)) / (2 * eps)
    return g


def test_compression_from_stretch_and_crimp_energies(cs):
    """StrhBend_C_L of every node is the sum over the stretch elements and the crimp
    bends (the bending element on the yarn copies x -/+ R n, with their construction-time
    Eulerian coordinates, src/Cloth.cpp:191-192, 321-325) of their forces -grad E, signed
    + on warp and - on weft elements; StrhBend_C_LE's self entries are the derivatives of
    those forces in the node's own Eulerian coordinate; the compression is
    relu(n . StrhBend_C_L / 2)."""
    W = 6
    spec = cs.heter_spec(W, "twill", 8)
    spec.update(wind_density=0.0, wind_drag=0.0)
    st = excited_state(cs, spec, 9, vel=0.0, evel=0.0)
    c = OracleCloth(spec)
    c.set_terms(["stretch", "crimp"])
    c.set_state(st)
    c.assemble(1e-3)
    sbc, sbc_le = c.compression_parts()
    nrm = c.normals()
    N, L, R = W * W, spec["L"], spec["R"]
    x, uv = st["LPos"], st["EPos"]
    inter = lambda k: 0 < k % W < W - 1 and 0 < k // W < W - 1   # noqa: E731
    yarn = lambda c_, k: (k // W) % 2 if c_ == 0 else (k % W) % 2   # noqa: E731
    exp_l = np.zeros((N, 3))
    exp_le = np.zeros((N, 2, 3))
    # stretch: E = ks D (|w| - 1)^2 / 2, w = (x1 - x0) / D, D = u1 - u0
    for c_ in (0, 1):
        sg = 1.0 if c_ == 0 else -1.0
        for k0 in range(N):
            k1 = k0 + 1 if c_ == 0 else k0 + W
            if (c_ == 0 and k0 % W == W - 1) or (c_ == 1 and k0 // W == W - 1):
                continue
            ks = _ks(spec, yarn(c_, k0))

            def energy(X, D):
                w = (X[1] - X[0]) / D
                return 0.5 * ks * D * (np.linalg.norm(w) - 1.0) ** 2
            X = np.stack([x[k0], x[k1]])
            D = uv[k1, c_] - uv[k0, c_]
            F = -_grad(lambda Y: energy(Y, D), X, 1e-9)
            exp_l[k0] += sg * F[0]
            exp_l[k1] += sg * F[1]
            hD = 1e-10
            dF = (-_grad(lambda Y: energy(Y, D + hD), X, 1e-9) + _grad(lambda Y: energy(Y, D - hD), X, 1e-9)) / (2 * hD)
            # D = u1 - u0: d/du1 = d/dD, d/du0 = -d/dD
            if inter(k0):
                exp_le[k0, c_] += sg * -dF[0]
            if inter(k1):
                exp_le[k1, c_] += sg * dF[1]
    # crimp bends: E = kb theta^2 / D on the copies, D from the construction-time uv (2 L)
    for c_ in (0, 1):
        sg = 1.0 if c_ == 0 else -1.0
        cp = np.array([x[k] - sg * R * nrm[k] if inter(k) else x[k] for k in range(N)])
        for k2 in range(N):
            k0 = k2 + 1 if c_ == 0 else k2 + W
            k1 = k2 + 2 if c_ == 0 else k2 + 2 * W
            if (c_ == 0 and k2 % W > W - 3) or (c_ == 1 and k2 // W > W - 3):
                continue
            kb = _kb(spec, yarn(c_, k2))

            def energy(X, D):
                a, b = X[2] - X[1], X[0] - X[1]
                th = np.arccos(-(a @ b) / (np.linalg.norm(a) * np.linalg.norm(b)))
                return kb * th ** 2 / D
            X = np.stack([cp[k2], cp[k0], cp[k1]])
            D = 2 * L
            F = -_grad(lambda Y: energy(Y, D), X, 1e-10)
            for i, k in enumerate((k2, k0, k1)):
                exp_l[k] += sg * F[i]
            hD = 1e-10
            dF = (-_grad(lambda Y: energy(Y, D + hD), X, 1e-10) + _grad(lambda Y: energy(Y, D - hD), X, 1e-10)) / (2 * hD)
            # D = u1 - u2: the n1 entry is dF1/du1 = dF1/dD, the n2 entry dF2/du2 = -dF2/dD
            if inter(k1):
                exp_le[k1, c_] += sg * dF[2]
            if inter(k2):
                exp_le[k2, c_] += sg * -dF[0]
    assert _rel(sbc, exp_l) < 1e-6
    ints = [k for k in range(N) if inter(k)]
    assert _rel(sbc_le[ints], exp_le[ints]) < 1e-5
    comp = c.crossing()[1]
    exp_comp = np.array([max(0.0, 0.5 * nrm[k] @ exp_l[k]) for k in ints])
    assert np.abs(comp - exp_comp).max() < 1e-6 * np.abs(exp_comp).max()
    assert (exp_comp > 0).sum() >= 3


# ---------------------------------------------------------------------------
# ParaColli (Q6)
# ---------------------------------------------------------------------------
def test_paracolli_value_entries(cs):
    """Plain weave: neighbouring interior crossings have opposite flips, so d = 4R = 2L
    exceeds Delta_uv ~ L and the repulsion is active (src/PhysicsCloth.cpp:989-1069).
    Its E force is differentiated exactly by differences; the E-E entries hold
    -kc L relu(d - Delta_uv), the value (src/PhysicsCloth.cpp:1028-1031)."""
    W = 5
    spec = cs.heter_spec(W, "plain", 3)
    spec.update(kc=3.0)
    st = excited_state(cs, spec, 11, vel=0.0, evel=0.0)
    ei = 2 * (W - 1) + 1          # warp edge (1,2)-(2,2)
    k0, k1 = 1 + 2 * W, 2 + 2 * W
    K = _one(spec, st, "edge", ei, ["paracolli"]).dense(2)
    Fv = _one(spec, st, "edge", ei, ["paracolli"]).vec("F")
    Duv = st["EPos"][k1, 0] - st["EPos"][k0, 0]
    d = 4 * spec["R"]
    val = spec["kc"] * spec["L"] * max(0.0, d - Duv)
    assert val > 0
    e0, e1 = _gdof(spec, k0, 3), _gdof(spec, k1, 3)
    assert Fv[e0] == pytest.approx(val, rel=1e-14) and Fv[e1] == pytest.approx(-val, rel=1e-14)
    for r in (e0, e1):
        for cc in (e0, e1):
            assert K[r, cc] == pytest.approx(-val, rel=1e-14)
    fd = _fd(spec, st, "edge", ei, ["paracolli"], (k0, k1))
    kcL = spec["kc"] * spec["L"]
    assert fd[e0][e0] == pytest.approx(kcL, rel=1e-6) and fd[e1][e0] == pytest.approx(-kcL, rel=1e-6)


# ---------------------------------------------------------------------------
# Edge kinetic terms: Inertia and M-dot (src/PhysicsCloth.cpp:1071-2577)
# ---------------------------------------------------------------------------
def _kron2(B):
    return np.block([[2 * B, B], [B, 2 * B]])


class _Edge:
    """One edge's kinetic quantities written from the definitions of
    src/PhysicsCloth.cpp:250-338 (M) and 1903-1961 (Mdot)."""

    def __init__(self, spec, st, k0, k1, c, yarn):
        self.rho = spec["yarns"][yarn]["rho"]
        self.x0, self.x1 = st["LPos"][k0].copy(), st["LPos"][k1].copy()
        self.u0, self.u1 = st["EPos"][k0, c], st["EPos"][k1, c]
        self.v0, self.v1 = st["LVel"][k0].copy(), st["LVel"][k1].copy()
        self.ud0, self.ud1 = st["EVel"][k0, c], st["EVel"][k1, c]
        self.q = np.concatenate([self.v0, [self.ud0], self.v1, [self.ud1]])

    def T(self, x0, u0, x1, u1, part="all"):
        """Kinetic energy q^T M q / 2, or its I-, w- or wTw-part."""
        D = u1 - u0
        w = (x1 - x0) / D
        B = np.zeros((4, 4))
        if part in ("I", "all"):
            B[:3, :3] = np.eye(3)
        if part in ("w", "all"):
            B[:3, 3] = -w
            B[3, :3] = -w
        if part in ("s", "all"):
            B[3, 3] = w @ w
        return 0.5 * self.q @ (self.rho * D / 6 * _kron2(B)) @ self.q

    def ws(self, x0, u0, x1, u1):
        D, Dd = u1 - u0, self.ud1 - self.ud0
        dx, dv = x1 - x0, self.v1 - self.v0
        w = dx / D
        wd = (dv * D - dx * Dd) / D ** 2
        return w, wd, D, Dd

    def f(self, x0, u0, x1, u1):
        """Mdot qdot (8)."""
        w, wd, D, Dd = self.ws(x0, u0, x1, u1)
        B = np.block([[np.eye(3), -w[:, None]], [-w[None, :], np.array([[w @ w]])]])
        Bd = np.block([[np.zeros((3, 3)), -wd[:, None]], [-wd[None, :], np.array([[2 * w @ wd]])]])
        return self.rho / 6 * (Dd * _kron2(B) + D * _kron2(Bd)) @ self.q


def _dvar(fun, args, i, a, eps):
    """Central difference of fun(*args) in argument i (component a for vectors)."""
    p, m = [np.array(x, dtype=float, copy=True) for x in args], [np.array(x, dtype=float, copy=True) for x in args]
    if np.ndim(args[i]) == 0:
        p[i], m[i] = args[i] + eps, args[i] - eps
    else:
        p[i][a] += eps
        m[i][a] -= eps
    return (fun(*p) - fun(*m)) / (2 * eps)


_XI, _UJ = (0, 2), (1, 3)   # argument positions of x0, x1 and u0, u1


def _edge_setup(cs, pattern, seed, warp):
    W = 6
    spec = cs.heter_spec(W, pattern, seed)
    spec.update(wind_density=0.0, wind_drag=0.0)
    st = excited_state(cs, spec, seed)
    if warp:
        v, u = 2, 2
        ei, k0, k1, c, yarn = v * (W - 1) + u, u + v * W, u + 1 + v * W, 0, v % 2
    else:
        u, v = 3, 2
        ei, k0, k1, c, yarn = W * (W - 1) + u * (W - 1) + v, u + v * W, u + (v + 1) * W, 1, u % 2
    return spec, st, ei, k0, k1, c, _Edge(spec, st, k0, k1, c, yarn)


@pytest.mark.parametrize("warp", [True, False])
def test_inertia_position_jacobian(cs, warp):
    spec, st, ei, k0, k1, c, ed = _edge_setup(cs, "twill", 21, warp)
    K = _one(spec, st, "edge", ei, ["inertia"]).dense(2)
    fd = _fd(spec, st, "edge", ei, ["inertia"], (k0, k1))
    args = (ed.x0, ed.u0, ed.x1, ed.u1)
    D = ed.u1 - ed.u0
    nodes = (k0, k1)
    # L-L: the second derivative of M in x written without the Delta_uv factor (:1250)
    for F in nodes:
        for Dn in nodes:
            assert _rel(_llT(K, F, Dn) * D, _fdblk(fd, F, Dn, spec)) < 1e-6
    # E-E: exact
    E = [_gdof(spec, k, 3 + c) for k in nodes]
    for i in range(2):
        for j in range(2):
            assert K[E[i], E[j]] == pytest.approx(fd[E[j]][E[i]], rel=1e-6)
    # mixed blocks: the w-part of dM/dx is differentiated in u only through the
    # Delta_uv prefactor (the written sign at :1362 cancels the rest in q^T (.) q)
    e_x, e_u = 1e-7, 1e-9
    for i, ki in enumerate(nodes):
        for j, kj in enumerate(nodes):
            dDdu = -1.0 if j == 0 else 1.0
            le, el = np.zeros(3), np.zeros(3)
            for a in range(3):
                dTw = _dvar(lambda *z: ed.T(*z, part="w"), args, _XI[i], a, e_x)
                dTs = _dvar(lambda *z: ed.T(*z, part="s"), args, _XI[i], a, e_x)
                d2Ts = _dvar(lambda *z: _dvar(lambda *y: ed.T(*y, part="s"), z, _XI[i], a, e_x), args, _UJ[j], 0, e_u)
                le[a] = dDdu / D * dTw + d2Ts
                el[a] = dDdu / D * dTw - d2Ts + 2 * dDdu / D * dTs
            assert _rel(K[3 * ki:3 * ki + 3, E[j]], le) < 1e-5
            assert _rel(K[E[j], 3 * ki:3 * ki + 3], el) < 1e-5
    # and the exact derivative would differ
    assert _rel(K[3 * k0:3 * k0 + 3, E[0]], fd[E[0]][3 * k0:3 * k0 + 3]) > 0.1


@pytest.mark.parametrize("warp", [True, False])
def test_mdot_force_and_position_jacobian(cs, warp):
    spec, st, ei, k0, k1, c, ed = _edge_setup(cs, "satin", 22, warp)
    one = _one(spec, st, "edge", ei, ["mdot"])
    K, Fv = one.dense(2), one.vec("F")
    args = (ed.x0, ed.u0, ed.x1, ed.u1)
    f = ed.f(*args)
    E = [_gdof(spec, k, 3 + c) for k in (k0, k1)]
    rows = list(range(3 * k0, 3 * k0 + 3)) + [E[0]] + list(range(3 * k1, 3 * k1 + 3)) + [E[1]]
    # MDotVDot_F: F = -Mdot qdot (:1939-1960)
    assert _rel(Fv[rows], -f) < 1e-12

    def kcol(col):   # K's column for DOF `col`, in the edge's q order (L blocks transposed back)
        out = np.zeros(8)
        for r, g in enumerate(rows):
            if r in (3, 7) or col >= 3 * spec["width"] * spec["length"]:
                out[r] = K[g, col]
            else:   # L row, L column: A[3F+v, 3D+u] = B[u][v]
                Fn, a = divmod(g, 3)
                Dn, b = divmod(col, 3)
                out[r] = K[3 * Fn + b, 3 * Dn + a]
        return out
    e_x, e_u = 1e-8, 1e-10
    # x columns: the exact derivative of Mdot qdot, added with + (:2212-2308)
    for i, ki in enumerate((k0, k1)):
        for a in range(3):
            dfx = _dvar(ed.f, args, _XI[i], a, e_x)
            assert _rel(kcol(3 * ki + a), dfx) < 1e-6
    # u columns: the written w_uv, wTw_uv signs, the 4 in w_dot_uv and the (rho/6) M_dot term
    rho = ed.rho
    w, wd, D, Dd = ed.ws(*args)
    dx = ed.x1 - ed.x0
    Bd = np.block([[np.zeros((3, 3)), -wd[:, None]], [-wd[None, :], np.array([[2 * w @ wd]])]])
    for j in range(2):
        dfu = _dvar(ed.f, args, _UJ[j], 0, e_u)
        dw = _dvar(lambda *z: ed.ws(*z)[0], args, _UJ[j], 0, e_u)
        dwTw = _dvar(lambda *z: ed.ws(*z)[0] @ ed.ws(*z)[0], args, _UJ[j], 0, e_u)
        dwd_extra = (-1.0 if j == 0 else 1.0) * 2 * Dd * dx / D ** 3
        dB = np.zeros((4, 4))
        dB[3, 3] = -2 * dwTw
        dBd = np.zeros((4, 4))
        dBd[:3, 3] = -dwd_extra
        dBd[3, :3] = -dwd_extra
        dBd[3, 3] = -4 * dw @ wd + 2 * dwd_extra @ w
        # the exact derivative's dDelta/du_j (rho/6) kron2(Bdot) qdot is written as -(rho/6) M_dot qdot
        dDdu = -1.0 if j == 0 else 1.0
        dev = (rho / 6 * Dd * _kron2(dB) @ ed.q - dDdu * rho / 6 * _kron2(Bd) @ ed.q - rho / 6 * f +
               rho / 6 * D * _kron2(dBd) @ ed.q)
        got = kcol(E[j])
        assert _rel(got, dfu + dev) < 1e-5
        assert _rel(got, dfu) > 0.1


# ---------------------------------------------------------------------------
# Friction (src/PhysicsCloth.cpp:2579-2641)
# ---------------------------------------------------------------------------
def test_friction_limits():
    kf, df, mu = 1e3, 1e3 * 0.7, 0.5
    for dlt in (2e-4, -3e-5, 0.07):
        th = np.tanh(10 * dlt)
        # stick: lim - F >> 0 (every tanh blend at +1): force mu comp tanh(10 dlt) - df udot,
        # EE position entry exactly -kf, velocity entry -df, is_slide 0
        comp, F, ud, cEE = 10.0, 0.3, 0.02, 0.37
        s, fr, fp, fv = friction_eval(dlt, comp, cEE, F, F, 1.7, 2.9, ud, mu, kf, df)
        assert s == 0.0
        assert fr == pytest.approx(th * mu * comp - df * ud, rel=1e-12, abs=1e-12)
        assert fp == pytest.approx(-kf, rel=1e-15) and fv == -df
        # slide: F >> lim (blends at -1): force kf dlt - df udot, EE position entry
        # -(1 - tanh^2) mu comp - tanh mu compEE, velocity entry -df, is_slide 1
        comp, F = 1.0, 30.0
        s, fr, fp, fv = friction_eval(dlt, comp, cEE, F, F, 1.7, 2.9, ud, mu, kf, df)
        assert s == 1.0
        assert fr == pytest.approx(kf * dlt - df * ud, rel=1e-12, abs=1e-12)
        assert fp == pytest.approx(-(1 - th ** 2) * mu * comp - th * mu * cEE, rel=1e-12, abs=1e-12)
        assert fv == -df
    # no slip displacement: the force is pure damping whatever the blend
    for F in (-3.0, 0.2, 0.45, 30.0):
        s, fr, fp, fv = friction_eval(0.0, 1.0, 0.3, F, F, 1.7, 2.9, 0.013, mu, kf, df)
        assert fr == -df * 0.013 and fv == -df
        tl = np.tanh(10 * (mu - F))
        assert fp == pytest.approx(-0.5 * (kf - mu) * tl - 0.5 * (kf + mu), rel=1e-14)
    # |F| = lim is the sigmoid's midpoint; the velocity entry's blend takes tanh(10 lim - F)
    # (Q3, :2630-2631), the position entry's tanh(10 (lim - F))
    s, fr, fp, fv = friction_eval(1e-3, 1.0, 0.0, 0.5, 0.5, 0.0, 2.9, 0.0, mu, kf, df)
    assert s == 0.5
    dlt, comp, F, Fv = 0.01, 1.0, 0.45, 2.9
    th = np.tanh(10 * dlt)
    _, _, _, fv = friction_eval(dlt, comp, 0.0, F, F, 0.0, Fv, 0.0, mu, kf, df)
    q3 = -0.5 * (kf * dlt - th * mu * comp) * (1 - np.tanh(10 * mu * comp - F) ** 2) * Fv - df
    alt = -0.5 * (kf * dlt - th * mu * comp) * (1 - np.tanh(10 * (mu * comp - F)) ** 2) * Fv - df
    assert fv == pytest.approx(q3, rel=1e-14) and abs(fv - alt) > 1e-3 * abs(alt)


def test_friction_wiring(cs):
    """The assembly feeds ConSlideFriction the no-slide E force and EE diagonals, the
    thread-partial E force (Q2), the compression and its EE input, uv - uvbar and the
    Eulerian velocity, and adds its outputs to F_E and the EE diagonals
    (src/PhysicsCloth.cpp:97-123, 2633-2640)."""
    W = 7
    spec = cs.heter_spec(W, "satin", 12)
    st = excited_state(cs, spec, 13)
    terms = ["mass", "gravity", "stretch", "inertia", "mdot", "paracolli", "bend", "crimp", "shear", "wind", "handle"]
    a, b = OracleCloth(spec), OracleCloth(spec)
    a.set_state(st)
    b.set_terms(terms)
    b.set_state(st)
    a.assemble(1e-3)
    b.assemble(1e-3)
    N = W * W
    Fa, Fb = a.vec("F"), b.vec("F")
    Kxa, Kxb, Kva, Kvb = a.dense(2), b.dense(2), a.dense(3), b.dense(3)
    sl, comp, fpart = a.crossing()
    _, sbc_le = a.compression_parts()
    nrm = a.normals()
    e = 0
    for v in range(1, W - 1):
        for u in range(1, W - 1):
            k = u + v * W
            for c in range(2):
                r = 3 * N + 2 * e + c
                out = friction_eval(st["EPos"][k, c] - st["EPosBar"][k, c], comp[e], 0.5 * nrm[k] @ sbc_le[k, c], Fb[r],
                                    fpart[e, c], Kxb[r, r], Kvb[r, r], st["EVel"][k, c], spec["mu"], spec["kf"], spec["df"])
                assert sl[e, c] == out[0]
                assert Fa[r] - Fb[r] == pytest.approx(out[1], rel=1e-9, abs=1e-12)
                assert Kxa[r, r] - Kxb[r, r] == pytest.approx(out[2], rel=1e-9, abs=1e-9)
                assert Kva[r, r] - Kvb[r, r] == pytest.approx(out[3], rel=1e-9, abs=1e-9)
            e += 1


# ---------------------------------------------------------------------------
# Shear (src/PhysicsCloth.cpp:2643-3089)
# ---------------------------------------------------------------------------
class _Shear:
    """The jamming lock-up of one shear segment, from the definitions of
    src/PhysicsCloth.cpp:2656-2699 (phi_bar = pi/2, phi_l = the jamming threshold
    2 asin(R/L) of src/Cloth.cpp:29, c = 3, sigma = 0.6)."""

    def __init__(self, spec, comp=0.0):
        self.L, self.R, self.S = spec["L"], spec["R"], spec["S"]
        self.pb, self.pl = PI / 2, 2 * np.arcsin(self.R / self.L)
        self.base = 0.5 * (comp + 1.0) * self.S * PI * self.R ** 2

    @staticmethod
    def phi(X):
        a, b = X[1] - X[0], X[2] - X[0]
        return np.arccos(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))

    def cfx(self, p):
        ps = p * (p - self.pl) * (p - self.pb)
        return self.pb ** 5 * (p - self.pl) / (ps ** 2 + self.pb ** 4 * 0.36)

    def gamma(self, p):
        return (self.L * np.sqrt(2.0) - 2 * self.L * np.sin(0.5 * p)) / self.R

    def kx(self, p):
        g3 = self.gamma(p) ** 3
        return self.base * ((1 + g3) + (1 - g3) * np.tanh(self.cfx(p)))

    def energy(self, X):
        p = self.phi(X)
        return 0.5 * self.kx(p) * self.L * (p - self.pb) ** 2


def _jam_setup(cs, angle=None):
    """W = 5 plain: the centre crossing (2, 2) is unflipped and its four neighbours are
    flipped, so its four shear segments jam (src/ClothUnit.h:226-230)."""
    W = 5
    spec = cs.heter_spec(W, "plain", 3)
    spec.update(wind_density=0.0, wind_drag=0.0)
    st = excited_state(cs, spec, 17, pos_jit=6e-5, vel=0.0, evel=0.0)
    k = 2 + 2 * W
    si = 4 * ((2 - 1) * (W - 2) + (2 - 1))     # segment (left, bottom) of crossing (2, 2)
    n1, n3 = k - 1, k + W
    if angle is not None:   # place n1 and n3 at the given angle around n0
        x0 = st["LPos"][k]
        a = np.array([-1.0, 0.1, 0.05])
        a /= np.linalg.norm(a)
        b = np.cross(a, [0.2, 1.0, 0.3])
        b /= np.linalg.norm(b)
        st["LPos"][n1] = x0 + 1.02e-3 * a
        st["LPos"][n3] = x0 + 0.97e-3 * (np.cos(angle) * a + np.sin(angle) * b)
    return spec, st, si, k, n1, n3


def test_jamming_force_is_the_energy_gradient_plus_the_broadcast(cs):
    """Shear alone (the compression, and with it kx's prefactor, is then zero): the
    jamming force equals -grad of E = kx(phi) L (phi - pi/2)^2 / 2 plus the term that
    (1 - con_func^2 * con_func_x) adds by broadcasting 1 over the vector
    (:2699-2705): -L (phi - pi/2)^2 base (1 - gamma^3) ((1,1,1) - grad con_func_x) / 2."""
    spec, st, si, k0, n1, n3 = _jam_setup(cs)
    Fv = _one(spec, st, "shear", si, ["shear"]).vec("F")
    sh = _Shear(spec)
    X = np.stack([st["LPos"][k0], st["LPos"][n1], st["LPos"][n3]])
    gE = _grad(sh.energy, X, 1e-10)
    gcfx = _grad(lambda Y: sh.cfx(sh.phi(Y)), X, 1e-10)
    p = sh.phi(X)
    assert abs(p - PI / 2) > 0.01
    pre = -0.5 * spec["L"] * (p - sh.pb) ** 2 * sh.base * (1 - sh.gamma(p) ** 3)
    exp1 = -gE[1] + pre * (np.ones(3) - gcfx[1])
    exp3 = -gE[2] + pre * (np.ones(3) - gcfx[2])
    for kk, ex in ((n1, exp1), (n3, exp3), (k0, -(exp1 + exp3))):
        assert _rel(Fv[3 * kk:3 * kk + 3], ex) < 1e-6
    # the broadcast term is not small: without it the force is off
    assert _rel(Fv[3 * n1:3 * n1 + 3], -gE[1]) > 1e-3


def test_jamming_hessian_in_the_right_angle_limit(cs):
    """Every Q4 term of the jamming Hessian (the precedence in con_func'' at :2914,2918,
    kx'' without pi at :2985, the duplicated terms of phi_s'' at :2810-2848) is multiplied
    by (phi - pi/2) or its square.  At phi -> pi/2 the blocks therefore tend to the
    derivative of the force, and their distance from it shrinks linearly with
    |phi - pi/2|."""
    devs = []
    for delta in (3e-2, 3e-3, 0.0):
        spec, st, si, k0, n1, n3 = _jam_setup(cs, angle=PI / 2 + delta)
        K = _one(spec, st, "shear", si, ["shear"]).dense(2)
        fd = _fd(spec, st, "shear", si, ["shear"], (k0, n1, n3), eps_x=1e-9)
        num = den = 0.0
        for F in (k0, n1, n3):
            for D in (k0, n1, n3):
                num += np.sum((_llT(K, F, D) - _fdblk(fd, F, D, spec)) ** 2)
                den += np.sum(_fdblk(fd, F, D, spec) ** 2)
        devs.append(np.sqrt(num / den))
    assert devs[2] < 1e-6                       # at phi = pi/2: the derivative, to difference error
    assert 0.05 < devs[1] / devs[0] < 0.2       # linear in phi - pi/2 (2.2 |phi - pi/2| here)


def test_non_jamming_cross_blocks(cs):
    """Non-jamming segments (3 x 3 plain: no flips): the outer nodes' self blocks are the
    force's derivative; the n1-n3 blocks are the derivative plus the written change: the
    (phi - pi/2) factor spans the 1/sin(phi) term and the P1 P3 term is absent
    (src/PhysicsCloth.cpp:3062-3066):
    kx L / (l1 l3 sin phi) ((1 - (phi - pi/2)) P1 d3 d1^T P3 / sin phi - (phi - pi/2) P1 P3)."""
    spec = cs.homo_spec(3)
    spec.update(wind_density=0.0, wind_drag=0.0)
    st = excited_state(cs, spec, 7, pos_jit=2e-4, vel=0.0, evel=0.0)
    kx = spec["S"] * PI * spec["R"] ** 2
    for q, (n1, n3) in enumerate(((3, 7), (3, 1), (5, 7), (5, 1))):
        K = _one(spec, st, "shear", q, ["shear"]).dense(2)
        fd = _fd(spec, st, "shear", q, ["shear"], (4, n1, n3), eps_x=1e-9)
        x = st["LPos"]
        a1, a3 = x[n1] - x[4], x[n3] - x[4]
        l1, l3 = np.linalg.norm(a1), np.linalg.norm(a3)
        d1, d3 = a1 / l1, a3 / l3
        phi = np.arccos(d1 @ d3)
        pb = phi - PI / 2
        P1, P3 = np.eye(3) - np.outer(d1, d1), np.eye(3) - np.outer(d3, d3)
        dev13 = kx * spec["L"] / (l1 * l3 * np.sin(phi)) * (
            (1 - pb) * np.outer(P1 @ d3, d1) @ P3 / np.sin(phi) - pb * P1 @ P3)
        dev31 = kx * spec["L"] / (l1 * l3 * np.sin(phi)) * (
            (1 - pb) * np.outer(P3 @ d1, d3) @ P1 / np.sin(phi) - pb * P3 @ P1)
        for kk in (n1, n3):
            assert _rel(_llT(K, kk, kk), _fdblk(fd, kk, kk, spec)) < 1e-6
        assert _rel(_llT(K, n1, n3), _fdblk(fd, n1, n3, spec) + dev13) < 1e-5
        assert _rel(_llT(K, n3, n1), _fdblk(fd, n3, n1, spec) + dev31) < 1e-5
