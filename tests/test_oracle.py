"""Pins the CPU oracle (oracle/yarn_oracle.cpp) against what the mathematics and the
reference's own data fix, independently of the HIP product.

- structure: DOF and stencil nonzero counts (closed forms), element counts, the
  libgomp static partition used for the friction input (brute force)
- exact derivatives: the Jacobian blocks the reference derives exactly (stretch,
  gravity, handles w.r.t. positions; inertia and M-dot w.r.t. velocities; bending
  cross blocks and LE) against central finite differences of the oracle's force
- invariants: translation invariance and zero net internal force; rotation
  equivariance of forces and normals; closed-form mass and gravity totals
- the boundary rule of the reference (cross terms of an edge are stored only when
  both ends are interior): the stretch-only stiffness then has the eigenvalue
  (sqrt(2)-1) ks / L, which a dense re-derivation predicts
- normals against an SVD of the centred five-point matrix (LAPACK via numpy)
- CG against a dense solve and a textbook Jacobi-PCG restatement
- the reference's shipped 5x5 trajectory (TrainingData/Table_2), loosely
- the non-jamming shear force against the gradient of its energy, and its
  self blocks against central differences
"""
import copy
import os

import numpy as np
import pytest

from oracle.oracle import OracleCloth, owner
from tests.golden_cases import excited_state

HERE = os.path.dirname(os.path.abspath(__file__))


def _spec(cs, W=5, **kw):
    s = cs.homo_spec(W)
    for k, v in kw.items():
        if k in ("Y", "B", "rho"):
            for y in s["yarns"]:
                y[k] = v
        else:
            s[k] = v
    return s


def _forces(spec, state, h=1e-3):
    c = OracleCloth(spec)
    c.set_state(state)
    c.assemble(h)
    return c


def test_structure_counts(cs):
    for W in (5, 6, 9, 17):
        c = _forces(cs.homo_spec(W), cs.initial_state(cs.homo_spec(W)))
        n = 3 * W * W + 2 * (W - 2) ** 2
        assert c.n == n
        nnz = 9 * (13 * W * W - 20 * W - 4) + 14 * (W - 2) * (3 * W - 8)
        assert c.csr(0).nnz == nnz
    assert _forces(cs.homo_spec(5), cs.initial_state(cs.homo_spec(5))).csr(0).nnz == 2283


def test_static_partition_brute_force():
    # libgomp schedule(static): thread t owns a contiguous chunk; the first n%T threads get one more
    for n in (1, 7, 8, 9, 40, 95, 2283):
        T = 8
        q, r = divmod(n, T)
        exp = []
        for t in range(T):
            cnt = q + (1 if t < r else 0)
            exp += [t] * cnt
        assert [owner(n, i) for i in range(n)] == exp


def test_rest_state_forces_are_gravity(cs):
    """At the constructor's rest pose (rest lengths, straight yarns, right angles, zero
    velocity, no wind) the only L force is gravity: -rho G Delta/2 per edge end."""
    W = 6
    s = _spec(cs, W, wind_density=0.0, wind_drag=0.0)
    c = _forces(s, cs.initial_state(s))
    F = c.vec("F")[: 3 * W * W].reshape(-1, 3)
    rho, L, g = s["yarns"][0]["rho"], s["L"], s["G"][2]
    nedge = np.zeros(W * W)
    for v in range(W):
        for u in range(W):
            nedge[u + v * W] = (u > 0) + (u < W - 1) + (v > 0) + (v < W - 1)
    # positions of magnitude `hight` = 5 m on a 1 mm grid carry strains of ~1e-13 from
    # rounding alone, i.e. forces of ks * 1e-13 ~ 4e-14 N per edge
    np.testing.assert_allclose(F[:, 2], -0.5 * rho * g * L * nedge, rtol=0, atol=1e-11)
    np.testing.assert_allclose(F[:, :2], 0.0, atol=1e-11)


def test_mass_closed_form(cs):
    """The assembled M is symmetric and its L-L part sums to rho * (total yarn length)
    per component (src/PhysicsCloth.cpp:257-270)."""
    s = _spec(cs, 6)
    st = excited_state(cs, s, 1)
    c = _forces(s, st)
    M = c.dense(1)
    W = 6
    assert np.abs(M - M.T).max() < 1e-18
    N = W * W
    tot = 0.0
    for v in range(W):
        for u in range(W):
            k = u + v * W
            if u < W - 1:
                tot += s["yarns"][0]["rho"] * (st["EPos"][k + 1, 0] - st["EPos"][k, 0])
            if v < W - 1:
                tot += s["yarns"][0]["rho"] * (st["EPos"][k + W, 1] - st["EPos"][k, 1])
    Mx = M[0:3 * N:3, 0:3 * N:3]
    assert abs(Mx.sum() - tot) < 1e-12 * tot


def _strip(cs, W, **kw):
    """Cloth with only the named terms left on."""
    s = _spec(cs, W, wind_density=0.0, wind_drag=0.0, kf=0.0, df=0.0, mu=0.0, kc=0.0)
    for k, v in kw.items():
        if k in ("Y", "B"):
            for y in s["yarns"]:
                y[k] = v
        else:
            s[k] = v
    return s


def _fd(spec, state, which, eps, cols=None):
    c = _forces(spec, state)
    K = c.dense(2 if which == "x" else 3)
    W = spec["width"]
    N = W * W
    n = c.n
    Kfd = np.zeros((n, n))
    for j in (range(n) if cols is None else cols):
        if j < 3 * N:
            k, a = divmod(j, 3)
            key = "LPos" if which == "x" else "LVel"
        else:
            e, a = divmod(j - 3 * N, 2)
            k = (e % (W - 2) + 1) + (e // (W - 2) + 1) * W
            key = "EPos" if which == "x" else "EVel"
        s1, s2 = copy.deepcopy(state), copy.deepcopy(state)
        s1[key][k, a] += eps
        s2[key][k, a] -= eps
        Kfd[:, j] = (_forces(spec, s1).vec("F") - _forces(spec, s2).vec("F")) / (2 * eps)
    return K, Kfd, N


def test_shear_non_jamming_force_and_outer_blocks(cs):
    """Shear alone (no stretch or bending, so the compression, and with it the shear
    stiffness kx = S pi R^2, is constant) on a 3 x 3 plain cloth: its one crossing has
    four non-jamming shear elements (the crossing and its four neighbours have equal
    weave flips, src/ClothUnit.h:226-230).  Their force is minus the gradient of
    E = sum 1/2 kx L (phi - pi/2)^2, phi the angle between the two yarn segments at the
    crossing, taken here by central differences of E written from that definition.  The
    Jacobian blocks of each outer node with itself are exact derivatives of the force;
    the blocks that couple two nodes are the reference's hand-derived values
    (src/PhysicsCloth.cpp:3056-3088) and differ from the derivative."""
    W = 3
    s = _strip(cs, W, Y=0.0, B=0.0, G=[0.0, 0.0, 0.0], handle_k=0.0)
    st = excited_state(cs, s, 7, pos_jit=2e-4, vel=0.0, evel=0.0)
    c = _forces(s, st)
    assert np.all(c.crossing()[1] == 0.0)
    kx = s["S"] * np.pi * s["R"] ** 2
    centre, left, right, top, bottom = 4, 3, 5, 1, 7

    def energy(X):
        e = 0.0
        for a, b in ((left, bottom), (left, top), (right, bottom), (right, top)):
            d1, d3 = X[a] - X[centre], X[b] - X[centre]
            phi = np.arccos(d1 @ d3 / (np.linalg.norm(d1) * np.linalg.norm(d3)))
            e += 0.5 * kx * s["L"] * (phi - np.pi / 2) ** 2
        return e

    X = st["LPos"]
    grad = np.zeros_like(X)
    for i in range(W * W):
        for a in range(3):
            Xp, Xm = X.copy(), X.copy()
            Xp[i, a] += 1e-9
            Xm[i, a] -= 1e-9
            grad[i, a] = (energy(Xp) - energy(Xm)) / 2e-9
    F = c.vec("F")
    assert np.linalg.norm(grad) > 1e-4
    assert np.linalg.norm(F[: 3 * W * W].reshape(-1, 3) + grad) < 1e-6 * np.linalg.norm(grad)
    assert np.all(F[3 * W * W:] == 0.0)
    K, Kfd, N = _fd(s, st, "x", 1e-9)
    for k in (left, right, top, bottom):
        blk = slice(3 * k, 3 * k + 3)
        assert np.linalg.norm(K[blk, blk] - Kfd[blk, blk]) < 1e-6 * np.linalg.norm(Kfd[blk, blk])
    blk1, blk3 = slice(3 * left, 3 * left + 3), slice(3 * bottom, 3 * bottom + 3)
    assert np.linalg.norm(K[blk1, blk3] - Kfd[blk1, blk3]) > 0.5 * np.linalg.norm(Kfd[blk1, blk3])


def _masked_rel(K, Kfd, rows, cols):
    m = K[rows, cols] != 0
    a, b = K[rows, cols][m], Kfd[rows, cols][m]
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_fd_stretch_gravity_handle_position_jacobian(cs):
    s = _strip(cs, 5, B=0.0, S=0.0)
    st = excited_state(cs, s, 3, vel=0.0, evel=0.0)
    K, Kfd, N = _fd(s, st, "x", 1e-8)
    L, E = slice(0, 3 * N), slice(3 * N, None)
    for r, c in [(L, L), (L, E), (E, L), (E, E)]:
        assert _masked_rel(K, Kfd, r, c) < 1e-6


def test_fd_inertia_velocity_jacobian(cs):
    s = _strip(cs, 5, Y=0.0, B=0.0, S=0.0, G=[0.0, 0.0, 0.0], handle_k=0.0)
    st = excited_state(cs, s, 4)
    K, Kfd, N = _fd(s, st, "v", 1e-6)
    L, E = slice(0, 3 * N), slice(3 * N, None)
    for r, c in [(L, L), (L, E), (E, L), (E, E)]:
        assert _masked_rel(K, Kfd, r, c) < 1e-6


def test_fd_bending_cross_blocks(cs):
    """Bending's n1-n2 blocks (two nodes apart, transposed by the LL placement) and its
    LE entries are exact derivatives of Bending_F; its diagonal, EL and EE entries are the reference's
    hand-derived values and are pinned by the GPU/oracle agreement only."""
    W = 5
    s = _strip(cs, W, Y=0.0, S=0.0, G=[0.0, 0.0, 0.0], handle_k=0.0)
    st = excited_state(cs, s, 5, vel=0.0, evel=0.0)
    K, Kfd, N = _fd(s, st, "x", 1e-9)
    # distance-2 LL blocks along both yarn directions
    errs = []
    for v in range(W):
        for u in range(W):
            for du, dv in ((2, 0), (0, 2)):
                if u + du < W and v + dv < W:
                    F, D = u + v * W, (u + du) + (v + dv) * W
                    # LL blocks enter A transposed (src/MatrixMp.cpp:65-76)
                    a = K[3 * F:3 * F + 3, 3 * D:3 * D + 3].T
                    b = Kfd[3 * F:3 * F + 3, 3 * D:3 * D + 3]
                    errs.append(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
    assert max(errs) < 1e-5
    assert _masked_rel(K, Kfd, slice(0, 3 * N), slice(3 * N, None)) < 1e-6


def test_boundary_rule_stretch_hessian(cs):
    """Cross blocks of an edge are stored only when both ends are interior
    (src/PhysicsCloth.cpp:310, 539).  At rest, the exact Hessian of one edge's energy
    1/2 ks D (|w| - 1)^2 over (x0, u0, x1, u1) is (ks/D) g g^T with
    g = (-d, 1, d, -1); the oracle's stretch-only K_x must equal minus the sum of these
    with every boundary coordinate dropped and the x_boundary-u_interior terms removed."""
    W = 4
    s = _strip(cs, W, B=0.0, S=0.0, G=[0.0, 0.0, 0.0], handle_k=0.0)
    c = _forces(s, cs.initial_state(s))
    K = c.dense(2)
    N, n = W * W, c.n
    ks = s["yarns"][0]["Y"] * 3.1415926535897 * s["R"] ** 2
    D = s["L"]
    Kp = np.zeros((n, n))
    inter = lambda u, v: 0 < u < W - 1 and 0 < v < W - 1
    for v in range(W):
        for u in range(W):
            for c_, (du, dv) in enumerate(((1, 0), (0, 1))):
                u1, v1 = u + du, v + dv
                if u1 >= W or v1 >= W:
                    continue
                k0, k1 = u + v * W, u1 + v1 * W
                d = np.array([1.0, 0.0, 0.0]) if c_ == 0 else np.array([0.0, 0.0, -1.0])  # upright pose
                idx, g = [], []
                for a in range(3):
                    idx.append(3 * k0 + a); g.append(-d[a])
                e0 = 3 * N + 2 * ((u - 1) + (v - 1) * (W - 2)) + c_ if inter(u, v) else None
                e1 = 3 * N + 2 * ((u1 - 1) + (v1 - 1) * (W - 2)) + c_ if inter(u1, v1) else None
                for a in range(3):
                    idx.append(3 * k1 + a); g.append(d[a])
                H = np.outer(g, g)
                for i, ri in enumerate(idx):
                    for j, cj in enumerate(idx):
                        Kp[ri, cj] -= ks / D * H[i, j]
                both = e0 is not None and e1 is not None
                for (ee, ge, kk) in ((e0, 1.0, k0), (e1, -1.0, k1)):
                    if ee is None:
                        continue
                    Kp[ee, ee] -= ks / D
                    for (kl, gl) in ((k0, -d), (k1, d)):
                        if kl != kk and not both:
                            continue
                        for a in range(3):
                            Kp[3 * kl + a, ee] -= ks / D * gl[a] * ge
                            Kp[ee, 3 * kl + a] -= ks / D * gl[a] * ge
                if both:
                    Kp[e0, e1] += ks / D
                    Kp[e1, e0] += ks / D
    assert np.abs(K - Kp).max() < 1e-9 * np.abs(Kp).max()


def _rotation(seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def test_rotation_and_translation_invariance(cs):
    """Internal forces (no gravity, wind or handles) rotate with the cloth, E forces
    and the compression are invariant, and a translation changes nothing.  Shear is
    off: the reference's jamming-shear force adds (1 - gamma^c) (1, 1, 1) through the
    broadcast in (1 - con_func^2 * con_func_x_x1) (src/PhysicsCloth.cpp:2699), a
    world-frame vector, so that term is not rotation-equivariant by construction."""
    s = cs.heter_spec(6, "twill", 3)
    # keep coordinates O(W L) so that rotating them costs no significant digits
    s.update(G=[0.0, 0.0, 0.0], wind_density=0.0, wind_drag=0.0, handle_k=0.0, hight=0.0, S=0.0)
    st = excited_state(cs, s, 6)
    Q = _rotation(1)
    st2 = copy.deepcopy(st)
    for k in ("LPos", "LPosFix", "LVel"):
        st2[k] = st[k] @ Q.T
    st3 = copy.deepcopy(st)
    st3["LPos"] = st["LPos"] + np.array([3e-3, -2e-3, 7e-3])
    st3["LPosFix"] = st["LPosFix"] + np.array([3e-3, -2e-3, 7e-3])
    c1, c2, c3 = _forces(s, st), _forces(s, st2), _forces(s, st3)
    N = 36
    F1, F2, F3 = c1.vec("F"), c2.vec("F"), c3.vec("F")
    FL1 = F1[:3 * N].reshape(-1, 3)
    np.testing.assert_allclose(F2[:3 * N].reshape(-1, 3), FL1 @ Q.T, atol=1e-9 * np.abs(FL1).max())
    np.testing.assert_allclose(F2[3 * N:], F1[3 * N:], atol=1e-9 * np.abs(F1[3 * N:]).max())
    np.testing.assert_allclose(F3, F1, atol=1e-8 * np.abs(F1).max())
    np.testing.assert_allclose(c2.normals(), c1.normals() @ Q.T, atol=1e-10)
    np.testing.assert_allclose(c2.crossing()[1], c1.crossing()[1], rtol=1e-8, atol=1e-14)


def test_zero_net_internal_force(cs):
    s = cs.heter_spec(7, "satin", 4)
    s.update(G=[0.0, 0.0, 0.0], wind_density=0.0, wind_drag=0.0, handle_k=0.0)
    st = excited_state(cs, s, 8, vel=0.0, evel=0.0)
    c = _forces(s, st)
    FL = c.vec("F")[:3 * 49].reshape(-1, 3)
    assert np.abs(FL.sum(0)).max() < 1e-10 * np.abs(FL).max()


def test_normals_against_svd(cs):
    s = cs.heter_spec(7, "twill", 5)
    st = excited_state(cs, s, 9)
    c = _forces(s, st)
    nr = c.normals()
    W = 7
    x = st["LPos"]
    for v in range(1, W - 1):
        for u in range(1, W - 1):
            k = u + v * W
            P = x[[k, k - 1, k + 1, k - W, k + W]]
            _, _, Vt = np.linalg.svd(P - P.mean(0))
            n = Vt[2] / np.linalg.norm(Vt[2])
            cr = np.cross(x[k + 1] - x[k - 1], x[k + W] - x[k - W])
            if cr @ n < 0:
                n = -n
            flip = {"twill": lambda i, j: ((i + j % 4) % 4) < 2}["twill"](u, v)
            if flip:
                n = -n
            np.testing.assert_allclose(nr[k], n, atol=1e-10)


def _pcg(A, b, tol, maxit):
    """Textbook Jacobi-PCG in the order of the published Eigen ConjugateGradient."""
    d = A.diagonal().copy()
    inv = np.where(d != 0, 1.0 / np.where(d != 0, d, 1.0), 1.0)
    x = np.zeros_like(b)
    r = b.copy()
    thr = max(tol * tol * (b @ b), np.finfo(float).tiny)
    if r @ r < thr:
        return x, 0
    p = inv * r
    absn = r @ p
    i = 0
    while i < maxit:
        t = A @ p
        a = absn / (p @ t)
        x += a * p
        r -= a * t
        if r @ r < thr:
            break
        z = inv * r
        abso, absn = absn, r @ z
        p = z + (absn / abso) * p
        i += 1
    return x, i


def test_cg_against_dense_and_textbook(cs):
    s = cs.heter_spec(6, "plain", 7)
    c = _forces(s, cs.initial_state(s))
    A, b = c.dense(0), c.vec("b")
    x, it, rr, st = c.solve(tol=1e-12, maxit=5000)
    assert st == 0 and rr < 1e-12
    xd = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8
    xt, itt = _pcg(A, b, 1e-12, 5000)
    # same recurrence; summation order of the products differs, which on this
    # nonsymmetric system can move the stopping iteration by a few
    assert abs(itt - it) <= 4
    assert np.linalg.norm(x - xt) / np.linalg.norm(xt) < 1e-10


def test_reference_fixture_trajectory(cs):
    """The reference's shipped 5x5 trajectory (TrainingData/Table_2/
    Wind_Yarn_1_2_Plain_Size5.h5, converted by tests/golden/make_fixture.py).  It was
    written by an earlier revision of the friction code (its EPosBar never moves,
    which the shipped soft is_slide cannot produce), so the check is loose: after the
    first 20 steps the displacement agrees within 5e-3 of its size."""
    d = np.load(os.path.join(HERE, "golden", "wind_yarn12_plain5.npz"))
    P = dict(zip([str(x) for x in d["prop_names"]], d["props"]))
    s = cs.base_spec(5)
    s["yarns"] = [{"rho": P["Yarn%d_rho" % i], "Y": P["Yarn%d_Y" % i], "B": P["Yarn%d_B" % i]} for i in (1, 2)]
    for k in ("kc", "kf", "df", "mu", "S"):
        s[k] = P[k]
    s["handle_k"] = P["handle_stiff"]
    s["wind_vel"] = [P["wind_velocity_x"], P["wind_velocity_y"], P["wind_velocity_z"]]
    s["wind_density"], s["wind_drag"] = P["wind_density"], P["wind_drag"]
    c = OracleCloth(s)
    c.set_state({"LPos": d["LPOS"][0], "LPosFix": d["LPOSFix"][0], "LVel": d["LVEL"][0],
                 "EPos": d["EPOS"][0], "EPosBar": d["EPOSBar"][0], "EVel": d["EVEL"][0]})
    rows = d["LPOS"].shape[0]
    for step in range(1, rows):
        st, _, _ = c.step(1e-3, tol=1e-12, maxit=20000)
        assert st == 0
        if step >= 20 and step % 20 == 0:
            x = c.get_state()["LPos"]
            disp = d["LPOS"][step] - d["LPOS"][0]
            assert np.linalg.norm(x - d["LPOS"][step]) < 5e-3 * np.linalg.norm(disp)


def test_update_rule(cs):
    """Cloth::update (src/Cloth.cpp:111-148): x += h v_new, xfix follows free nodes,
    uv += h uvd, uvbar += is_slide * uvd * h."""
    s = cs.heter_spec(5, "plain", 9)
    st = cs.initial_state(s)
    c = OracleCloth(s)
    c.set_state(st)
    c.assemble(1e-3)
    x, _, _, _ = c.solve(tol=1e-12, maxit=5000)
    sl, _, _ = c.crossing()
    c.update(x, 1e-3)
    g = c.get_state()
    N = 25
    np.testing.assert_allclose(g["LPos"], st["LPos"] + 1e-3 * x[:3 * N].reshape(-1, 3), rtol=0, atol=1e-15)
    free = np.ones(N, bool)
    free[[0, 4]] = False
    np.testing.assert_array_equal(g["LPosFix"][free], g["LPos"][free])
    np.testing.assert_array_equal(g["LPosFix"][~free], st["LPosFix"][~free])
    k = 1 + 1 * 5
    ud = x[3 * N:3 * N + 2]
    np.testing.assert_allclose(g["EPosBar"][k], st["EPosBar"][k] + sl[0] * ud * 1e-3, rtol=0, atol=1e-18)


@pytest.mark.parametrize("W,Lg", [(7, 5), (5, 7)])
def test_non_square_rest_state_and_state_order(cs, W, Lg):
    """Non-square cloths: at the rest pose the only L force is gravity (as for square
    ones), and set/get state use index u + v*width (the reference's u + v*length for
    the square cloths it supports) -- a round trip returns every node's own values."""
    s = _spec(cs, W, wind_density=0.0, wind_drag=0.0)
    s["length"] = Lg
    st = cs.initial_state(s)
    N = W * Lg
    assert st["LPos"].shape == (N, 3)
    for v in range(Lg):
        for u in range(W):
            assert np.allclose(st["LPos"][u + v * W], (u * s["L"], 0.0, s["hight"] - v * s["L"]))
    c = _forces(s, st)
    F = c.vec("F")[: 3 * N].reshape(-1, 3)
    rho, L, g = s["yarns"][0]["rho"], s["L"], s["G"][2]
    nedge = np.array([(u > 0) + (u < W - 1) + (v > 0) + (v < Lg - 1) for v in range(Lg) for u in range(W)])
    np.testing.assert_allclose(F[:, 2], -0.5 * rho * g * L * nedge, rtol=0, atol=1e-11)
    np.testing.assert_allclose(F[:, :2], 0.0, atol=1e-11)
    rng = np.random.default_rng(W * 10 + Lg)
    mark = {k: rng.normal(size=v.shape) for k, v in st.items()}
    c.set_state(mark)
    back = c.get_state()
    for k in ("LPos", "LPosFix", "LVel"):
        assert np.array_equal(back[k], mark[k])
    inter = np.array([0 < u < W - 1 and 0 < v < Lg - 1 for v in range(Lg) for u in range(W)])
    for k in ("EPos", "EPosBar", "EVel"):
        assert np.array_equal(back[k][inter], mark[k][inter])


def test_parallel_assembly_and_dense_layout_match_serial(cs):
    """The oracle's multi-threaded assembly (each thread owns grid rows and walks the
    element lists in the reference's order) is bit for bit the single-thread one, and the
    dense-layout cost model (the reference's per-thread N^2 block arrays, dense
    To_Sparse_Mat matrices and n^2 scan) steps to the same state as the sparse path."""
    s = cs.heter_spec(9, "satin", 31)
    st = excited_state(cs, s, 32)
    out = []
    for threads in (1, 3, 8):
        c = OracleCloth(s)
        c.set_threads(threads)
        c.set_state(st)
        c.assemble(1e-3)
        out.append((c.csr(1), c.csr(2), c.csr(3), c.vec("F"), c.vec("b"), c.crossing(), c.normals()))
    for o in out[1:]:
        for x, y in zip(o[:3], out[0][:3]):
            assert (x != y).nnz == 0 and np.array_equal(x.indices, y.indices)
        for x, y in zip(o[3:5] + (o[6],), out[0][3:5] + (out[0][6],)):
            assert np.array_equal(x, y)
        for x, y in zip(o[5], out[0][5]):
            assert np.array_equal(x, y)
    s0 = cs.initial_state(s)
    a, b = OracleCloth(s), OracleCloth(s)
    a.set_state(s0)
    b.set_state(s0)
    for _ in range(3):
        ra = a.step(1e-3, tol=1e-10, maxit=5000)
        rb = b.step_dense(1e-3, tol=1e-10, maxit=5000)
        assert ra == rb and ra[0] == 0
    sa, sb = a.get_state(), b.get_state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
