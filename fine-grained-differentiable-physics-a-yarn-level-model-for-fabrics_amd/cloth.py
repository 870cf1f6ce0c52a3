"""Host mirror of the reference step API over the C ABI.

``Cloth`` owns one device context (src/Cloth.h:31-82 + src/PhysicsCloth.h:6-78
collapsed into one object, since the HIP context holds both the state and the
system); ``PhysicsCloth`` is a thin view that keeps the reference's method names
(FillMV, get_SpaM-style exports); ``SimStep`` mirrors src/TrainHelper.cpp:15-97 and
returns False on NaN/Inf as the reference does.
"""
import ctypes

import numpy as np

from . import _abi

_dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
_NULL = ctypes.POINTER(ctypes.c_double)()


def _arr(a, shape):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.size != int(np.prod(shape)):
        raise ValueError("expected %s doubles, got %d" % (shape, a.size))
    return a


def configure_solver(cloth, cg="persistent", precond="block2"):
    """Selects the CG form and preconditioner for a rollout, falling back to the fused form
    with the Jacobi preconditioner where the persistent one does not fit the device."""
    try:
        cloth.set_cg_variant(cg)
    except _abi.YcError:
        cloth.set_cg_variant("fused")
        return
    if cg == "persistent" and precond != "jacobi":
        try:
            cloth.set_preconditioner(precond)
        except _abi.YcError:
            pass


class Cloth:
    """One cloth on the current HIP device (reference Cloth + PhysicsCloth)."""

    def __init__(self, spec):
        L = _abi.load_library()
        self.spec = spec
        self._desc = _abi.make_desc(spec)
        h = ctypes.c_void_p()
        _abi.check(L.yc_create(ctypes.byref(self._desc), ctypes.byref(h)))
        self._h = h
        N, Nc, n, ns = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _abi.check(L.yc_sizes(h, ctypes.byref(N), ctypes.byref(Nc), ctypes.byref(n), ctypes.byref(ns)))
        self.N, self.Nc, self.n, self.nslots = N.value, Nc.value, n.value, ns.value
        self.width, self.length = spec["width"], spec["length"]

    def close(self):
        if getattr(self, "_h", None):
            _abi.load_library().yc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- src/Cloth.cpp:350-438 ---
    def get_ClothState(self):
        N = self.N
        s = {k: np.zeros((N, 3)) for k in ("LPos", "LPosFix", "LVel")}
        s.update({k: np.zeros((N, 2)) for k in ("EPos", "EPosBar", "EVel")})
        _abi.check(_abi.load_library().yc_get_state(self._h, _dp(s["LPos"]), _dp(s["LPosFix"]), _dp(s["LVel"]),
                                                    _dp(s["EPos"]), _dp(s["EPosBar"]), _dp(s["EVel"])))
        return s

    def set_ClothState(self, s):
        N = self.N
        a = [_arr(s[k], (N, 3)) for k in ("LPos", "LPosFix", "LVel")] + \
            [_arr(s[k], (N, 2)) for k in ("EPos", "EPosBar", "EVel")]
        _abi.check(_abi.load_library().yc_set_state(self._h, *[_dp(x) for x in a]))

    # --- state in device memory (torch tensors on the context's device) ---
    def set_ClothState_device(self, s, stream=None):
        """s: the get_ClothState keys as contiguous float64 CUDA tensors ((N, 3) / (N, 2));
        the copy is ordered on `stream` (default: torch's current stream)."""
        import torch
        t = [s[k] for k in self.STATE_KEYS]
        for k, x in zip(self.STATE_KEYS, t):
            if not (x.is_cuda and x.dtype == torch.float64 and x.is_contiguous()) or x.numel() != self.N * (3 if k[0] == "L" else 2):
                raise ValueError("%s: expected a contiguous float64 CUDA tensor of %d x %d" % (k, self.N, 3 if k[0] == "L" else 2))
        st = (stream or torch.cuda.current_stream()).cuda_stream
        _abi.check(_abi.load_library().yc_set_state_device(self._h, *[x.data_ptr() for x in t], st))

    def get_ClothState_device(self, stream=None):
        """The state as float64 CUDA tensors, written in `stream` order (default: torch's
        current stream)."""
        import torch
        st_ = stream or torch.cuda.current_stream()
        out = {k: torch.empty((self.N, 3 if k[0] == "L" else 2), dtype=torch.float64, device="cuda") for k in self.STATE_KEYS}
        _abi.check(_abi.load_library().yc_get_state_device(self._h, *[out[k].data_ptr() for k in self.STATE_KEYS], st_.cuda_stream))
        return out

    def set_ext_force(self, f):
        if f is None:
            _abi.check(_abi.load_library().yc_set_ext_force(self._h, _NULL))
        else:
            self._ext = _arr(f, (self.n,))
            _abi.check(_abi.load_library().yc_set_ext_force(self._h, _dp(self._ext)))

    # --- FillMV + system (src/PhysicsCloth.cpp:46-156, src/TrainHelper.cpp:19-20) ---
    def assemble(self, h):
        _abi.check(_abi.load_library().yc_assemble(self._h, float(h)))

    MATRICES = {"A": 0, "M": 1, "Kx": 2, "Kv": 3}

    def csr(self, which="A"):
        """A (as assembled), or M, K_x or K_v alone (get_SpaM / get_SpaFPos / get_SpaFVel)
        -> scipy.sparse.csr_matrix in the reference's global layout."""
        import scipy.sparse as sp
        L = _abi.load_library()
        nnz = ctypes.c_int()
        ip = ctypes.POINTER(ctypes.c_int)
        kind = self.MATRICES[which]
        _abi.check(L.yc_export_matrix(self._h, kind, ip(), ip(), _NULL, ctypes.byref(nnz)))
        rp = np.zeros(self.n + 1, dtype=np.int32)
        cl = np.zeros(nnz.value, dtype=np.int32)
        val = np.zeros(nnz.value)
        _abi.check(L.yc_export_matrix(self._h, kind, rp.ctypes.data_as(ip), cl.ctypes.data_as(ip), _dp(val),
                                      ctypes.byref(nnz)))
        return sp.csr_matrix((val, cl, rp), shape=(self.n, self.n))

    VECTORS = {"F": 0, "b": 1, "v": 2, "x": 3, "pos": 4}

    def vec(self, which):
        out = np.zeros(self.n)
        _abi.check(_abi.load_library().yc_export_vec(self._h, self.VECTORS[which], _dp(out)))
        return out

    def load_system(self, A, b):
        """Makes a caller-formed system the context's (EigenSolver::apply's input side):
        A a scipy sparse matrix inside the yarn stencil, b an n-vector."""
        import scipy.sparse as sp
        A = sp.csr_matrix(A)
        if A.shape != (self.n, self.n):
            raise ValueError("expected an %d x %d matrix" % (self.n, self.n))
        ip = ctypes.POINTER(ctypes.c_int)
        rp = np.ascontiguousarray(A.indptr, dtype=np.int32)
        cl = np.ascontiguousarray(A.indices, dtype=np.int32)
        val = np.ascontiguousarray(A.data, dtype=np.float64)
        bb = _arr(b, (self.n,))
        _abi.check(_abi.load_library().yc_load_system(self._h, rp.ctypes.data_as(ip), cl.ctypes.data_as(ip), _dp(val),
                                                       _dp(bb)))

    def update_vel(self, new_vel, is_slide, h):
        """Cloth::update(new_vel, F_E_is_Slide, h) (src/Cloth.cpp:111-148) with a caller's
        velocity; is_slide (Nc, 2) or None for the last assembly's."""
        nv = _arr(new_vel, (self.n,))
        sl = _NULL if is_slide is None else _dp(_arr(is_slide, (self.Nc, 2)))
        _abi.check(_abi.load_library().yc_update_vel(self._h, _dp(nv), sl, float(h)))

    def crossing(self):
        sl, cp, fp = np.zeros(2 * self.Nc), np.zeros(self.Nc), np.zeros(2 * self.Nc)
        _abi.check(_abi.load_library().yc_export_crossing(self._h, _dp(sl), _dp(cp), _dp(fp)))
        return sl.reshape(-1, 2), cp, fp.reshape(-1, 2)

    def normals(self):
        out = np.zeros((self.N, 3))
        _abi.check(_abi.load_library().yc_export_normals(self._h, _dp(out)))
        return out

    # --- EigenSolver::apply forward (src/Solvers.cpp:3-31) ---
    def solve(self, tol=1e-8, maxit=None, x0=None, raise_on_nonconv=False):
        if maxit is None:
            maxit = 2 * self.n
        it, rr = ctypes.c_int(), ctypes.c_double()
        x0a = _NULL if x0 is None else _dp(_arr(x0, (self.n,)))
        rc = _abi.load_library().yc_solve(self._h, x0a, float(tol), int(maxit), ctypes.byref(it), ctypes.byref(rr))
        if rc == _abi.YC_ERR_NOT_CONVERGED and not raise_on_nonconv:
            return it.value, rr.value, False
        _abi.check(rc)
        return it.value, rr.value, True

    # --- Cloth::update (src/Cloth.cpp:111-148) ---
    def update(self, h):
        _abi.check(_abi.load_library().yc_update(self._h, float(h)))

    def step(self, h, tol=1e-8, maxit=None, x0=None):
        """SimStep; returns (iters, relres, status)."""
        if maxit is None:
            maxit = 2 * self.n
        info = _abi.YcStepInfo()
        x0a = _NULL if x0 is None else _dp(_arr(x0, (self.n,)))
        rc = _abi.load_library().yc_step(self._h, float(h), x0a, float(tol), int(maxit), ctypes.byref(info))
        if rc not in (_abi.YC_OK, _abi.YC_ERR_NOT_CONVERGED, _abi.YC_ERR_NONFINITE):
            _abi.check(rc)
        return info.cg_iters, info.cg_relres, rc

    def run(self, h, steps, tol=1e-8, maxit=None):
        if maxit is None:
            maxit = 2 * self.n
        tot = ctypes.c_longlong()
        rc = _abi.load_library().yc_run(self._h, float(h), int(steps), float(tol), int(maxit), ctypes.byref(tot))
        if rc not in (_abi.YC_OK, _abi.YC_ERR_NOT_CONVERGED):
            _abi.check(rc)
        return tot.value, rc

    GUESSES = {"zero": 0, "prev": 1, "extrap": 2, "quad": 3}

    def run_guess(self, guess, h, steps, tol=1e-8, maxit=None):
        """run() from a cheap initial guess: "prev" (v_n), "extrap" (2 v_n - v_{n-1}) or
        "quad" (3 v_n - 3 v_{n-1} + v_{n-2}); persistent CG only.  Returns (total CG
        iterations, status)."""
        if maxit is None:
            maxit = 2 * self.n
        tot = ctypes.c_longlong()
        rc = _abi.load_library().yc_run_guess(self._h, self.GUESSES[guess], float(h), int(steps), float(tol), int(maxit),
                                              ctypes.byref(tot))
        if rc not in (_abi.YC_OK, _abi.YC_ERR_NOT_CONVERGED):
            _abi.check(rc)
        return tot.value, rc

    # --- learned warm start ---
    def features(self):
        out = np.zeros((self.N, 32), dtype=np.float32)
        _abi.check(_abi.load_library().yc_features(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def targets(self):
        """After a solve: x - x_q per node (N x 5), x_q = 3 v_n - 3 v_{n-1} + v_{n-2}
        the quadratic extrapolation the network corrects."""
        out = np.zeros((self.N, 5), dtype=np.float32)
        _abi.check(_abi.load_library().yc_targets(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def time_warm_guess(self, mlp, reps=100):
        """(ms per warm guess, ms per launch of its features + network kernel), after
        assemble(); leaves the solution vector unspecified."""
        a, b = ctypes.c_double(), ctypes.c_double()
        _abi.check(_abi.load_library().yc_time_warm_guess(self._h, mlp._h, int(reps), ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def warm_guess(self, mlp):
        """After assemble(): the network's initial guess, left in the solution vector
        (vec("x")) without solving."""
        _abi.check(_abi.load_library().yc_warm_guess(self._h, mlp._h))

    def node_mask(self):
        m = np.ones((self.N, 5), dtype=np.float32)
        W, Lg = self.width, self.length
        for v in range(Lg):
            for u in range(W):
                if not (0 < u < W - 1 and 0 < v < Lg - 1):
                    m[u + v * W, 3:] = 0.0
        return m

    def step_warm(self, mlp, h, tol=1e-8, maxit=None):
        if maxit is None:
            maxit = 2 * self.n
        info = _abi.YcStepInfo()
        rc = _abi.load_library().yc_step_warm(self._h, mlp._h, float(h), float(tol), int(maxit), ctypes.byref(info))
        if rc not in (_abi.YC_OK, _abi.YC_ERR_NOT_CONVERGED, _abi.YC_ERR_NONFINITE):
            _abi.check(rc)
        return info.cg_iters, info.cg_relres, rc

    def run_warm(self, mlp, h, steps, tol=1e-8, maxit=None):
        if maxit is None:
            maxit = 2 * self.n
        tot = ctypes.c_longlong()
        rc = _abi.load_library().yc_run_warm(self._h, mlp._h, float(h), int(steps), float(tol), int(maxit), ctypes.byref(tot))
        if rc not in (_abi.YC_OK, _abi.YC_ERR_NOT_CONVERGED):
            _abi.check(rc)
        return tot.value, rc

    # --- forward-mode tangents of SimStep (differentiation) ---
    # parameter order of a direction (YC_P_* in include/yarnstep.h)
    PARAMS = ("rho0", "rho1", "Y0", "Y1", "B0", "B1", "S", "mu", "kc", "kf", "df", "handle_k")
    STATE_KEYS = ("LPos", "LPosFix", "LVel", "EPos", "EPosBar", "EVel")
    GRAD_MODES = {"ref": 0, "exact": 1}

    def tangent_begin(self, dparams, dstates=None):
        """dparams: (ndir, 12) parameter tangents; dstates: list of ndir state dicts
        (the get_ClothState keys) or None for zero state tangents."""
        dp = np.ascontiguousarray(np.atleast_2d(dparams), dtype=np.float64)
        if dp.shape[1] != len(self.PARAMS):
            raise ValueError("expected %d parameter tangents per direction" % len(self.PARAMS))
        ndir = dp.shape[0]
        ds = _NULL
        if dstates is not None:
            flat = np.concatenate([np.concatenate([np.ascontiguousarray(st[k], dtype=np.float64).ravel()
                                                   for k in self.STATE_KEYS]) for st in dstates])
            self._dstate = flat
            ds = _dp(flat)
        self._dparam = dp
        _abi.check(_abi.load_library().yc_tangent_begin(self._h, ndir, _dp(dp), ds))
        self.ndir = ndir

    def step_tangent(self, h, mode="exact", tol=1e-10, maxit=None):
        """SimStep plus the tangents of all directions through it; returns (iters, relres, status)."""
        if maxit is None:
            maxit = 2 * self.n
        info = _abi.YcStepInfo()
        rc = _abi.load_library().yc_step_tangent(self._h, float(h), self.GRAD_MODES[mode], float(tol), int(maxit),
                                                  ctypes.byref(info))
        _abi.check(rc)
        return info.cg_iters, info.cg_relres, rc

    def tangent_set_ext(self, dext):
        """Tangents of the external force, (ndir, n) in the global layout, for the following
        step_tangent calls; None clears them."""
        if dext is None:
            _abi.check(_abi.load_library().yc_tangent_set_ext(self._h, _NULL))
            return
        a = _arr(dext, (self.ndir, self.n))
        _abi.check(_abi.load_library().yc_tangent_set_ext(self._h, _dp(a)))

    def tangent_get(self, d):
        N = self.N
        s = {k: np.zeros((N, 3)) for k in ("LPos", "LPosFix", "LVel")}
        s.update({k: np.zeros((N, 2)) for k in ("EPos", "EPosBar", "EVel")})
        _abi.check(_abi.load_library().yc_tangent_get(self._h, int(d), _dp(s["LPos"]), _dp(s["LPosFix"]), _dp(s["LVel"]),
                                                      _dp(s["EPos"]), _dp(s["EPosBar"]), _dp(s["EVel"])))
        return s

    def tangent_end(self):
        _abi.check(_abi.load_library().yc_tangent_end(self._h))

    # --- reverse mode (adjoint) of a recorded rollout (differentiation) ---
    def adjoint_begin(self, max_steps, checkpoint_every=1):
        """Starts recording at the current state, for up to max_steps adjoint_step calls;
        the state is stored every checkpoint_every steps (the backward pass recomputes
        the others)."""
        _abi.check(_abi.load_library().yc_adjoint_begin(self._h, int(max_steps), int(checkpoint_every)))
        self._adj_max = int(max_steps)
        self._adj_steps = 0

    def adjoint_step(self, h, tol=1e-10, maxit=None):
        """A recorded SimStep (zero CG guess); returns (iters, relres, status).  A step whose
        solve does not converge raises (NotConverged, NonFinite) and is neither taken nor
        recorded."""
        if maxit is None:
            maxit = 2 * self.n
        info = _abi.YcStepInfo()
        rc = _abi.load_library().yc_adjoint_step(self._h, float(h), float(tol), int(maxit), ctypes.byref(info))
        _abi.check(rc)
        self._adj_steps += 1
        return info.cg_iters, info.cg_relres, rc

    def adjoint_run(self, h, steps, tol=1e-10, maxit=None):
        """steps recorded SimSteps queued without host round trips; returns the CG
        iterations.  A failure raises and voids the recording."""
        if maxit is None:
            maxit = 2 * self.n
        tot = ctypes.c_longlong()
        rc = _abi.load_library().yc_adjoint_run(self._h, float(h), int(steps), float(tol), int(maxit), ctypes.byref(tot))
        if rc != _abi.YC_OK:
            self._adj_steps = 0
        _abi.check(rc)
        self._adj_steps += int(steps)
        return tot.value

    def adjoint_set_targets(self, lpos, epos, weight):
        """Targets of the reference's position loss, (max_steps, N, 3) and (max_steps, N, 2):
        step t's loss is weight (sum |LPos - lpos[t-1]|^2 + sum |EPos - epos[t-1]|^2)."""
        if lpos is None:
            _abi.check(_abi.load_library().yc_adjoint_set_targets(self._h, _NULL, _NULL, 0.0))
            return
        tx = _arr(lpos, (self._adj_max, self.N, 3))
        tu = _arr(epos, (self._adj_max, self.N, 2))
        _abi.check(_abi.load_library().yc_adjoint_set_targets(self._h, _dp(tx), _dp(tu), float(weight)))

    def adjoint_backward(self, h, mode="ref", gstates=None, tol=1e-10, maxit=None, want_state=False, want_ext=False):
        """Backward pass over the recorded steps.  gstates: per recorded step, the loss
        gradient with respect to the state after it (a list of get_ClothState-keyed dicts,
        missing keys zero), or None for the targets' loss.  Returns a dict: "param" (the
        12 parameter derivatives, PARAMS order), "loss" (the targets' loss), "cg_iters",
        and on request "state0" (a get_ClothState-keyed dict: the gradient with respect to
        the recorded initial state) and "ext" ((steps, n): with respect to each step's
        external force)."""
        if maxit is None:
            maxit = 2 * self.n
        steps, N = self._adj_steps, self.N
        gs = _NULL
        if gstates is not None:
            if len(gstates) != steps:
                raise ValueError("expected %d state gradients" % steps)
            flat = np.zeros((steps, 15 * N))
            offs = {"LPos": (0, 3), "LPosFix": (3, 3), "LVel": (6, 3), "EPos": (9, 2), "EPosBar": (11, 2), "EVel": (13, 2)}
            for t, g in enumerate(gstates):
                for k, arr in g.items():
                    o, w = offs[k]
                    flat[t, o * N:(o + w) * N] = np.asarray(arr, dtype=np.float64).reshape(N * w)
            gs = _dp(flat)
        dparam = np.zeros(len(self.PARAMS))
        st = np.zeros(15 * N) if want_state else None
        ext = np.zeros((steps, self.n)) if want_ext else None
        loss = ctypes.c_double()
        its = ctypes.c_longlong()
        rc = _abi.load_library().yc_adjoint_backward(self._h, self.GRAD_MODES[mode], float(h), float(tol), int(maxit), gs,
                                                      _dp(dparam), _dp(st) if st is not None else _NULL,
                                                      _dp(ext) if ext is not None else _NULL, ctypes.byref(loss),
                                                      ctypes.byref(its))
        _abi.check(rc)
        out = {"param": dparam, "loss": loss.value, "cg_iters": its.value}
        if want_state:
            out["state0"] = {"LPos": st[:3 * N].reshape(N, 3), "LPosFix": st[3 * N:6 * N].reshape(N, 3),
                             "LVel": st[6 * N:9 * N].reshape(N, 3), "EPos": st[9 * N:11 * N].reshape(N, 2),
                             "EPosBar": st[11 * N:13 * N].reshape(N, 2), "EVel": st[13 * N:15 * N].reshape(N, 2)}
        if want_ext:
            out["ext"] = ext
        return out

    def adjoint_end(self):
        _abi.check(_abi.load_library().yc_adjoint_end(self._h))

    # CG variants: "reference" (the reference solver's three-kernel order), "fused" (the
    # same recurrences with one kernel and one reduction per iteration) or "persistent"
    # (the fused recurrences in one launch per solve)
    CG_VARIANTS = {"reference": 0, "fused": 1, "persistent": 2}

    def set_cg_variant(self, name):
        _abi.check(_abi.load_library().yc_set_cg_variant(self._h, self.CG_VARIANTS[name]))

    # preconditioners: "jacobi" (the reference solver's diagonal one) or "block2" (exact
    # inverses of A on 2 x 2 groups of nodes; persistent CG only)
    PRECONDITIONERS = {"jacobi": 0, "block2": 1}

    def set_preconditioner(self, name):
        _abi.check(_abi.load_library().yc_set_preconditioner(self._h, self.PRECONDITIONERS[name]))

    def precond_rows(self):
        """The block-Jacobi rows of the assembled A, (n, 20) (see yc_export_precond)."""
        out = np.zeros((self.n, 20))
        _abi.check(_abi.load_library().yc_export_precond(self._h, _dp(out)))
        return out

    def time_spmv(self, reps=200):
        ms = ctypes.c_double()
        _abi.check(_abi.load_library().yc_time_spmv(self._h, int(reps), ctypes.byref(ms)))
        return ms.value

    def time_solve(self, tol=1e-8, maxit=None, reps=5):
        """Mean duration (ms) of a persistent-CG solve of the assembled system, and its
        iteration count."""
        if maxit is None:
            maxit = 2 * self.n
        ms, it = ctypes.c_double(), ctypes.c_int()
        _abi.check(_abi.load_library().yc_time_solve(self._h, float(tol), int(maxit), int(reps), ctypes.byref(ms),
                                                     ctypes.byref(it)))
        return ms.value, it.value

    def time_assemble(self, h, reps=20):
        """Mean duration (ms) of one assembly of A and b (HIP events, no status read-back)."""
        ms = ctypes.c_double()
        _abi.check(_abi.load_library().yc_time_assemble(self._h, float(h), int(reps), ctypes.byref(ms)))
        return ms.value

    def cg_iteration_bytes(self):
        return _abi.load_library().yc_cg_iteration_bytes(self._h)

    def spmv_bytes(self):
        return _abi.load_library().yc_spmv_bytes(self._h)


class PhysicsCloth:
    """Reference-named view of the assembled system (src/PhysicsCloth.h:27-77).  The
    get_Spa* matrices are scipy CSR (dense n x n tensors in the reference), vectors are
    n-element numpy arrays, both in the reference's global layout."""

    def __init__(self, cloth):
        self.cloth = cloth
        self._h = None

    def FillMV(self, h):
        """FillMV; the step h is the one A and b are formed with (SimStep's)."""
        self.cloth.assemble(h)
        self._h = h

    def ResetMV(self):
        """Nothing to reset: the next assembly overwrites the system."""

    def get_SpaA(self):
        return self.cloth.csr("A")

    def get_SpaM(self):
        return self.cloth.csr("M")

    def get_SpaFPos(self):
        return self.cloth.csr("Kx")

    def get_SpaFVel(self):
        return self.cloth.csr("Kv")

    def get_SpaF(self):
        return self.cloth.vec("F")

    def get_SpaPos(self):
        return self.cloth.vec("pos")

    def get_SpaVel(self):
        return self.cloth.vec("v")

    @property
    def F_E_is_Slides(self):
        """[F_E_is_Slides[0]]: the is_slide of every crossing, (Nc, 2)."""
        return [self.cloth.crossing()[0]]


class EigenSolver:
    """EigenSolver::apply (src/Solvers.h:35-52) forward on the GPU: the caller's system
    becomes the cloth's and is solved with the reference's Jacobi-PCG to tol."""

    def __init__(self, cloth, tol=1e-8, maxit=None):
        self.cloth, self.tol, self.maxit = cloth, tol, maxit

    def apply(self, A, b):
        self.cloth.load_system(A, b)
        self.cloth.solve(tol=self.tol, maxit=self.maxit, raise_on_nonconv=False)
        return self.cloth.vec("x")


def SimStep(cloth, sim_phy, h, ext_force=None, tol=1e-8, maxit=None):
    """src/TrainHelper.cpp:15-97: False on NaN/Inf in A, b or the new velocity."""
    if ext_force is not None:
        cloth.set_ext_force(ext_force)
    _, _, rc = cloth.step(h, tol=tol, maxit=maxit)
    if ext_force is not None:
        cloth.set_ext_force(None)
    return rc != _abi.YC_ERR_NONFINITE
