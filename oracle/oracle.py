"""ctypes wrapper of the CPU oracle (test infrastructure only).

The oracle restates the reference's FP64 implicit step on the CPU (see
oracle/yarn_oracle.cpp).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load it, and only as the checker.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
MAX_HANDLES = 16


class YoParams(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int), ("length", ctypes.c_int),
        ("hight", ctypes.c_double), ("L", ctypes.c_double), ("R", ctypes.c_double),
        ("n_yarns", ctypes.c_int),
        ("rho", ctypes.c_double * 2), ("Y", ctypes.c_double * 2), ("B", ctypes.c_double * 2),
        ("kc", ctypes.c_double), ("kf", ctypes.c_double), ("df", ctypes.c_double),
        ("mu", ctypes.c_double), ("S", ctypes.c_double), ("handle_k", ctypes.c_double),
        ("pattern", ctypes.c_int), ("pose", ctypes.c_int),
        ("G", ctypes.c_double * 3),
        ("wind_vel", ctypes.c_double * 3), ("wind_density", ctypes.c_double), ("wind_drag", ctypes.c_double),
        ("n_handles", ctypes.c_int),
        ("handles", ctypes.c_int * (2 * MAX_HANDLES)),
    ]


def build():
    """Compile liboracle.so with the recipe in oracle/Makefile."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.yo_create.restype = ctypes.c_void_p
        L.yo_create.argtypes = [ctypes.POINTER(YoParams)]
        L.yo_destroy.argtypes = [ctypes.c_void_p]
        L.yo_ndof.argtypes = [ctypes.c_void_p]
        L.yo_get_state.argtypes = [ctypes.c_void_p] + [dp] * 6
        L.yo_set_state.argtypes = [ctypes.c_void_p] + [dp] * 6
        L.yo_assemble.argtypes = [ctypes.c_void_p, ctypes.c_double, dp]
        L.yo_nnz.argtypes = [ctypes.c_void_p]
        L.yo_get_csr.argtypes = [ctypes.c_void_p, ctypes.c_int, ip, ip, dp]
        L.yo_get_dense.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
        L.yo_get_vec.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
        L.yo_get_crossing.argtypes = [ctypes.c_void_p, dp, dp, dp]
        L.yo_get_normals.argtypes = [ctypes.c_void_p, dp]
        L.yo_solve.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, dp, dp, ip, dp]
        L.yo_update.argtypes = [ctypes.c_void_p, dp, ctypes.c_double]
        L.yo_step.argtypes = [ctypes.c_void_p, ctypes.c_double, dp, ctypes.c_double, ctypes.c_int, ip, dp]
        L.yo_owner.argtypes = [ctypes.c_int, ctypes.c_int]
        L.yo_freeze_shear_angle.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.yo_set_frozen_shear_angle.argtypes = [ctypes.c_void_p, dp, ctypes.c_int]
        L.yo_get_frozen_shear_angle.argtypes = [ctypes.c_void_p, dp]
        L.yo_set_terms.argtypes = [ctypes.c_void_p, ctypes.c_uint]
        L.yo_set_element.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.yo_friction_eval.argtypes = [dp, dp]
        L.yo_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.yo_step_dense.argtypes = [ctypes.c_void_p, ctypes.c_double, dp, ctypes.c_double, ctypes.c_int, ip, dp]
        L.yo_get_compression_parts.argtypes = [ctypes.c_void_p, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    if a is None:
        return ctypes.POINTER(ctypes.c_double)()
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_params(spec):
    """spec: dict with the cloth description (see the package's cloth_spec)."""
    p = YoParams()
    p.width, p.length = int(spec["width"]), int(spec["length"])
    p.hight, p.L, p.R = float(spec["hight"]), float(spec["L"]), float(spec["R"])
    yarns = spec["yarns"]
    p.n_yarns = len(yarns)
    for i, y in enumerate(yarns):
        p.rho[i], p.Y[i], p.B[i] = float(y["rho"]), float(y["Y"]), float(y["B"])
    for k in ("kc", "kf", "df", "mu", "S", "handle_k"):
        setattr(p, k, float(spec[k]))
    p.pattern = {"plain": 0, "twill": 1, "satin": 2}[spec["pattern"]]
    p.pose = {"upright": 0, "flat": 1}[spec["pose"]]
    for i in range(3):
        p.G[i] = float(spec["G"][i])
        p.wind_vel[i] = float(spec["wind_vel"][i])
    p.wind_density, p.wind_drag = float(spec["wind_density"]), float(spec["wind_drag"])
    hs = spec["handles"]
    p.n_handles = len(hs)
    for i, (u, v) in enumerate(hs):
        p.handles[2 * i], p.handles[2 * i + 1] = int(u), int(v)
    return p


class OracleCloth:
    """One cloth in the oracle: Cloth + PhysicsCloth + SimStep on the CPU."""

    def __init__(self, spec):
        self.spec = spec
        self._p = make_params(spec)
        self._h = lib().yo_create(ctypes.byref(self._p))
        if not self._h:
            raise ValueError("oracle rejected the cloth description")
        self.W, self.Lg = spec["width"], spec["length"]
        self.N = self.W * self.Lg
        self.Nc = (self.W - 2) * (self.Lg - 2)
        self.n = lib().yo_ndof(self._h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().yo_destroy(h)
            self._h = None

    # state: dict of LPos (N,3), LPosFix, LVel, EPos (N,2), EPosBar, EVel
    def get_state(self):
        N = self.N
        s = {k: np.zeros((N, 3)) for k in ("LPos", "LPosFix", "LVel")}
        s.update({k: np.zeros((N, 2)) for k in ("EPos", "EPosBar", "EVel")})
        lib().yo_get_state(self._h, _dp(s["LPos"]), _dp(s["LPosFix"]), _dp(s["LVel"]),
                           _dp(s["EPos"]), _dp(s["EPosBar"]), _dp(s["EVel"]))
        return s

    def set_state(self, s):
        a = {k: np.ascontiguousarray(s[k], dtype=np.float64) for k in
             ("LPos", "LPosFix", "LVel", "EPos", "EPosBar", "EVel")}
        lib().yo_set_state(self._h, _dp(a["LPos"]), _dp(a["LPosFix"]), _dp(a["LVel"]),
                           _dp(a["EPos"]), _dp(a["EPosBar"]), _dp(a["EVel"]))

    def assemble(self, h, ext_force=None):
        ef = None if ext_force is None else np.ascontiguousarray(ext_force, dtype=np.float64)
        return lib().yo_assemble(self._h, float(h), _dp(ef))

    def csr(self, which=0):
        """which: 0 A, 1 M, 2 K_x, 3 K_v -> scipy.sparse.csr_matrix"""
        import scipy.sparse as sp
        nnz = lib().yo_nnz(self._h)
        rp = np.zeros(self.n + 1, dtype=np.int32)
        cl = np.zeros(nnz, dtype=np.int32)
        val = np.zeros(nnz)
        lib().yo_get_csr(self._h, which, rp.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                         cl.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dp(val))
        return sp.csr_matrix((val, cl, rp), shape=(self.n, self.n))

    def dense(self, which=0):
        out = np.zeros((self.n, self.n))
        lib().yo_get_dense(self._h, which, _dp(out))
        return out

    def vec(self, which):
        """which: 'F', 'b' or 'v'"""
        out = np.zeros(self.n)
        lib().yo_get_vec(self._h, {"F": 0, "b": 1, "v": 2}[which], _dp(out))
        return out

    def crossing(self):
        sl = np.zeros(2 * self.Nc)
        cp = np.zeros(self.Nc)
        fp = np.zeros(2 * self.Nc)
        lib().yo_get_crossing(self._h, _dp(sl), _dp(cp), _dp(fp))
        return sl.reshape(-1, 2), cp, fp.reshape(-1, 2)

    def normals(self):
        out = np.zeros((self.N, 3))
        lib().yo_get_normals(self._h, _dp(out))
        return out

    def solve(self, tol=1e-12, maxit=None, x0=None):
        if maxit is None:
            maxit = 2 * self.n
        x = np.zeros(self.n)
        it = ctypes.c_int(0)
        rr = ctypes.c_double(0)
        x0a = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
        st = lib().yo_solve(self._h, float(tol), int(maxit), _dp(x0a), _dp(x), ctypes.byref(it), ctypes.byref(rr))
        return x, it.value, rr.value, st

    TERMS = {"mass": 1, "gravity": 2, "stretch": 4, "inertia": 8, "mdot": 16, "paracolli": 32, "bend": 64,
             "crimp": 128, "friction": 256, "shear": 512, "wind": 1024, "handle": 2048}

    def set_terms(self, names=None):
        """Evaluate only the named element terms in the following assemblies (None: all)."""
        mask = 4095 if names is None else sum(self.TERMS[n] for n in names)
        lib().yo_set_terms(self._h, mask)

    ELEMENTS = {"all": 0, "edge": 1, "bend": 2, "shear": 3, "face": 4, "crossing": 5}

    def set_element(self, kind="all", index=-1):
        """Assemble only element `index` of the named list (edge, bend, shear, face) or the
        friction of crossing `index`; "all" restores the full assembly."""
        lib().yo_set_element(self._h, self.ELEMENTS[kind], int(index))

    def set_threads(self, n):
        """OpenMP threads of the assembly (bit for bit the single-thread result), the
        normals and the CG (chunked inner products when n > 1)."""
        lib().yo_set_threads(self._h, int(n))

    def compression_parts(self):
        """StrhBend_C_L per node (N, 3) and the self StrhBend_C_LE entries (N, 2, 3)."""
        a, b = np.zeros((self.N, 3)), np.zeros((self.N, 2, 3))
        lib().yo_get_compression_parts(self._h, _dp(a), _dp(b))
        return a, b

    def freeze_shear_angle(self, on=True):
        """Hold the non-jamming shear's (phi - pi/2) at the current state's values."""
        lib().yo_freeze_shear_angle(self._h, 1 if on else 0)

    def frozen_shear_angle(self):
        out = np.zeros(8 * self.n + 16)
        k = lib().yo_get_frozen_shear_angle(self._h, _dp(out))
        return out[:k].copy()

    def set_frozen_shear_angle(self, pb):
        pb = np.ascontiguousarray(pb, dtype=np.float64)
        lib().yo_set_frozen_shear_angle(self._h, _dp(pb), int(pb.size))

    def update(self, new_vel, h):
        nv = np.ascontiguousarray(new_vel, dtype=np.float64)
        lib().yo_update(self._h, _dp(nv), float(h))

    def step_dense(self, h, ext_force=None, tol=1e-12, maxit=None):
        """step() with the reference's dense storage work (cost model; same result)."""
        if maxit is None:
            maxit = 2 * self.n
        ef = None if ext_force is None else np.ascontiguousarray(ext_force, dtype=np.float64)
        it = ctypes.c_int(0)
        rr = ctypes.c_double(0)
        st = lib().yo_step_dense(self._h, float(h), _dp(ef), float(tol), int(maxit), ctypes.byref(it), ctypes.byref(rr))
        return st, it.value, rr.value

    def step(self, h, ext_force=None, tol=1e-12, maxit=None):
        if maxit is None:
            maxit = 2 * self.n
        ef = None if ext_force is None else np.ascontiguousarray(ext_force, dtype=np.float64)
        it = ctypes.c_int(0)
        rr = ctypes.c_double(0)
        st = lib().yo_step(self._h, float(h), _dp(ef), float(tol), int(maxit), ctypes.byref(it), ctypes.byref(rr))
        return st, it.value, rr.value


def friction_eval(dlt, comp, compEE, FNS, Fpt, Fpos, Fvel, uvd, mu, kf, df):
    """The oracle's ConSlideFriction arithmetic for one component -> (is_slide, force,
    EE position entry, EE velocity entry)."""
    i = np.array([dlt, comp, compEE, FNS, Fpt, Fpos, Fvel, uvd, mu, kf, df], dtype=np.float64)
    o = np.zeros(4)
    lib().yo_friction_eval(_dp(i), _dp(o))
    return tuple(o)


def owner(n_iter, i):
    return lib().yo_owner(int(n_iter), int(i))
