"""Learned CG warm start: dataset generation from the simulator's own solves, GPU
training with the library's BF16-MFMA kernels, and the network wrapper.

The network (32-128-128-128-5, SiLU) corrects the quadratic extrapolation
x_q = 3 v_n - 3 v_{n-1} + v_{n-2} of the implicit step's velocity: per node it maps the
Jacobi correction of x_q's residual (at the node and its four neighbours), the last two
steps' velocity changes, the sliding offset and three flags to x - x_q.  CG starts from
x_q + out_std * y.  Zero-guess CG (Cloth.step) and the extrapolations alone
(Cloth.run_guess) are the yardsticks: every warm-started solve stops at the same
relative residual.
"""
import ctypes
import os

import numpy as np

from . import _abi
from . import cloth_spec as cs

FEATURE_NAMES = (["jac_%d" % i for i in range(5)] + ["dv_prev_%d" % i for i in range(5)] +
                 ["dv_prev2_%d" % i for i in range(5)] + ["slide_u", "slide_v"] +
                 ["jac_%s_%s" % (n, c) for n in ("left", "right", "down", "up") for c in "xyz"] +
                 ["interior", "handle", "flip"])
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "warmstart_weights.npz")

_fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))  # noqa: E731
_ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))  # noqa: E731


def init_params(seed):
    """Seeded initial parameters (uniform +-1/sqrt(fan_in), zero biases)."""
    rng = np.random.default_rng(seed)
    parts = []
    for fin, fout in ((32, 128), (128, 128), (128, 128), (128, 5)):
        lim = 1.0 / np.sqrt(fin)
        parts.append(rng.uniform(-lim, lim, (fout, fin)).astype(np.float32).ravel())
        parts.append(np.zeros(fout, np.float32))
    return np.concatenate(parts)


class WarmStartMLP:
    def __init__(self, params=None, seed=0):
        L = _abi.load_library()
        h = ctypes.c_void_p()
        _abi.check(L.yc_mlp_create(ctypes.byref(h)))
        self._h = h
        self.nparam = L.yc_mlp_param_count()
        self.in_mean = np.zeros(32, np.float32)
        self.in_std = np.ones(32, np.float32)
        self.out_std = np.ones(5, np.float32)
        self.scale = 1.0
        self.set_params(init_params(seed) if params is None else params)

    def close(self):
        if getattr(self, "_h", None):
            _abi.load_library().yc_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.nparam:
            raise ValueError("expected %d parameters, got %d" % (self.nparam, p.size))
        _abi.check(_abi.load_library().yc_mlp_set_params(self._h, _fp(p)))

    def params(self):
        p = np.zeros(self.nparam, np.float32)
        _abi.check(_abi.load_library().yc_mlp_get_params(self._h, _fp(p)))
        return p

    def set_norm(self, in_mean, in_std, out_std):
        """Input standardisation and the output scale; the device applies out_std times
        the correction scale (set_scale)."""
        self.in_mean = np.ascontiguousarray(in_mean, dtype=np.float32)
        self.in_std = np.ascontiguousarray(in_std, dtype=np.float32)
        self.out_std = np.ascontiguousarray(out_std, dtype=np.float32)
        self._upload_norm()

    def set_scale(self, s):
        """Scales the correction the network adds to the extrapolation; kept with the
        weights."""
        self.scale = float(s)
        self._upload_norm()

    def _upload_norm(self):
        dev_out = np.ascontiguousarray(self.out_std * np.float32(getattr(self, "scale", 1.0)), dtype=np.float32)
        _abi.check(_abi.load_library().yc_mlp_set_norm(self._h, _fp(self.in_mean), _fp(self.in_std), _fp(dev_out)))

    def set_data(self, X, T, M):
        self._X = np.ascontiguousarray(X, dtype=np.float32)
        self._T = np.ascontiguousarray(T, dtype=np.float32)
        self._M = np.ascontiguousarray(M, dtype=np.float32)
        _abi.check(_abi.load_library().yc_mlp_set_data(self._h, _fp(self._X), _fp(self._T), _fp(self._M), len(self._X)))

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((len(X), 5), np.float32)
        _abi.check(_abi.load_library().yc_mlp_forward(self._h, _fp(X), len(X), _fp(Y)))
        return Y

    def grad(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        g = np.zeros(self.nparam, np.float32)
        loss = ctypes.c_float()
        _abi.check(_abi.load_library().yc_mlp_grad(self._h, _ip(idx), len(idx), _fp(g), ctypes.byref(loss)))
        return g, loss.value

    def train_epoch(self, idx, batch, lr):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        loss = ctypes.c_float()
        _abi.check(_abi.load_library().yc_mlp_train(self._h, _ip(idx), len(idx), int(batch), float(lr), ctypes.byref(loss)))
        return loss.value

    def time_forward(self, rows, reps=100):
        ms = ctypes.c_double()
        _abi.check(_abi.load_library().yc_mlp_time_forward(self._h, int(rows), int(reps), ctypes.byref(ms)))
        return ms.value

    def save(self, path):
        np.savez(path, params=self.params(), in_mean=self.in_mean, in_std=self.in_std, out_std=self.out_std,
                 scale=np.float32(getattr(self, "scale", 1.0)))

    @classmethod
    def load(cls, path):
        d = np.load(path)
        m = cls(params=d["params"])
        m.scale = float(d["scale"]) if "scale" in d.files else 1.0
        m.set_norm(d["in_mean"], d["in_std"], d["out_std"])
        return m


def generate(cloth_specs, steps, skip=5, tol=1e-10, node_stride=1):
    """Run the simulator and collect (features, target, mask) for every node of every
    step after the first `skip` steps: the features before each solve and the converged
    solution's difference from the quadratic extrapolation after it.  node_stride
    subsamples nodes."""
    from .cloth import Cloth
    Xs, Ts, Ms = [], [], []
    for spec in cloth_specs:
        c = Cloth(spec)
        c.set_ClothState(cs.initial_state(spec))
        mask = c.node_mask()
        h = spec["h"]
        for s in range(steps):
            c.assemble(h)
            X = c.features()
            it, rr, ok = c.solve(tol=tol, maxit=20000)
            if not ok:
                raise RuntimeError("data generation: CG did not converge")
            T = c.targets()
            c.update(h)
            if s >= skip:
                Xs.append(X[::node_stride]); Ts.append(T[::node_stride]); Ms.append(mask[::node_stride])
        c.close()
    return np.concatenate(Xs), np.concatenate(Ts), np.concatenate(Ms)


def training_cloths(n=8, seed=1000, lo=48, hi=128):
    """Seeded cloths of lo x lo to hi x hi crossings, all three weaves, seeded wind (the
    generator of the benchmark's cloth, other seeds)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        W = int(rng.integers(lo, hi + 1))
        spec = cs.heter_spec(W, ("plain", "twill", "satin")[i % 3], seed + i)
        spec["wind_vel"] = [0.0, float(rng.uniform(2.0, 8.0)), 0.0]
        out.append(spec)
    return out


def validation_cloth():
    """A held-out cloth of the training generator (another seed, weave and wind) on which
    the correction scale is chosen."""
    spec = cs.heter_spec(128, "plain", 2024)
    spec["wind_vel"] = [0.0, 5.0, 0.0]
    return spec


def choose_scale(net, spec, scales=(0.0, 0.25, 0.5, 0.75, 1.0), skip=5, steps=40, tol=1e-8, log=None):
    """The correction scale with the fewest CG iterations on `spec` (persistent CG,
    block-Jacobi preconditioner, the benchmark's tolerance), after `skip` steps."""
    from .cloth import Cloth
    best = None
    for s in scales:
        net.set_scale(s)
        c = Cloth(spec)
        c.set_cg_variant("persistent")
        c.set_preconditioner("block2")
        c.set_ClothState(cs.initial_state(spec))
        c.run_warm(net, spec["h"], skip, tol=tol)
        it, _ = c.run_warm(net, spec["h"], steps, tol=tol)
        c.close()
        if log:
            log("scale %.2f: %.2f CG iterations per step" % (s, it / steps))
        if best is None or it < best[1]:
            best = (s, it)
    net.set_scale(best[0])
    return best[0]


def fit(X, T, M, epochs=12, batch=4096, lr=2e-3, seed=0, log=None):
    """Standardise, train on the GPU, return the network (with its normalisation)."""
    in_mean = X.mean(0)
    in_std = X.std(0)
    in_std[in_std == 0] = 1.0
    out_std = np.sqrt((T ** 2 * M).sum(0) / np.maximum(M.sum(0), 1.0)).astype(np.float32)
    out_std[out_std == 0] = 1.0
    net = WarmStartMLP(seed=seed)
    net.set_norm(in_mean, in_std, out_std)
    net.set_data(X, T / out_std, M)
    rng = np.random.default_rng(seed)
    losses = []
    for ep in range(epochs):
        lr_e = lr * 0.5 * (1.0 + np.cos(np.pi * ep / epochs))
        loss = net.train_epoch(rng.permutation(len(X)).astype(np.int32), batch, lr_e)
        losses.append(loss)
        if log:
            log("epoch %d loss %.4e" % (ep, loss))
    return net, losses
