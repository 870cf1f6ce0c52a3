"""Many cloths of uneven size stepped together, and the launch-shape autotuner.

``ClothBatch`` wraps yc_batch_*: each phase of a step is one launch over every cloth
(the persistent CG one launch whose blocks walk groups of cloths that fit on the device at
once), replayed from a hipGraph.  Each cloth's result is bit-identical to stepping it
alone with ``Cloth.run`` and the persistent CG.  The reference has no batching: its
training drivers loop over cloths one SimStep at a time (src/TrainHeter.h:81-146).

``autotune_batch`` and ``autotune_cloth`` time the launch options on the caller's own
workload (a few steps from the current state, which is restored afterwards) and keep
the fastest; results can be cached in a JSON file keyed by the workload's shape.
"""
import ctypes
import itertools
import json
import os
import time

import numpy as np

from . import _abi
from .cloth import Cloth


class ClothBatch:
    """Cloths (``Cloth`` objects on one device) stepped by shared launches.  The cloths
    stay owned by the caller and must outlive the batch."""

    NODE_BLOCKS = (64, 128, 256)

    def __init__(self, cloths):
        self.cloths = list(cloths)
        if not self.cloths:
            raise ValueError("empty batch")
        L = _abi.load_library()
        arr = (ctypes.c_void_p * len(self.cloths))(*[c._h.value for c in self.cloths])
        h = ctypes.c_void_p()
        _abi.check(L.yc_batch_create(arr, len(self.cloths), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _abi.load_library().yc_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.cloths)

    def set_opts(self, graph=None, group_order=None, node_block=None):
        o = self.opts()
        if graph is not None:
            o["graph"] = int(graph)
        if group_order is not None:
            o["group_order"] = int(group_order)
        if node_block is not None:
            o["node_block"] = int(node_block)
        c = _abi.YcBatchOpts(o["graph"], o["group_order"], o["node_block"])
        _abi.check(_abi.load_library().yc_batch_set_opts(self._h, ctypes.byref(c)))

    def opts(self):
        c = _abi.YcBatchOpts()
        _abi.check(_abi.load_library().yc_batch_get_opts(self._h, ctypes.byref(c)))
        return {"graph": c.graph, "group_order": c.group_order, "node_block": c.node_block}

    def info(self):
        g, u, t, cap = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _abi.check(_abi.load_library().yc_batch_info(self._h, ctypes.byref(g), ctypes.byref(u), ctypes.byref(t),
                                                     ctypes.byref(cap)))
        return {"groups": g.value, "graph_used": bool(u.value), "tiles": t.value, "capacity": cap.value}

    def run(self, h, steps, tol=1e-8, maxit=None):
        """`steps` SimSteps of every cloth.  Returns (iters, status, rc): CG iterations
        per cloth summed over the steps, each cloth's worst status, and the worst
        overall.  Raises on HIP errors and NaN/Inf (as ``Cloth.run``); a solve stopped
        at maxit is reported in status / rc, not raised."""
        n = len(self.cloths)
        it = (ctypes.c_longlong * n)()
        st = (ctypes.c_int * n)()
        rc = _abi.load_library().yc_batch_run(self._h, float(h), int(steps), float(tol),
                                              -1 if maxit is None else int(maxit), it, st)
        if rc not in (_abi.YC_OK, _abi.YC_ERR_NOT_CONVERGED):
            _abi.check(rc)
        return np.array(it[:], dtype=np.int64), np.array(st[:], dtype=np.int32), rc


    def adjoint_run(self, h, steps, tol=1e-8, maxit=None):
        """`steps` recorded SimSteps of every cloth (each cloth's adjoint_begin first,
        checkpoint interval 1).  Returns (iters, status); raises on failure, which voids
        every member's recording."""
        n = len(self.cloths)
        it = (ctypes.c_longlong * n)()
        st = (ctypes.c_int * n)()
        rc = _abi.load_library().yc_batch_adjoint_run(self._h, float(h), int(steps), float(tol),
                                                      -1 if maxit is None else int(maxit), it, st)
        for c in self.cloths:
            c._adj_steps = 0 if rc != _abi.YC_OK else c._adj_steps + int(steps)
        _abi.check(rc)
        return np.array(it[:], dtype=np.int64), np.array(st[:], dtype=np.int32)

    def adjoint_backward(self, h, mode="ref", tol=1e-8, maxit=None):
        """The backward passes of every cloth's recording against its targets
        (Cloth.adjoint_set_targets).  Returns (dparam (n, 12) in Cloth.PARAMS order,
        loss (n,), backward CG iterations (n,))."""
        from .cloth import Cloth
        n = len(self.cloths)
        dparam = np.zeros((n, len(Cloth.PARAMS)))
        loss = np.zeros(n)
        it = (ctypes.c_longlong * n)()
        st = (ctypes.c_int * n)()
        dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
        rc = _abi.load_library().yc_batch_adjoint_backward(self._h, Cloth.GRAD_MODES[mode], float(h), float(tol),
                                                           -1 if maxit is None else int(maxit), dp(dparam), dp(loss), it, st)
        _abi.check(rc)
        return dparam, loss, np.array(it[:], dtype=np.int64)


def batch_signature(cloths):
    return "batch:" + ",".join("%dx%d" % (c.width, c.length) for c in cloths)


def _load_cache(path):
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def _store_cache(path, key, value):
    if not path:
        return
    data = _load_cache(path)
    data[key] = value
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def autotune_batch(batch, h, steps=3, tol=1e-8, maxit=None, cache=None, candidates=None):
    """Times every launch option of `batch` on its own cloths and keeps the fastest.
    Options: graph replay on/off, layout order (given / by decreasing size: decides
    which cloths share a resident group of the solve launch) and the per-node block size.  Each
    candidate runs one untimed step (graph capture) and `steps` timed steps from the
    same saved state; the state is restored at the end.  Returns (best options,
    {option tuple: seconds per step})."""
    key = batch_signature(batch.cloths)
    cached = _load_cache(cache).get(key)
    if cached:
        batch.set_opts(**cached["best"])
        return cached["best"], {}
    saved = [c.get_ClothState() for c in batch.cloths]

    def restore():
        for c, s in zip(batch.cloths, saved):
            c.set_ClothState(s)

    if candidates is None:
        candidates = list(itertools.product((1, 0), (0, 1), ClothBatch.NODE_BLOCKS))
    times = {}
    for graph, order, nb in candidates:
        batch.set_opts(graph=graph, group_order=order, node_block=nb)
        restore()
        batch.run(h, 1, tol, maxit)
        restore()
        t0 = time.perf_counter()
        batch.run(h, steps, tol, maxit)
        times[(graph, order, nb)] = (time.perf_counter() - t0) / steps
    restore()
    g, o, nb = min(times, key=times.get)
    best = {"graph": g, "group_order": o, "node_block": nb}
    batch.set_opts(**best)
    _store_cache(cache, key, {"best": best, "s_per_step": {"%d,%d,%d" % k: v for k, v in times.items()}})
    return best, times


def autotune_cloth(cloth, h, steps=3, tol=1e-8, maxit=None, cache=None):
    """Picks the fastest CG form (reference / fused / persistent) for one cloth on its
    own state, as ``autotune_batch`` does for a batch.  Returns (name, {name: s/step})."""
    key = "cloth:%dx%d" % (cloth.width, cloth.length)
    cached = _load_cache(cache).get(key)
    if cached:
        cloth.set_cg_variant(cached["best"])
        return cached["best"], {}
    saved = cloth.get_ClothState()
    times = {}
    for name in Cloth.CG_VARIANTS:
        try:
            cloth.set_cg_variant(name)
        except _abi.YcError:
            continue   # the form does not fit this cloth on this device
        cloth.set_ClothState(saved)
        cloth.run(h, 1, tol, maxit)
        cloth.set_ClothState(saved)
        t0 = time.perf_counter()
        cloth.run(h, steps, tol, maxit)
        times[name] = (time.perf_counter() - t0) / steps
    cloth.set_ClothState(saved)
    best = min(times, key=times.get)
    cloth.set_cg_variant(best)
    _store_cache(cache, key, {"best": best, "s_per_step": times})
    return best, times
