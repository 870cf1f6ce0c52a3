"""Benchmark: implicit yarn-level steps/s on one MI355X, learned warm start vs plain CG.

Workload: the 128x128 two-yarn twill cloth with seeded wind (cloth_spec.workload
"twill128"), h = 1 ms, FP64, the reference's step (assemble, solve, update) per step.
Two runs from the same initial state: plain zero-guess PCG, and PCG started from the
learned warm start; both stop at the same 1e-8 relative residual (persistent form,
2x2-block-Jacobi preconditioner by default).  `value` is the warm-started steps/s; the
plain rate, the ratio and the Jacobi-preconditioned rates (the reference solver's
preconditioner, persistent and reference forms) are in `config`.  Inputs
live in HBM before the timed region.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
MLP_FLOP_PER_NODE = 2 * (32 * 128 + 128 * 128 * 2 + 128 * 5)   # 75,008


def _cpu_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, os.cpu_count() or 1, avail


def cpu_baseline(cs, name, nsteps):
    """The CPU oracle (the reference's step restated in C++, FP64, the same Jacobi-PCG and
    stopping rule) on steps 2..nsteps+1 of the same workload, on this host's cores: with
    every core available to the process (capped by OMP_NUM_THREADS when it is set), with 8
    threads (the reference's NThrd, src/Constants.h:26) and with one.  The assembly with
    several threads is bit for bit the single-thread one.  Beside it, `ref_dense` times
    the same oracle with the reference's dense storage work on a 25 x 25 cloth, the
    largest size that layout handles (a lower bound of the reference's own cost)."""
    from oracle.oracle import OracleCloth
    model, nproc, avail = _cpu_info()
    env_t = os.environ.get("OMP_NUM_THREADS")
    threads_all = min(avail, int(env_t)) if env_t and env_t.isdigit() else avail
    spec, state, ext, _ = cs.workload(name)
    h = spec["h"]

    def rate(threads, cspec, cstate, cext, steps, dense=False):
        c = OracleCloth(cspec)
        c.set_threads(threads)
        c.set_state(cstate)
        stepf = c.step_dense if dense else c.step
        stepf(h, cext, tol=1e-8, maxit=20000)   # setup / first step untimed
        it_tot = 0
        t0 = time.perf_counter()
        for _ in range(steps):
            st, it, _ = stepf(h, cext, tol=1e-8, maxit=20000)
            it_tot += it
        dt = time.perf_counter() - t0
        return {"steps_per_s": steps / dt, "ms_per_step": 1e3 * dt / steps, "cg_iters_per_step": it_tot / steps,
                "threads": threads}
    main = rate(threads_all, spec, state, ext, nsteps)
    t8 = rate(8, spec, state, ext, nsteps)
    t1 = rate(1, spec, state, ext, max(2, nsteps // 2))
    dspec = cs.heter_spec(25, "twill", 5)
    dstate = cs.initial_state(dspec)
    dense = rate(8, dspec, dstate, None, 3, dense=True)
    sparse25 = rate(8, dspec, dstate, None, 3)
    return {"value": main["steps_per_s"], "unit": "steps/s", "cores": threads_all, "kind": "port",
            "sample": "steps 2..%d of the %s workload, oracle FP64 (block-sparse storage), Jacobi-PCG to 1e-8, "
                      "%d OpenMP threads" % (nsteps + 1, name, threads_all),
            "cpu_model": model, "nproc": nproc, "cores_available": avail,
            "ms_per_step": main["ms_per_step"], "cg_iters_per_step": main["cg_iters_per_step"],
            "threads_8": t8, "threads_1": t1,
            "ref_dense_w25": {"workload": "25x25 two-yarn twill, steps 2..4, 8 threads", "dense_layout": dense,
                              "block_sparse": sparse25,
                              "dense_over_sparse_time": dense["ms_per_step"] / sparse25["ms_per_step"]}}


def batch_measure(ys, steps, warmup=2, precond="block2"):
    """64 cloths of uneven size (sides 10..64, two-yarn blends) stepped by shared
    launches with the autotuned options, against the same cloths stepped one after
    another on their own (persistent CG, queued steps, the same preconditioner)."""
    cs = ys.cloth_spec
    items = cs.batch_workload()
    h, tol = items[0][0]["h"], 1e-8
    cloths = []
    for spec, state in items:
        c = ys.Cloth(spec)
        c.set_cg_variant("persistent")
        c.set_preconditioner(precond)
        c.set_ClothState(state)
        cloths.append(c)
    b = ys.ClothBatch(cloths)
    best, _ = ys.autotune_batch(b, h, steps=2, tol=tol)
    b.run(h, warmup, tol=tol)
    t0 = time.perf_counter()
    iters, status, rc = b.run(h, steps, tol=tol)
    dt = time.perf_counter() - t0
    info = b.info()
    b.close()
    for c, (spec, state) in zip(cloths, items):
        c.set_ClothState(state)
        c.run(h, warmup, tol=tol)
    t0 = time.perf_counter()
    for c in cloths:
        c.run(h, steps, tol=tol)
    dt_seq = time.perf_counter() - t0
    for c in cloths:
        c.close()
    nodes = sum(s["width"] * s["length"] for s, _ in items)
    return {"workload": "64 cloths, sides 10..64 (seed 3), two-yarn blends in plain/twill/satin, h=0.001, "
                        "%s-PCG to 1e-08, zero guess" % ("2x2-block-Jacobi" if precond == "block2" else "Jacobi"),
            "steps": steps,
            "cloth_steps_per_s": len(items) * steps / dt, "node_steps_per_s": nodes * steps / dt,
            "ms_per_step": 1e3 * dt / steps, "sequential_ms_per_step": 1e3 * dt_seq / steps,
            "speedup_over_sequential": dt_seq / dt, "cg_iters_per_cloth_step": float(iters.sum()) / (len(items) * steps),
            "status": int(rc), "options": best, "solve_groups": info["groups"], "graph_replay": info["graph_used"],
            "tiles": info["tiles"], "resident_tiles": info["capacity"]}


def large_measure(ys, steps):
    """The 256 x 256 homogeneous plain cloth with a seeded +z pressure (about 327k DOF):
    plain-CG steps/s (the fused form: the persistent one needs every tile resident),
    and the assembly and the stencil product against the HBM peak."""
    cs = ys.cloth_spec
    spec, state, ext, _ = cs.workload("homo256")
    h, tol = spec["h"], 1e-8
    c = ys.Cloth(spec)
    c.set_cg_variant("fused")
    c.set_ClothState(state)
    c.set_ext_force(ext)
    c.run(h, 2, tol=tol)
    t0 = time.perf_counter()
    iters, rc = c.run(h, steps, tol=tol)
    dt = time.perf_counter() - t0
    c.assemble(h)
    asm_ms = c.time_assemble(h, 20)
    spmv_ms = c.time_spmv(200)
    spmv_bytes = c.spmv_bytes()
    W, n = spec["width"], c.n
    nnz = (spmv_bytes - 16 * n) / 8
    asm_bytes = 8 * nnz + 16 * n + 8 * 16 * W * W
    c.close()
    return {"workload": "homo256: 256x256 crossings, plain weave, cotton, seeded +z pressure, h=0.001, "
                        "Jacobi-PCG (fused form) to 1e-08", "steps": steps, "steps_per_s": steps / dt,
            "ms_per_step": 1e3 * dt / steps, "cg_iters_per_step": iters / steps, "status": int(rc),
            "spmv": {"bound": "hbm", "achieved": spmv_bytes / (spmv_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": spmv_bytes / (spmv_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "bytes_per_launch": spmv_bytes, "avg_launch_ms": spmv_ms},
            "assembly": {"bound": "hbm", "achieved": asm_bytes / (asm_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": asm_bytes / (asm_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "bytes_per_assembly": asm_bytes, "avg_ms": asm_ms}}


def tangent_measure(ys, spec, state, steps, precond):
    """The differentiation path on the same workload: SimSteps with the forward-mode
    tangents of the eight learned material parameters (the reference's gradient, and the
    exact one), as a parameter-learning rollout runs them."""
    names = ("rho0", "rho1", "Y0", "Y1", "B0", "B1", "S", "mu")
    out = {"workload": "the same cloth and state, %d steps, tangents of %s" % (steps, ", ".join(names)),
           "directions": len(names)}
    for mode in ("ref", "exact"):
        c = ys.Cloth(spec)
        c.set_cg_variant("persistent")
        c.set_preconditioner(precond)
        c.set_ClothState(state)
        dp = np.zeros((len(names), len(c.PARAMS)))
        for i, n in enumerate(names):
            dp[i, c.PARAMS.index(n)] = 1.0
        c.tangent_begin(dp)
        c.step_tangent(spec["h"], mode, tol=1e-8)   # untimed
        t0 = time.perf_counter()
        for _ in range(steps):
            c.step_tangent(spec["h"], mode, tol=1e-8)
        dt = time.perf_counter() - t0
        c.tangent_end()
        c.close()
        out[mode] = {"steps_per_s": steps / dt, "ms_per_step": 1e3 * dt / steps}
    return out


def adjoint_measure(ys, spec, state, steps, precond):
    """The reverse mode on the same workload: a rollout of `steps` recorded SimSteps and
    its backward pass (one backward solve per step) giving the position loss's gradient
    with respect to all twelve parameters, against targets from the rollout of the cloth
    with its first yarn 10 % heavier (made beforehand, untimed)."""
    tspec = dict(spec)
    tspec["yarns"] = [dict(spec["yarns"][0], rho=spec["yarns"][0]["rho"] * 1.1)] + list(spec["yarns"][1:])
    t = ys.Cloth(tspec)
    t.set_cg_variant("persistent")
    t.set_preconditioner(precond)
    t.set_ClothState(state)
    X, U = [], []
    for _ in range(steps):
        t.step(spec["h"], tol=1e-8)
        st = t.get_ClothState()
        X.append(st["LPos"])
        U.append(st["EPos"])
    t.close()
    X, U = np.array(X), np.array(U)
    out = {"workload": "the same cloth and state, %d recorded steps + the backward pass, gradient of the position "
                       "loss with respect to the 12 parameters" % steps}
    for mode in ("ref", "exact"):
        c = ys.Cloth(spec)
        c.set_cg_variant("persistent")
        c.set_preconditioner(precond)
        runs = []
        for rep in range(2):   # the first run is untimed (allocation, kernel loading)
            c.set_ClothState(state)
            t0 = time.perf_counter()
            c.adjoint_begin(steps)
            for _ in range(steps):
                c.adjoint_step(spec["h"], tol=1e-8)
            c.adjoint_set_targets(X, U, 1.0 / c.N)
            t1 = time.perf_counter()
            r = c.adjoint_backward(spec["h"], mode, tol=1e-8)
            t2 = time.perf_counter()
            c.adjoint_end()
            runs.append((t1 - t0, t2 - t1, r))
        c.close()
        fwd, bwd, r = runs[-1]
        out[mode] = {"ms_per_gradient_step": 1e3 * (fwd + bwd) / steps, "forward_ms_per_step": 1e3 * fwd / steps,
                     "backward_ms_per_step": 1e3 * bwd / steps, "backward_cg_iters_per_step": r["cg_iters"] / steps,
                     "gradient_steps_per_s": steps / (fwd + bwd)}
    return out


def gradient_measure(ys, steps, precond):
    """Batched differentiable learning (64 cloths of uneven size, the batch workload):
    one loss-and-gradient evaluation of the eight material parameters per cloth, each a
    rollout of `steps` steps from the first frame of its ground truth with perturbed
    parameters, as one epoch of parameter learning runs it.  The ground truth comes from
    a batch run beforehand (untimed).  The gradient is the reference's (ref mode), from
    forward-mode tangents: eight directions and the primal solve per step."""
    import yarnstep.learn as ln
    items = ys.cloth_spec.batch_workload()
    frames = ln.batch_trajectories(items, steps)
    out = {"workload": "64 cloths, sides 10..64 (seed 3), %d-step rollouts, gradients of %s" % (steps, ", ".join(ln.NAMES))}
    grads = {}
    for form, fn in (("tangents", ln.sample_loss_terms), ("adjoint", ln.sample_loss_terms_adjoint)):
        t0 = time.perf_counter()
        ok = 0
        gs = []
        for i, ((spec, _), f) in enumerate(zip(items, frames)):
            sp = ln.perturbed_spec(spec, ln.NAMES, 500 + i)
            _, g, ok_ = fn(sp, f, ln.NAMES, mode="ref", tol=1e-8, precond=precond)
            ok += ok_
            gs.append(g)
        dt = time.perf_counter() - t0
        grads[form] = np.array(gs)
        out[form] = {"form": ("forward-mode tangents, 8 directions + the primal solve per step" if form == "tangents" else
                              "reverse mode: recorded rollout, one backward solve per step, all parameters at once") +
                             ", one cloth at a time",
                     "cloth_steps": ok, "cloth_gradient_steps_per_s": ok / dt, "ms_per_cloth_step": 1e3 * dt / max(ok, 1),
                     "s_per_epoch": dt}
    # the reverse mode of all 64 cloths at once: recorded together, backward passes together
    specs = [ln.perturbed_spec(spec, ln.NAMES, 500 + i) for i, (spec, _) in enumerate(items)]
    ln.batch_loss_terms_adjoint(specs[:2], frames[:2], ln.NAMES, mode="ref", tol=1e-8)   # untimed: kernels, allocations
    t0 = time.perf_counter()
    res = ln.batch_loss_terms_adjoint(specs, frames, ln.NAMES, mode="ref", tol=1e-8)
    dt = time.perf_counter() - t0
    ok = sum(r[2] for r in res)
    grads["batched_adjoint"] = np.array([r[1] for r in res])
    out["batched_adjoint"] = {"form": "reverse mode of all 64 cloths together: one launch per phase over all cloths, the "
                                      "solves (forward and backward) one launch, Jacobi preconditioner",
                              "cloth_steps": ok, "cloth_gradient_steps_per_s": ok / dt, "ms_per_cloth_step": 1e3 * dt / max(ok, 1),
                              "s_per_epoch": dt}
    out["batched_adjoint_over_tangents"] = out["batched_adjoint"]["cloth_gradient_steps_per_s"] / \
        out["tangents"]["cloth_gradient_steps_per_s"]
    out["batched_max_rel_gradient_difference"] = float(np.max(np.abs(grads["batched_adjoint"] - grads["tangents"])) /
                                                       max(np.max(np.abs(grads["tangents"])), 1e-300))
    a, t = grads["adjoint"], grads["tangents"]
    out["adjoint_over_tangents"] = out["adjoint"]["cloth_gradient_steps_per_s"] / out["tangents"]["cloth_gradient_steps_per_s"]
    out["max_rel_gradient_difference"] = float(np.max(np.abs(a - t)) / max(np.max(np.abs(t)), 1e-300))
    return out


def hbm_stream_measure(launches=20):
    """Sustained HBM bandwidth on this GPU: a STREAM-style FP64 negation of 1 GiB into
    another 1 GiB (2 GiB moved per launch), timed with HIP events on torch's stream."""
    import torch
    x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.neg(x, out=y)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(launches):
        torch.neg(x, out=y)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / launches
    nbytes = 2 * x.numel() * 8
    del x, y
    torch.cuda.empty_cache()
    return {"kernel": "torch.neg, FP64, 1 GiB -> 1 GiB", "bytes_per_launch": nbytes, "avg_launch_ms": ms,
            "achieved": nbytes / (ms * 1e-3) / 1e9, "unit": "GB/s", "peak": HBM_PEAK_GBS}


def measured_traffic(kernel, workload, iterations=None):
    """HBM bytes per launch of `kernel` from the committed counter summary (FETCH_SIZE +
    WRITE_SIZE runs, scaled by the copy calibration), or None when there is none.  A
    solve kernel's summary gives bytes per CG iteration; `iterations` scales it to the
    launch being reported."""
    path = os.path.join(ROOT, "profiles", "%s_%s_pmc.md" % (kernel, workload))
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
        if line.startswith("traffic_bytes_per_iteration:") and iterations:
            return float(line.split(":", 1)[1]) * iterations
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="twill128")
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-only", action="store_true", help="time only the zero-guess CG path (profiling runs)")
    ap.add_argument("--warm-only", action="store_true",
                    help="time only the learned warm-start line (`value`'s path; profiling runs), print its steps/s")
    ap.add_argument("--batch-steps", type=int, default=20, help="steps of the 64-cloth batch line (0: skip it)")
    ap.add_argument("--large-steps", type=int, default=10, help="steps of the 256x256 line (0: skip it)")
    ap.add_argument("--tangent-steps", type=int, default=10,
                    help="steps of the differentiation line (8 parameter tangents; 0: skip it)")
    ap.add_argument("--cg", default="persistent", choices=["persistent", "fused", "reference"],
                    help="CG form of the timed runs (the reference form is also timed once, zero guess)")
    ap.add_argument("--precond", default="block2", choices=["block2", "jacobi"],
                    help="preconditioner of the persistent-CG runs (the Jacobi one, the reference solver's, is "
                         "also timed once, zero guess)")
    ap.add_argument("--gradient-steps", type=int, default=5,
                    help="rollout steps of the 64-cloth parameter-gradient line (0: skip it)")
    ap.add_argument("--tangent-only", action="store_true",
                    help="time only the differentiation line and print it (profiling runs)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="time only the reverse-mode line on the workload and the 64-cloth gradient line, and print "
                         "them (profiling runs)")
    ap.add_argument("--copy-calibrate", action="store_true",
                    help="first run 5 launches of a 256 MiB -> 256 MiB FP64 negation, the known-size kernel "
                         "that scales FETCH_SIZE / WRITE_SIZE counter runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build measures a single MI355X (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import __graft_entry__
    ys = __graft_entry__.load_package()
    if args.copy_calibrate:
        import torch
        x = torch.ones(32 << 20, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            torch.neg(x, out=y)
        torch.cuda.synchronize()
        del x, y
    import yarnstep.warmstart as ws
    cs = ys.cloth_spec
    spec, state, ext, stable_steps = cs.workload(args.workload)
    if args.warmup + args.steps > stable_steps:
        # every timed run starts from rest; beyond this many steps the cloth's implicit step
        # loses definiteness (compressed yarn segments), in the oracle as on the GPU
        print("bench.py: the %s workload stays stable for %d steps from rest; --warmup + --steps must not exceed "
              "that" % (args.workload, stable_steps), file=sys.stderr)
        sys.exit(2)
    h = spec["h"]
    tol = 1e-8
    if args.tangent_only:
        print(json.dumps({"tangents": tangent_measure(ys, spec, state, args.tangent_steps, args.precond)}), flush=True)
        return
    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_measure(ys, spec, state, args.tangent_steps, args.precond)}), flush=True)
        if args.gradient_steps > 0:
            print(json.dumps({"batch_gradient": gradient_measure(ys, args.gradient_steps, args.precond)}), flush=True)
        return
    if args.plain_only:
        net = None
    elif os.path.exists(ws.WEIGHTS):
        net = ws.WarmStartMLP.load(ws.WEIGHTS)
    else:   # train on the GPU before the timed region
        X, T, M = ws.generate(ws.training_cloths(6), steps=40, skip=5)
        net, _ = ws.fit(X, T, M, epochs=16)

    def timed(warm, variant=None, precond=None, guess=None):
        cloth = ys.Cloth(spec)
        try:
            cloth.set_cg_variant(variant or args.cg)
        except ys.YcError:   # the persistent form needs every tile resident at once
            args.cg = "fused"
            args.precond = "jacobi"
            cloth.set_cg_variant(args.cg)
        if (variant or args.cg) == "persistent":
            cloth.set_preconditioner(precond or args.precond)
        cloth.set_ClothState(state)
        cloth.set_ext_force(ext)
        if warm:
            run = lambda n: cloth.run_warm(net, h, n, tol=tol)   # noqa: E731
        elif guess:
            run = lambda n: cloth.run_guess(guess, h, n, tol=tol)   # noqa: E731
        else:
            run = lambda n: cloth.run(h, n, tol=tol)   # noqa: E731
        run(args.warmup)
        t0 = time.perf_counter()
        iters, rc = run(args.steps)
        dt = time.perf_counter() - t0
        if rc != 0:
            print("warning: a CG solve stopped at maxit above tolerance", file=sys.stderr)
        return cloth, dt, iters

    if args.warm_only and net is not None:
        cloth, dt_w, it_w = timed(True)
        cloth.close()
        net.close()
        import torch
        torch.cuda.synchronize()
        print(json.dumps({"warm_only": {"workload": args.workload, "steps": args.steps, "warmup": args.warmup,
                                        "steps_per_s": args.steps / dt_w, "cg_iters_per_step": it_w / args.steps}}),
              flush=True)
        return

    ref_form = None
    if args.cg != "reference" and not args.plain_only:
        c_ref, dt_ref, it_ref = timed(False, "reference")
        c_ref.close()
        ref_form = {"steps_per_s": args.steps / dt_ref, "cg_iters_per_step": it_ref / args.steps}
    jacobi = None
    if args.cg == "persistent" and args.precond != "jacobi" and not args.plain_only:
        c_j, dt_j, it_j = timed(False, "persistent", "jacobi")
        c_j.close()
        jacobi = {"steps_per_s": args.steps / dt_j, "cg_iters_per_step": it_j / args.steps}
    cloth, dt_plain, it_plain = timed(False)
    # the stencil SpMV (the reference-form CG's product), timed with HIP events on the
    # context stream
    cloth.assemble(h)
    spmv_ms = cloth.time_spmv(200)
    spmv_bytes = cloth.spmv_bytes()
    spmv_gbs = spmv_bytes / (spmv_ms * 1e-3) / 1e9
    # the dominant kernel: one persistent-CG solve per step (k_cgp), timed with an event
    # pair around each solve's launch; its algorithmic bytes are the solve's
    # iterations times one PCG iteration streamed from memory
    dominant = None
    if args.cg == "persistent":
        cloth.assemble(h)
        solve_ms, solve_iters = cloth.time_solve(tol=tol, reps=10)
        # algorithmic bytes of one Jacobi-PCG iteration streamed from memory (A once, 15
        # vector passes of n doubles): 8 nnz + 120 n.  The block-Jacobi preconditioner's
        # rows add 152 n; that unit is reported beside it, labelled.
        n = cloth.n
        it_jacobi = spmv_bytes - 16.0 * n + 120.0 * n
        it_block = cloth.cg_iteration_bytes() if args.precond == "block2" else it_jacobi
        solve_bytes = solve_iters * it_jacobi
        bj = args.precond == "block2"
        traffic = measured_traffic("k_cgp_bj" if bj else "k_cgp", args.workload, solve_iters)
        secs = solve_ms * 1e-3
        dominant = {"bound": "hbm", "achieved": solve_bytes / secs / 1e9, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "traffic": traffic,
                    "kernel": "k_cgp (one persistent %s-PCG solve)" % ("2x2-block-Jacobi" if bj else "Jacobi"),
                    "bytes_per_launch": solve_bytes, "bytes_unit": "Jacobi-PCG iteration: 8 nnz + 120 n",
                    "avg_launch_ms": solve_ms, "iterations_per_launch": solve_iters,
                    "us_per_iteration": 1e3 * solve_ms / max(solve_iters, 1)}
        dominant["frac"] = dominant["achieved"] / HBM_PEAK_GBS
        dominant["frac_algorithmic_jacobi"] = dominant["frac"]
        if bj:
            dominant["frac_algorithmic_block_jacobi"] = solve_iters * it_block / secs / 1e9 / HBM_PEAK_GBS
        # HBM bytes the counters saw (FETCH_SIZE + WRITE_SIZE, copy-calibrated) over the
        # same launch time: the fraction of the HBM peak the kernel really moves
        dominant["frac_counter"] = traffic / secs / 1e9 / HBM_PEAK_GBS if traffic else None
    cloth.close()
    guesses = None
    if net is not None:
        cloth, dt_warm, it_warm = timed(True)
        # the guess's cost: x_q, A x_q and the one-launch features + network (the last
        # alone is the network's line)
        cloth.assemble(h)
        guess_ms, mlp_ms = cloth.time_warm_guess(net, 100)
        cloth.close()
        if args.cg == "persistent":
            # the free guesses the network has to beat, same preconditioner and tolerance
            guesses = {"zero": {"steps_per_s": args.steps / dt_plain, "cg_iters_per_step": it_plain / args.steps},
                       "mlp": {"steps_per_s": args.steps / dt_warm, "cg_iters_per_step": it_warm / args.steps}}
            for g in ("prev", "extrap", "quad"):
                cg_, dt_g, it_g = timed(False, guess=g)
                cg_.close()
                guesses[g] = {"steps_per_s": args.steps / dt_g, "cg_iters_per_step": it_g / args.steps}
    else:
        dt_warm, it_warm, mlp_ms, guess_ms = dt_plain, it_plain, None, None
    mlp_tflops = MLP_FLOP_PER_NODE * spec["width"] * spec["length"] / (mlp_ms * 1e-3) / 1e12 if net else None
    steps_per_s = args.steps / dt_warm
    plain_steps_per_s = args.steps / dt_plain
    dt = dt_warm
    iters = it_warm
    out = {
        "metric": "implicit steps/s at 128x128 crossings: warm-start vs plain CG, same residual",
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "%s: %dx%d crossings, %s weave, %d yarns, seeded wind, h=%g, PCG to %g relative residual" % (
            args.workload, spec["width"], spec["length"], spec["pattern"], len(spec["yarns"]), h, tol),
            "solver": "learned warm start (32-128-128-128-5 BF16 MLP) + %s-PCG (%s form)" % (
                "2x2-block-Jacobi" if args.precond == "block2" else "Jacobi", args.cg),
            "cg_iters_per_step": it_warm / args.steps,
            "plain_cg_steps_per_s": plain_steps_per_s, "plain_cg_ms_per_step": 1e3 * dt_plain / args.steps,
            "plain_cg_iters_per_step": it_plain / args.steps,
            "warm_over_plain": steps_per_s / plain_steps_per_s,
            # against the reference's own method: zero-guess Jacobi-PCG (persistent form)
            "warm_over_plain_jacobi": steps_per_s / jacobi["steps_per_s"] if jacobi else None,
            "initial_guesses": guesses,
            "mlp_over_extrapolation": (guesses["mlp"]["steps_per_s"] / guesses["extrap"]["steps_per_s"]) if guesses else None,
            "plain_cg_reference_form": ref_form,
            "plain_cg_jacobi_persistent": jacobi,
            "mlp": {"bound": "mfma", "achieved": mlp_tflops, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": mlp_tflops / BF16_PEAK_TFLOPS, "avg_launch_ms": mlp_ms,
                    "kernel": "k_mlp_warm (features, network and correction, one launch)",
                    "warm_guess_ms": guess_ms, "correction_scale": getattr(net, "scale", None),
                    "rows": spec["width"] * spec["length"]} if net else None},
        "roofline": dominant,
    }
    spmv = {"bound": "hbm", "achieved": spmv_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": spmv_gbs / HBM_PEAK_GBS, "traffic": measured_traffic("k_spmv", args.workload),
            "kernel": "k_spmv", "bytes_per_launch": spmv_bytes, "avg_launch_ms": spmv_ms}
    if dominant is None:
        out["roofline"] = spmv
    else:
        out["config"]["spmv"] = spmv
    out["config"]["hbm_stream"] = hbm_stream_measure()
    if args.batch_steps > 0 and not args.plain_only:
        # Jacobi: two solve blocks per CU, twice the resident tiles of the block-Jacobi form,
        # which makes the batch faster despite the iterations (6.6 ms against 7.1 ms per step)
        out["config"]["batch"] = batch_measure(ys, args.batch_steps, precond="jacobi")
    if args.large_steps > 0 and not args.plain_only:
        out["config"]["homo256"] = large_measure(ys, args.large_steps)
    if args.tangent_steps > 0 and not args.plain_only and args.cg == "persistent":
        out["config"]["tangents"] = tangent_measure(ys, spec, state, args.tangent_steps, args.precond)
        out["config"]["adjoint"] = adjoint_measure(ys, spec, state, args.tangent_steps, args.precond)
    if args.gradient_steps > 0 and not args.plain_only and args.cg == "persistent":
        out["config"]["batch_gradient"] = gradient_measure(ys, args.gradient_steps, args.precond)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cs, args.workload, args.cpu_steps)
    # release every HIP object before the interpreter shuts down (no destructor runs
    # during teardown, after a profiler's own exit handlers)
    if net is not None:
        net.close()
    import torch
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
