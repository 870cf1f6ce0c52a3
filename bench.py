"""Benchmark: implicit yarn-level steps/s on one MI355X.

Workload: the 128x128 two-yarn twill cloth with seeded wind (cloth_spec.workload
"twill128"), h = 1 ms, Jacobi-PCG to a 1e-8 relative residual from a zero guess,
the reference's step (assemble, solve, update) per step, FP64.  Inputs live in HBM
before the timed region.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def cpu_baseline(cs, name, nsteps):
    """The CPU oracle (single thread) on the first nsteps steps of the same workload."""
    from oracle.oracle import OracleCloth
    spec, state, ext, _ = cs.workload(name)
    c = OracleCloth(spec)
    c.set_state(state)
    h = spec["h"]
    c.step(h, ext, tol=1e-8, maxit=20000)  # setup / first step untimed
    t0 = time.perf_counter()
    for _ in range(nsteps):
        c.step(h, ext, tol=1e-8, maxit=20000)
    dt = time.perf_counter() - t0
    return {"value": nsteps / dt, "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": "steps 2..%d of the %s workload, oracle single-thread FP64 CG to 1e-8" % (nsteps + 1, name)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="twill128")
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build measures a single MI355X (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import __graft_entry__
    ys = __graft_entry__.load_package()
    cs = ys.cloth_spec
    spec, state, ext, _ = cs.workload(args.workload)
    h = spec["h"]
    cloth = ys.Cloth(spec)
    cloth.set_ClothState(state)
    cloth.set_ext_force(ext)
    tol = 1e-8
    it_w, rc = cloth.run(h, args.warmup, tol=tol)
    t0 = time.perf_counter()
    iters, rc = cloth.run(h, args.steps, tol=tol)
    dt = time.perf_counter() - t0
    if rc != 0:
        print("warning: a CG solve stopped at maxit above tolerance", file=sys.stderr)
    steps_per_s = args.steps / dt
    # dominant kernel: stencil SpMV, timed with HIP events on the context stream
    spmv_ms = cloth.time_spmv(200)
    spmv_bytes = cloth.spmv_bytes()
    achieved = spmv_bytes / (spmv_ms * 1e-3) / 1e9
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
        "config": {"workload": "%s: %dx%d crossings, %s weave, %d yarns, h=%g, plain zero-guess Jacobi-PCG to %g" % (
            args.workload, spec["width"], spec["length"], spec["pattern"], len(spec["yarns"]), h, tol),
            "solver": "plain-cg", "cg_iters_per_step": iters / args.steps},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_spmv", "bytes_per_launch": spmv_bytes, "avg_launch_ms": spmv_ms},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cs, args.workload, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
