"""Trains the warm-start network on the GPU and writes warmstart_weights.npz.

Data: the simulator's own converged solves on seeded cloths of 32x32 to 96x96
crossings (all three weaves, seeded wind), every node of every step after the first
five.  Training: the library's BF16-MFMA forward/backward kernels with Adam.
Usage: python train_warmstart.py [n_cloths] [steps] [epochs]
"""
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def main():
    import __graft_entry__
    ys = __graft_entry__.load_package()
    import yarnstep.warmstart as ws
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 6
    steps = int(sys.argv[2]) if len(sys.argv) > 2 else 40
    epochs = int(sys.argv[3]) if len(sys.argv) > 3 else 16
    t0 = time.time()
    specs = ws.training_cloths(n)
    print("cloths:", [s["width"] for s in specs], flush=True)
    X, T, M = ws.generate(specs, steps=steps, skip=5)
    print("pairs: %d (state, dv) node samples from %d solves, %.1fs" % (len(X), n * (steps - 5), time.time() - t0), flush=True)
    net, losses = ws.fit(X, T, M, epochs=epochs, batch=4096, lr=2e-3, log=lambda m: print(m, flush=True))
    net.save(ws.WEIGHTS)
    print("wrote", ws.WEIGHTS, "%.1fs" % (time.time() - t0), flush=True)
    _ = ys


if __name__ == "__main__":
    main()
