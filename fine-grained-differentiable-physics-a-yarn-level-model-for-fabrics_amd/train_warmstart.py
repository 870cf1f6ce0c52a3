"""Trains the warm-start network on the GPU and writes warmstart_weights.npz.

Data: the simulator's own converged solves on seeded cloths of 48x48 to 128x128
crossings (all three weaves, seeded wind), every other node of every step after the
first five.  Training: the library's BF16-MFMA forward/backward kernels with Adam; then
the scale of the network's correction (0 to 1) that needs the fewest CG iterations on a
held-out validation cloth.
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
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    steps = int(sys.argv[2]) if len(sys.argv) > 2 else 120
    epochs = int(sys.argv[3]) if len(sys.argv) > 3 else 20
    t0 = time.time()
    specs = ws.training_cloths(n)
    print("cloths:", [s["width"] for s in specs], flush=True)
    X, T, M = ws.generate(specs, steps=steps, skip=5, node_stride=2)
    print("pairs: %d (state, dv) node samples from %d solves, %.1fs" % (len(X), n * (steps - 5), time.time() - t0), flush=True)
    net, losses = ws.fit(X, T, M, epochs=epochs, batch=4096, lr=2e-3, log=lambda m: print(m, flush=True))
    s = ws.choose_scale(net, ws.validation_cloth(), log=lambda m: print(m, flush=True))
    print("correction scale %.2f (chosen on the validation cloth)" % s, flush=True)
    net.save(ws.WEIGHTS)
    print("wrote", ws.WEIGHTS, "%.1fs" % (time.time() - t0), flush=True)
    _ = ys


if __name__ == "__main__":
    main()
