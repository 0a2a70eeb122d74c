"""Train the ADR warm-start network on batched-ADR load-step solutions.

    python -m pd_gpu1_amd.train_warmstart --plates 20000 --increments 1 \\
        --epochs 3 --out pd_gpu1_amd/data/mpnn_warmstart.npy

Training plates come from the seeded plate family (seed 0x5EED0004 + k, side
40..200 particles, crack length 0.1..0.5 of the side, angle -30..30 degrees);
each is loaded in ``--increments`` steps of 1e-6 m and relaxed by the batched
ADR to relative residual 1e-6, ``--chunk`` plates per batched launch.  Every
converged increment that the network predicts is one sample.  With one
increment per plate the samples are kept compactly (one graph per plate
side, packed intact masks, float32 targets), which is what lets tens of
thousands of plates fit in host memory.  The parameters are written as
float32 (.npy) with a JSON log beside them.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import batch as PB
from . import warmstart as W
from .mpnn import Mpnn


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--plates", type=int, default=512)
    ap.add_argument("--increments", type=int, default=1)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--n-min", type=int, default=40)
    ap.add_argument("--n-max", type=int, default=200)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-end", type=float, default=1e-4)
    ap.add_argument("--seed", type=int, default=0x5EED0004)
    ap.add_argument("--out", default="pd_gpu1_amd/data/mpnn_warmstart.npy")
    ap.add_argument("--log", default=None)
    ap.add_argument("--chunk", type=int, default=512, help="plates per batched-ADR launch")
    ap.add_argument("--init", default=None, help="start from these parameters (.npy) instead of the seeded ones")
    ap.add_argument("--max-seconds", type=float, default=None,
                    help="training time limit (the learning rate decays to --lr-end by then)")
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    specs = PB.plate_family(a.plates, seed_base=a.seed, n_min=a.n_min, n_max=a.n_max, dx=W.DX)
    p = PB.plate_params(W.DX)
    compact = a.increments == 1
    recs, iters, t_data, sides = [], [], 0.0, {}
    for k in range(0, a.plates, a.chunk):  # one batched launch per chunk of plates
        if compact:
            r, it, td = W.generate_first(specs[k:k + a.chunk], sides, p.delta, log=log)
            it = it[None, :]
        else:
            r, it, td = W.generate(specs[k:k + a.chunk], a.increments, log=log)
        recs += r
        iters.append(it)
        t_data += td
    iters = np.concatenate(iters, axis=1)
    log(f"[data] {a.plates} plates x {a.increments} increments in {t_data:.1f} s")
    t0 = time.perf_counter()
    steps = None
    with Mpnn(seed=a.seed) as net:
        if a.init:
            net.set_params(np.load(a.init))
        if compact:
            hist, steps = W.train_first(net, recs, sides, a.epochs, lr=a.lr, lr_end=a.lr_end, seed=a.seed, log=log,
                                        max_seconds=a.max_seconds)
        else:
            hist = W.train(net, recs, p.delta, a.epochs, lr=a.lr, lr_end=a.lr_end, log=log)
        theta = net.get_params()
    t_train = time.perf_counter() - t0
    np.save(a.out, theta.astype(np.float32))
    info = dict(command="python -m pd_gpu1_amd.train_warmstart " + " ".join(argv if argv is not None else sys.argv[1:]),
                plates=a.plates, increments=a.increments, samples=len(recs) if compact else a.plates * a.increments,
                epochs=a.epochs, steps=steps, n_range=[a.n_min, a.n_max], seed=a.seed, lr=[a.lr, a.lr_end],
                init=a.init, max_seconds=a.max_seconds, loss_per_epoch=hist, data_seconds=t_data,
                train_seconds=t_train, adr_iterations_median_per_increment=[int(np.median(r)) for r in iters])
    with open(a.log or a.out.replace(".npy", ".json"), "w") as f:
        json.dump(info, f, indent=1)
    log(f"[train] {t_train:.1f} s; parameters written to {a.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
