"""Train the learned Poisson preconditioner on the GPU with libsuli's own
kernels (BF16 MFMA forward and backward, Adam), from residuals that the
solver itself produces.

Data come from seeded dam-break runs on a 1.0 x 0.5 x 0.5 m tank (water
column width U(0.2, 0.4) m, height U(0.3, 0.45) m drawn with splitmix64 from
the seed; water/air 1000/1.2 kg/m^3; dt = 0.2 h / (2 sqrt(g H))):

1. the scaled residuals at every inner iteration of multigrid-preconditioned
   FCG solves;
2. then, round after round, the residuals met by FCG preconditioned with the
   network being trained, from states advanced a varying number of steps, so
   that the network learns the residuals of its own iterations (a solve that
   stalls falls back to multigrid, whose iterations are captured as well).

With --entry e >= 1 the network is the coarse solver of multigrid level e
(two-grid form, see suli_unet_set_entry): the samples are then level e's
problems (right-hand side, block density and face conductances) met by the
V-cycle in both phases, and the loss is taken on that level's operator.

    python3 train_unet.py --nx 128 --levels 6 --out weights/unet_dambreak_L6.npy
    python3 train_unet.py --nx 256 --entry 1 --levels 5 --out weights/unet_dambreak_e1_L5.npy

--eval-seed adds a held-out column geometry: after training, one projection
of that tank is solved with Jacobi, multigrid and the learned preconditioner
and the times are written to the log.

Writes the FP32 weight vector (.npy) and a JSON log beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import pysuli as S  # noqa: E402


def splitmix64(seed):
    x = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        z = z ^ (z >> 31)
        yield (z >> 11) / float(1 << 53)


def column_geometry(seed):
    r = splitmix64(seed)
    return 0.2 + 0.2 * next(r), 0.3 + 0.15 * next(r)


class Tank:
    """One GPU context whose state is reset to a new dam break per seed."""

    def __init__(self, nx, precision=S.PREC_MIXED):
        self.nx, self.ny, self.nz = nx, nx // 2, nx // 2
        self.h = 1.0 / nx
        self.s = S.Solver(self.nx, self.ny, self.nz, self.h, precision=precision, poisson=S.POISSON_FCG,
                          precond=S.PRECOND_GMG, tol_rel=1e-8)
        self.zeros = {n: np.zeros(self.s.shape(n)) for n in ("u", "v", "w", "prd1")}
        self.dt = None

    def reset(self, seed):
        width, height = column_geometry(seed)
        for n, z in self.zeros.items():
            self.s.upload(n, z)
        self.s.init_dambreak(width, height)
        self.s.contorno(1)
        self.s.contorno(3)
        self.dt = 0.2 * self.h / (2.0 * np.sqrt(9.81 * height))
        return width, height

    def steps(self, n, precond):
        self.s.set_solver(S.POISSON_FCG, precond)
        out = []
        for _ in range(n):
            _, ps = self.s.step(self.dt)
            out += [(p["iters"], p["fallbacks"]) for p in ps]
        return out


def evaluate(tank, seed, warm_steps, repeats=3):
    """ms per solve of one projection of a held-out tank with each preconditioner."""
    s = tank.s
    tank.reset(seed)
    tank.steps(warm_steps, S.PRECOND_GMG)
    dt = tank.dt
    s.level_set(dt)
    s.contorno(3)
    s.visco()
    s.convdiff()
    s.tempo(1, 0.5 * dt)
    s.contorno(2)
    p0 = s.download("prd1")
    out = {}
    for name, pc in (("jacobi", S.PRECOND_JACOBI), ("gmg", S.PRECOND_GMG), ("learned", S.PRECOND_LEARNED)):
        s.set_solver(S.POISSON_FCG, pc)
        runs = []
        for _ in range(repeats):
            s.upload("prd1", p0)
            runs.append(s.graddin(0.5 * dt))
        ms = sorted(r["ms"] for r in runs)
        out[name] = {"ms": ms[len(ms) // 2], "iters": runs[-1]["iters"], "fallbacks": runs[-1]["fallbacks"],
                     "res": runs[-1]["res"]}
    return out


def train(args):
    t0 = time.time()
    tank = Tank(args.nx)
    s = tank.s
    s.unet_configure(args.levels)
    s.unet_set_entry(args.entry)
    if args.init:
        s.unet_set_weights(np.load(args.init))
    else:
        s.unet_init(args.seed)
    log = {"nx": args.nx, "ny": tank.ny, "nz": tank.nz, "levels": args.levels, "entry": args.entry,
           "nparams": s.unet_nparams(), "rounds": []}
    seeds = list(range(args.seed, args.seed + args.seeds))
    # 1. multigrid trajectories
    s.unet_capture(args.capacity, 1)
    for sd in seeds:
        tank.reset(sd)
        tank.steps(args.warm_steps, S.PRECOND_GMG)
    s.unet_capture(0)
    if not args.init:
        loss = s.unet_train(args.train_steps, args.lr, args.seed)
        log["rounds"].append({"round": -1, "samples": s.unet_dataset_size(), "loss": loss,
                              "seconds": time.time() - t0})
        print(f"multigrid data {s.unet_dataset_size()} loss {loss:.4f} {time.time() - t0:.1f}s", flush=True)
    # 2. trajectories of the network itself
    for rnd in range(args.rounds):
        sd = seeds[rnd % len(seeds)]
        tank.reset(sd)
        advance = (rnd * 7) % (args.warm_steps + 1)
        tank.steps(advance, S.PRECOND_GMG)
        s.unet_capture(args.capacity, 2)
        its = tank.steps(args.round_steps, S.PRECOND_LEARNED)
        s.unet_capture(0)
        lr = args.lr * 0.5 ** (rnd / max(args.rounds / 4, 1))
        loss = s.unet_train(args.train_steps, lr, args.seed + rnd + 1)
        rec = {"round": rnd, "seed": sd, "advance": advance, "solves": its, "samples": s.unet_dataset_size(),
               "lr": lr, "loss": loss, "seconds": time.time() - t0}
        log["rounds"].append(rec)
        print(f"round {rnd} seed {sd} +{advance} solves {its} samples {rec['samples']} lr {lr:.1e} "
              f"loss {loss:.4f} {rec['seconds']:.1f}s", flush=True)
        if rec["samples"] >= args.capacity:  # keep the newest trajectories: restart the set
            s.unet_clear()
    if args.eval_seed is not None:
        log["eval"] = {"seed": args.eval_seed, **evaluate(tank, args.eval_seed, args.warm_steps)}
        print("eval", json.dumps(log["eval"]), flush=True)
    w = s.unet_get_weights()
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, w)
    log["seconds"] = time.time() - t0
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(log, f, indent=1)
    s.close()
    print(f"saved {args.out} ({w.size} parameters) in {log['seconds']:.1f}s", flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--nx", type=int, default=128, help="grid nx (ny = nz = nx/2)")
    ap.add_argument("--levels", type=int, default=6)
    ap.add_argument("--entry", type=int, default=0, help="multigrid level the network solves (0: whole grid)")
    ap.add_argument("--eval-seed", type=int, default=None, help="held-out geometry timed after training")
    ap.add_argument("--seeds", type=int, default=4, help="number of column geometries")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--warm-steps", type=int, default=8, help="multigrid steps captured per geometry")
    ap.add_argument("--rounds", type=int, default=12)
    ap.add_argument("--round-steps", type=int, default=3, help="learned steps captured per round")
    ap.add_argument("--train-steps", type=int, default=2000, help="Adam steps per round")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--capacity", type=int, default=3000, help="training samples held on the GPU")
    ap.add_argument("--init", default=None, help="start from these weights (fine-tuning)")
    ap.add_argument("--out", default=os.path.join(HERE, "weights", "unet_dambreak_L6.npy"))
    train(ap.parse_args())


if __name__ == "__main__":
    main()
