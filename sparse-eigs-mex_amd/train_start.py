"""Train the learned-start network and measure what it gains.

Training family: 3-D graph Laplacians on an N^3 grid (N = 32) whose edge
coefficients are exp of a seeded Gaussian random field (correlation length 8,
log-std 1) plus 1e-3 I (gen.grf_lap3d).  For every training instance the plain
FP64 GPU solver returns the k = 16 smallest eigenvectors; the network is then
trained on the GPU (its own forward/backward kernels, Adam) to maximise the
part of that eigenspace its 32-vector output block captures.  Instances are
used as they are produced plus a replay buffer of recent ones, so memory stays
bounded at any instance count.

Evaluation: held-out seeds of the same family at 64^3, k = 16, tol 1e-10, the
block eigensolver (precision 2) and the Lanczos path (precision 0), each with
the uniform start and with the learned start; reports operator applications,
iterations and time.  Prints one JSON line at the end.

    python sparse-eigs-mex_amd/train_start.py --instances 2000 --out sparse-eigs-mex_amd/models/start_grf32.bin
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def _pkg():
    import importlib.util
    if "speigs_amd" in sys.modules:
        return sys.modules["speigs_amd"]
    spec = importlib.util.spec_from_file_location("speigs_amd", HERE / "__init__.py",
                                                  submodule_search_locations=[str(HERE)])
    m = importlib.util.module_from_spec(spec)
    sys.modules["speigs_amd"] = m
    spec.loader.exec_module(m)
    return m


def eigvecs(pkg, A, k, tol=1e-8):
    r = pkg.speigs(A, k, "sm", opts={"tol": tol})
    if r.rc != 0:
        raise RuntimeError(f"plain solver failed: {r.rc}")
    return r.V


def train(pkg, args, log):
    rng = np.random.default_rng(args.seed)
    model = pkg.StartModel(3, seed=args.seed)
    replay = []
    losses = []
    t0 = time.time()
    step = 0
    for i in range(args.instances):
        A = pkg.gen.grf_lap3d(args.N, seed=args.seed0 + i)
        U = eigvecs(pkg, A, args.k)
        replay.append((A, U))
        if len(replay) > args.replay:
            replay.pop(0)
        frac = i / max(1, args.instances - 1)
        lr = args.lr * (1.0 - 0.9 * frac)               # linear decay to 10%
        for s in range(args.steps):
            A2, U2 = (A, U) if s == 0 else replay[int(rng.integers(len(replay)))]
            losses.append(model.train_step(A2, U2, seed=int(rng.integers(1 << 62)), lr=lr))
            step += 1
        if (i + 1) % args.report == 0 or i + 1 == args.instances:
            log(f"[train] instance {i + 1}/{args.instances} steps {step} "
                f"mean loss (last {args.report * args.steps}) {np.mean(losses[-args.report * args.steps:]):.4f} "
                f"elapsed {time.time() - t0:.0f}s")
    return model, losses


def evaluate(pkg, model, args, log):
    rows = []
    for j in range(args.eval):
        A = pkg.gen.grf_lap3d(args.eval_N, seed=args.eval_seed0 + j)
        for prec in (2, 0):
            res = {}
            for start in ("uniform", "learned"):
                ho = {"precision": prec}
                if start == "learned":
                    ho.update({"start": 2, "model": model.handle.value})
                t = time.time()
                r = pkg.speigs(A, args.k, "sm", opts={"tol": 1e-10}, hip_opts=ho)
                dt = time.time() - t
                res[start] = (r, dt)
                log(f"[eval] inst {j} precision {prec} {start}: {dt:.3f}s nops {r.info['nops']} "
                    f"restarts {r.info['restarts']} t_iter {r.info['t_iter']:.3f} flag {r.info['flag']}")
            ru, rl = res["uniform"][0], res["learned"][0]
            rows.append({"instance": j, "precision": prec,
                         "nops_uniform": ru.info["nops"], "nops_learned": rl.info["nops"],
                         "iters_uniform": ru.info["restarts"], "iters_learned": rl.info["restarts"],
                         "t_iter_uniform": ru.info["t_iter"], "t_iter_learned": rl.info["t_iter"],
                         "total_uniform": res["uniform"][1], "total_learned": res["learned"][1],
                         "max_lambda_diff": float(np.abs(ru.lam - rl.lam).max() / np.abs(ru.lam).max())})
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--instances", type=int, default=2000)
    ap.add_argument("--N", type=int, default=32)
    ap.add_argument("--k", type=int, default=16)
    ap.add_argument("--steps", type=int, default=3, help="training steps per new instance")
    ap.add_argument("--replay", type=int, default=64)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--seed0", type=int, default=10_000)
    ap.add_argument("--report", type=int, default=50)
    ap.add_argument("--eval", type=int, default=3)
    ap.add_argument("--eval-N", dest="eval_N", type=int, default=64)
    ap.add_argument("--eval-seed0", dest="eval_seed0", type=int, default=900_000)
    ap.add_argument("--out", default=str(HERE / "models" / "start_grf32.bin"))
    ap.add_argument("--load", default=None, help="skip training, evaluate these parameters")
    args = ap.parse_args()
    pkg = _pkg()

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    if args.load:
        model = pkg.StartModel(3, seed=args.seed)
        model.load(args.load)
        losses = []
    else:
        model, losses = train(pkg, args, log)
        out = Path(args.out)
        out.parent.mkdir(parents=True, exist_ok=True)
        model.save(out)
        meta = {"family": f"grf_lap3d({args.N}, corr 8, logstd 1, shift 1e-3)", "k": args.k,
                "instances": args.instances, "steps": len(losses), "layers": 3, "width": 32,
                "first_loss": float(np.mean(losses[:50])), "final_loss": float(np.mean(losses[-200:]))}
        out.with_suffix(".json").write_text(json.dumps(meta, indent=1) + "\n")
        log(f"[train] saved {out} ({model.nparams} parameters)")
    rows = evaluate(pkg, model, args, log) if args.eval > 0 else []
    summary = {"final_loss": float(np.mean(losses[-200:])) if losses else None, "eval": rows}
    print(json.dumps(summary))


if __name__ == "__main__":
    main()
