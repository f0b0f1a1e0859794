"""Write the golden fixtures of tests/golden/ from the FP64 oracle.

    python tests/golden/make_golden.py small    # 32^3 cross-ply fixture (4 layers, r 2.5, vf 0.4), seconds
    python tests/golden/make_golden.py big      # 256^3 cross-ply C_eff (16 layers, r 6, vf 0.5), ~2 h on 8 cores
                                                #   (resumable)

The microstructures come from the seeded cross-ply generator (RSA placement,
splitmix64), restated in oracle/microstructure_oracle.py; the fields are not
stored, only the generator parameters, because placement is bit-reproducible.
NumPy random streams are not used for these fixtures.
"""

import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import ls_oracle as O  # noqa: E402
from oracle import microstructure_oracle as MO  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

CONFIG1 = dict(shape=(32, 32, 32), extents=(32.0, 32.0, 32.0), nlayers=4, radius=2.5, vf=0.4, seed=1)
CONFIG2 = dict(shape=(256, 256, 256), extents=(256.0, 256.0, 256.0), nlayers=16, radius=6.0, vf=0.5, seed=2)


def fields(cfg):
    lam, mu, vf, fib = MO.crossply(cfg["shape"], cfg["extents"], cfg["nlayers"], cfg["radius"], cfg["vf"],
                                   cfg["seed"])
    return lam.astype(np.float64), mu.astype(np.float64), vf, len(fib)


def small():
    cfg = CONFIG1
    lam, mu, vf, nf = fields(cfg)
    gs = O.GridSpec(*cfg["shape"], *cfg["extents"])
    C6, its6 = O.effective_tensor(lam, mu, gs, rtol=1e-6, depth=0, maxiter=5000)
    C10, its10 = O.effective_tensor(lam, mu, gs, rtol=1e-10, depth=0, maxiter=20000)
    out = dict(config=cfg, achieved_vf=vf, n_fibres=nf, ceff_rtol1e6=C6.tolist(), iters_rtol1e6=its6,
               ceff_rtol1e10=C10.tolist(), iters_rtol1e10=its10)
    with open(os.path.join(HERE, "crossply32_ceff.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote crossply32_ceff.json", its6, its10)


def big():
    """One load case at a time; finished rows are kept in a side file so an
    interrupted run resumes where it stopped."""
    cfg = CONFIG2
    t0 = time.time()
    lam, mu, vf, nf = fields(cfg)
    gs = O.GridSpec(*cfg["shape"], *cfg["extents"])
    workers = int(os.environ.get("GOLDEN_WORKERS", len(os.sched_getaffinity(0))))
    partial = os.path.join(HERE, "crossply256_ceff.partial.json")
    done = {}
    if os.path.exists(partial):
        with open(partial) as f:
            done = {int(k): v for k, v in json.load(f).items()}
    for k in range(6):
        if k in done:
            continue
        e = np.zeros(6)
        e[k] = 1.0
        _, sigma, it, hist = O.lippmann_schwinger_rfft(
            lam, mu, e, gs, rtol=1e-9, depth=5, maxiter=2000, workers=workers,
            callback=lambda i, r: print(f"  load case {k} it {i} R {r:.3e} t={time.time() - t0:.0f}s", flush=True)
            if i % 10 == 0 else None)
        done[k] = dict(row=sigma.mean(axis=(1, 2, 3)).tolist(), iters=int(it))
        with open(partial, "w") as f:
            json.dump({str(i): v for i, v in done.items()}, f)
        print(f"load case {k}: {it} iterations, {time.time() - t0:.0f}s", flush=True)
    C = np.array([done[k]["row"] for k in range(6)])
    its = [done[k]["iters"] for k in range(6)]
    out = dict(config=cfg, achieved_vf=vf, n_fibres=nf, method="anderson m=5, float64, rtol 1e-9",
               ceff=C.tolist(), iters=its)
    with open(os.path.join(HERE, "crossply256_ceff.json"), "w") as f:
        json.dump(out, f, indent=1)
    os.remove(partial)
    print("wrote crossply256_ceff.json", its, f"{time.time() - t0:.0f}s")


if __name__ == "__main__":
    {"small": small, "big": big}[sys.argv[1]]()
