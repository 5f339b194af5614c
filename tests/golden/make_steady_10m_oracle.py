"""Writes tests/golden/steady_10m_oracle.json / steady_10m_heads_sample.npz:
the CPU oracle's FP64 natural-order ILU0-PCG solve of the first outer
iteration of the 10 x 1000 x 1000 steady workload (configs.steady_10m(),
default seed).  About half an hour of one CPU core; run once, offline.

The JSON keeps the iteration count and head statistics; the npz keeps the
heads of 4096 seeded cells (the full field is 80 MB)."""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from modflow6_amd import configs  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    case = configs.steady_10m()
    sim = O.GwfSimulation(case)
    sim.kstp = 1
    sim.advance(1.0, True)
    sim.formulate(1)
    d = sim.dis
    O.lib().orc_sln_ls_fixes(sim.neq, d.ia, d.ja, sim.ibound, 1, sim.amat,
                             sim.rhs, sim.x, sim.xtemp)
    t0 = time.time()
    icn, it, l2, stop, _ = sim.imslinear.apply(sim.amat, sim.rhs, sim.x, kstp=1)
    write(icn, it, l2, time.time() - t0, sim.x)


def write(icn, it, l2, secs, x):
    json.dump(dict(workload="steady_10m 10x1000x1000 seed 20260102, first outer "
                            "iteration from strt, CG + natural-order ILU0, "
                            "rclose = dvclose = 1e-6",
                   icnvg=int(icn), iterations=int(it), l2norm0=float(l2),
                   solve_s_one_core=float(secs), head_sum=float(x.sum()),
                   head_min=float(x.min()), head_max=float(x.max())),
              open(os.path.join(HERE, "steady_10m_oracle.json"), "w"), indent=1)
    rng = np.random.default_rng(4096)
    cells = np.sort(rng.choice(x.size, 4096, replace=False))
    np.savez(os.path.join(HERE, "steady_10m_heads_sample.npz"), cells=cells,
             heads=x[cells])


if __name__ == "__main__":
    if len(sys.argv) > 2 and sys.argv[1] == "--from-npy":
        # reuse a finished run: heads .npy + its json summary
        x = np.load(sys.argv[2])
        meta = json.load(open(sys.argv[3]))
        write(meta["icnvg"], meta["iterations"], meta["l2norm0"], meta["solve_s"], x)
    else:
        main()
