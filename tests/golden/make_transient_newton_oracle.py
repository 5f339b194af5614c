"""Writes tests/golden/transient_newton_oracle.json / .npz: the CPU
oracle's run of the transient Newton workload (configs.transient_newton(),
5 x 500 x 500, 10 periods x 5 steps, BiCGSTAB + ILU0, NEWTON
UNDER_RELAXATION, default seed): outer iterations per time step and the
heads of 4096 seeded cells after the first and the last time step.  About
five minutes of one CPU core; run once, offline.  Rerun after any change to
the oracle's GWF formulate or outer loop: the device tests compare with it."""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from modflow6_amd import configs  # noqa: E402
from oracle import oracle as O  # noqa: E402


def host_cpu():
    """CPU model of the host the oracle ran on (its one-core time depends on it)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    case = configs.transient_newton()
    sim = O.GwfSimulation(case)
    rng = np.random.default_rng(4096)
    cells = np.sort(rng.choice(case.nodes, 4096, replace=False))
    first = None
    t0 = time.time()
    nstep = 0
    for kper, (perlen, nstp, steady) in enumerate(case.periods, start=1):
        sim.kper = kper
        sim.kstp = 0
        for _ in range(nstp):
            sim.run_timestep(perlen / nstp, steady)
            nstep += 1
            if first is None:
                first = sim.x[cells].copy()
            print(f"step {nstep}: {len(sim.log[-1])} outer, "
                  f"{sum(r['inner'] for r in sim.log[-1])} inner, "
                  f"{time.time() - t0:.0f} s", flush=True)
    json.dump(dict(workload="transient_newton 5x500x500 seed 20260103, 10 periods x 5 "
                            "steps of 6 d, BiCGSTAB + ILU0, NEWTON UNDER_RELAXATION",
                   outer_per_step=[len(s) for s in sim.log],
                   inner_per_step=[int(sum(r["inner"] for r in s)) for s in sim.log],
                   converged=[bool(s[-1]["converged"]) for s in sim.log],
                   run_s_one_core=time.time() - t0, host=host_cpu(),
                   head_min=float(sim.x.min()), head_max=float(sim.x.max())),
              open(os.path.join(HERE, "transient_newton_oracle.json"), "w"), indent=1)
    np.savez(os.path.join(HERE, "transient_newton_heads_sample.npz"), cells=cells,
             heads_first_step=first, heads_last_step=sim.x[cells])


if __name__ == "__main__":
    main()
