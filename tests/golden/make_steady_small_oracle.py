"""Writes tests/golden/steady_small_oracle.npz: the CPU oracle's FP64
natural-order ILU0-PCG solve of the first outer iteration of the
3 x 100 x 100 steady model (configs.steady_small(), default seed), all
30 000 heads, with the iteration count.  Seconds of one CPU core."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from modflow6_amd import configs  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    case = configs.steady_small()
    sim = O.GwfSimulation(case)
    sim.kstp = 1
    sim.advance(1.0, True)
    sim.formulate(1)
    d = sim.dis
    O.lib().orc_sln_ls_fixes(sim.neq, d.ia, d.ja, sim.ibound, 1, sim.amat,
                             sim.rhs, sim.x, sim.xtemp)
    icn, it, l2, stop, _ = sim.imslinear.apply(sim.amat, sim.rhs, sim.x, kstp=1)
    np.savez(os.path.join(HERE, "steady_small_oracle.npz"), heads=sim.x,
             iterations=np.int64(it), icnvg=np.int64(icn))


if __name__ == "__main__":
    main()
