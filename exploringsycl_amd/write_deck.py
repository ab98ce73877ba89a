"""Writes a deck of the time-to-solution field for one solver, to run
bin/tealeaf on it directly (for instance under a profiler):

  python -m exploringsycl_amd.write_deck DIR --solver S [--cells N] [--max-iters K]

DIR/tea.in is the 64-rectangle field of the time-to-solution benchmark
(log10 density in [-2, 2], K = 1/density, one timestep to 1e-10) with the
keys of solver S (cg, mpcg, jpcg, fcg, fcgs); for fcg the network file is
copied to DIR/net.bin.
"""

import argparse
import os
import shutil
import sys

from .deck import SOLVER_KEYS, tts_deck


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("dir")
    ap.add_argument("--solver", choices=sorted(SOLVER_KEYS), required=True)
    ap.add_argument("--cells", type=int, default=4096)
    ap.add_argument("--max-iters", type=int, default=12000)
    a = ap.parse_args(argv)
    os.makedirs(a.dir, exist_ok=True)
    tts_deck(a.solver, a.cells, a.max_iters).write(os.path.join(a.dir, "tea.in"))
    if a.solver == "fcg":
        from .learned import WEIGHTS_PATH
        shutil.copy(WEIGHTS_PATH, os.path.join(a.dir, "net.bin"))
    return 0


if __name__ == "__main__":
    sys.exit(main())
