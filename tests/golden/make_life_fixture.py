"""Writes tests/golden/life_16384x16384.json: SHA-256 and popcount of the
Game of Life state of the 16384 x 16384 Bernoulli(0.3) input (seed
20261202, dead Cube edge) after 0, 1, 31, 100 and 1000 generations, computed
by the C oracle (test infrastructure) on the CPU.  The GPU test compares the
device result of the full-size run with these numbers.

Run from the repository root after building the oracle:
    python tests/golden/make_life_fixture.py [--threads N]
(~3 minutes with 8 threads)."""
import argparse
import hashlib
import json
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(REPO))
sys.path.insert(0, str(REPO / "tests"))

from libgeodecomp_amd import synthetic  # noqa: E402
import oracle_lib  # noqa: E402

N = 16384
CHECKPOINTS = (0, 1, 31, 100, 1000)


def digest(grid):
    g = np.ascontiguousarray(grid, dtype=np.uint8)
    return hashlib.sha256(g.tobytes()).hexdigest(), int(np.count_nonzero(g))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    args = ap.parse_args()
    oracle_lib.load()
    grid = synthetic.life_initial(N, N, 0.3, synthetic.DEFAULT_SEED)
    out = {"grid": [N, N], "density": 0.3, "seed": synthetic.DEFAULT_SEED, "edge": "dead (Cube)",
           "layout": "uint8, row-major (y, x), x fastest", "checkpoints": {}}
    done = 0
    t0 = time.time()
    for gen in CHECKPOINTS:
        if gen > done:
            grid = oracle_lib.life2d(grid, gen - done, 0, args.threads)
            done = gen
        sha, pop = digest(grid)
        out["checkpoints"][str(gen)] = {"sha256": sha, "popcount": pop}
        print(f"generation {gen}: popcount {pop} sha256 {sha[:16]}... ({time.time() - t0:.0f} s)", flush=True)
    dst = Path(__file__).resolve().parent / "life_16384x16384.json"
    dst.write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", dst)


if __name__ == "__main__":
    main()
