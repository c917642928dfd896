"""Writes the front-end golden fixtures: random ANML files (tests/nfa_helpers
random_anml_states) and the group dumps that the reference's own front-end
(oracle/_ref/ref_frontend_dump, built by oracle/Makefile from the reference
sources) produces for them.  Run in a container that has the reference tree:

    make -C oracle ref && python tests/golden/make_frontend_golden.py
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from ngap_amd.groups import write_anml  # noqa: E402
from tests import nfa_helpers as H  # noqa: E402

CASES = [(seed, n, gnum, maxsz) for seed, n, gnum, maxsz in
         [(0, 60, 3, -1), (1, 80, 2, -1), (2, 80, 10, -1), (3, 120, 4, 40), (4, 40, 1, -1),
          (5, 150, 7, -1)]]


def main():
    out_dir = os.path.join(HERE, "frontend")
    os.makedirs(out_dir, exist_ok=True)
    for seed, n, gnum, maxsz in CASES:
        base = os.path.join(out_dir, f"rand{seed}")
        write_anml(base + ".anml", H.random_anml_states(seed, n=n))
        dump = oracle.ref_dump(base + ".anml", gnum, maxsz)
        with open(f"{base}_g{gnum}_m{maxsz}.dump", "w") as f:
            f.write(dump)
    dump = oracle.ref_dump(os.path.join(HERE, "apple.anml"), 1)
    with open(os.path.join(out_dir, "apple_g1_m-1.dump"), "w") as f:
        f.write(dump)


if __name__ == "__main__":
    main()
