"""Writes nsad_1080p.json: the algorithmic SAD-term counts (N_sad = the sum
over the candidates the reference's search evaluates of Xl * Yl) per stage,
counted by the oracle, for the 1080p pictures bench.py times (synth.config1
seeds 100..107; 16/12 blocks, depth 4, quarter pel, lambda 10, 2 refs).

bench.py reads the file for its SAD-term roofline; tests/test_oracle.py
re-counts seed 100.  Run from the repository root:
    python3 tests/golden/make_nsad.py
"""
import json
import os
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "dirac-research_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle_py as O  # noqa: E402
import dirac_me  # noqa: E402
from dirac_me import synth  # noqa: E402

SEEDS = list(range(100, 108))


def counts(seed):
    d = synth.config1(seed)
    p = dirac_me.make_params(*d["block"], precision=2, depth=4, lam=10.0)
    t = {}
    for name, stop in (("pel", O.STOP_PEL), ("subpel", O.STOP_SUBPEL), ("mode", O.STOP_MODE)):
        t[name] = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"], stop=stop)[0].sad_terms
    return {"seed": seed, "pel": t["pel"], "subpel": t["subpel"] - t["pel"], "mode": t["mode"] - t["subpel"]}


def main():
    with ThreadPoolExecutor(max_workers=8) as ex:
        rows = list(ex.map(counts, SEEDS))
    out = {"workload": "synth.config1(seed): 1920x1080 B picture, 2 refs, blocks 16/12, depth 4, quarter pel, lambda 10",
           "definition": "sum over the candidates evaluated by the reference's search of Xl * Yl (SAD terms), "
                         "per stage: pel search (PixelMatcher), sub-pel refinement (SubpelRefine), mode decision "
                         "(ModeDecider)",
           "frames": rows}
    with open(os.path.join(ROOT, "tests", "golden", "nsad_1080p.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(rows))


if __name__ == "__main__":
    main()
