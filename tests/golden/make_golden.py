"""Writes the DepthTree fixtures under tests/golden/trees/ with the reference's
own front end (pareas-ast, compiled from the reference checkout into
oracle/_ref/pareas_dump by oracle/Makefile).  Inputs: the corpus files in
tests/golden/corpus/ (copied from the reference's testfiles/) and a few seeded
generated programs.  Run in the build container: python tests/golden/make_golden.py
"""
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import oracle  # noqa: E402
from ptilopsis_amd import workloads as W  # noqa: E402

GENERATED = [("gen_small_%d" % s, dict(seed=9000 + s, n_functions=6)) for s in range(3)] + [
    ("gen_stress", dict(seed=9100, n_functions=3, base=W.STRESS)),
]


def main():
    if not os.path.exists(oracle.REF_DUMP):
        oracle.build()
    out_dir = os.path.join(HERE, "trees")
    os.makedirs(out_dir, exist_ok=True)
    tmp = tempfile.mkdtemp()
    jobs = [(os.path.splitext(f)[0], os.path.join(HERE, "corpus", f)) for f in sorted(os.listdir(os.path.join(HERE, "corpus")))]
    for name, kw in GENERATED:
        src, _ = W.program_source(kw["seed"], kw["n_functions"], base=kw.get("base", W.BASE))
        p = os.path.join(tmp, name + ".ptl")
        open(p, "w").write(src)
        jobs.append((name, p))
    for name, path in jobs:
        t = oracle.ref_front_end(path, os.path.join(tmp, "dump.bin"))
        np.savez_compressed(os.path.join(out_dir, name + ".npz"), node_types=t.node_types, result_types=t.result_types,
                            parents=t.parents, depth=t.depth, child_idx=t.child_idx, node_data=t.node_data,
                            max_depth=np.array([t.max_depth]), source=np.frombuffer(open(path, "rb").read(), np.uint8))
        print(name, t.n)


if __name__ == "__main__":
    main()
