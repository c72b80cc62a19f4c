"""Write the golden fixture of the reference's output files (run where
/root/reference is present; the fixture then travels with the repository).

The compiled reference (oracle/Makefile.ref: unmodified sources + state-dump
hook, which with MPD_REF_WRITERS=1 also calls the reference's own
Output_Flowfield / Output_Diag / Output_MidX / Output_MidY at each dumped step)
runs the shipped AC+HCZ drop case for 1000 steps.  Recorded: the name and
SHA-256 of Case.ark and of every flow-field and profile file at steps 0 and 100,
the full text of the SumRho and Residual records written at step 1000, and the
step-0 flow field of the same case read from a mesh file whose cells are
numbered in a shuffled order (seed 7).

    python tests/golden/make_writers_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def sha(path):
    with open(path, "rb") as fh:
        return hashlib.sha256(fh.read()).hexdigest()


def main():
    import ref_run
    ref_run.run("hcz", [0, 100], 1001, flowfield=True, writers=True)
    ff = os.path.join(ref_run.HERE, "_ref", "hcz", "run", "FlowField")
    rec = {"config": "hcz", "nx": 128, "ny": 128, "steps": [0, 100], "files": {}, "text": {}}
    for sub in ("", "global", "Convergence"):
        d = os.path.join(ff, sub)
        for name in sorted(os.listdir(d)):
            p = os.path.join(d, name)
            if not os.path.isfile(p):
                continue
            key = (sub + "/" if sub else "") + name
            if name.startswith(("SumRho_", "Residual_")):
                with open(p) as fh:
                    rec["text"][key] = fh.read()
            else:
                rec["files"][key] = {"sha256": sha(p), "bytes": os.path.getsize(p)}
    # the same case on a mesh file whose cells are numbered in a shuffled order:
    # the reference writes its flow field in that (CellArray) order
    import numpy as np
    perm = np.random.default_rng(7).permutation(128 * 128)
    ref_run.run("hcz", [], 1, flowfield=True, writers=True, cell_order=perm)
    p = os.path.join(ff, "global", "step0Ma0_Car128_CFL0.5_T0.dat")
    rec["permuted"] = {"seed": 7, "file": "global/step0Ma0_Car128_CFL0.5_T0.dat", "sha256": sha(p),
                       "bytes": os.path.getsize(p)}
    with open(os.path.join(HERE, "ref_writers.json"), "w") as fh:
        json.dump(rec, fh, indent=1, sort_keys=True)
    print(json.dumps(rec, indent=1, sort_keys=True))


if __name__ == "__main__":
    main()
