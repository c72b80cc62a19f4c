"""Write the golden fixtures of the reference solver (run where /root/reference
is present; the fixtures then travel with the repository).

For each configuration the compiled reference (oracle/Makefile.ref, unmodified
sources + state-dump hook) is run from its own initialiser and the full cell
state after steps 0, 1, 2, 10 and 100 is recorded as SHA-256 digests of the
float64 planes in the library's upload layout (tilde_h, tilde_f, each macro
field), plus a few raw values for diagnostics.  A bit-exact restatement must
reproduce every digest.

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from mpdugks_amd.abi import MACRO_FIELDS  # noqa: E402

REF_KEYS = {"phi": "Phi", "rho": "Rho", "u": "U", "v": "V", "p": "p", "phix": "Phi_x",
            "phiy": "Phi_y", "rhox": "Rho_x", "rhoy": "Rho_y", "fx": "Fx", "fy": "Fy",
            "lap": "laplacianPhi", "ppu": "prevPhiU", "ppv": "prevPhiV", "psi": "Psi",
            "scg": "scG", "mu": "Mu", "tau": "ftau"}
STEPS = [0, 1, 2, 10, 100]
# name -> (model, steps recorded)
CONFIGS = {"hcz": ("ac_hcz", STEPS), "acp": ("ac_p", STEPS), "hcz_wall": ("ac_hcz", STEPS),
           "acp_wall": ("ac_p", STEPS),
           # walls with non-equilibrium extrapolation (Wall_3_NEE)
           "hcz_nee": ("ac_hcz", STEPS), "acp_nee": ("ac_p", STEPS),
           # UW face scheme, periodic and with walls (the reference keeps two copies of
           # each periodic face; the oracle follows both, tests/test_oracle_ref_live.py)
           "hcz_uw": ("ac_hcz", STEPS), "acp_uw": ("ac_p", STEPS), "hcz_uw_wall": ("ac_hcz", STEPS),
           "hcz_uw_nee": ("ac_hcz", STEPS), "acp_uw_nee": ("ac_p", STEPS), "acp_uw_wall": ("ac_p", STEPS),
           # pseudopotential: the face force depends on the face visiting order; the
           # reference runs with one thread, which the oracle's reference-order mode follows
           "pp": ("pp_d2q9", STEPS),
           # the same drop shifted cyclically by (60, 70) cells: it crosses both periodic lines
           "pp_shift": ("pp_d2q9", STEPS),
           # SCMP_FlatInterface: a liquid band across the periodic x line
           "pp_flat": ("pp_d2q9", STEPS),
           # SCMP_Drop with UW faces
           "pp_uw": ("pp_d2q9", STEPS)}
SHIFT = {"pp_shift": [60, 70]}  # (x, y) cells, oracle/refcfg/pp_shift/ZeroCondition.h


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def state_digests(d):
    n = d["f"].shape[1] * d["f"].shape[2]
    out = {"f": digest(d["f"].reshape(9, n))}
    if "h" in d:
        out["h"] = digest(d["h"].reshape(9, n))
    out["m"] = {k: digest(d[REF_KEYS[k]].reshape(n)) for k in MACRO_FIELDS}
    return out


def main(names=None):
    from ref_run import load_cells, run
    for cfg, (model, steps) in CONFIGS.items():
        if names and cfg not in names:
            continue
        out = run(cfg, steps, max(steps))
        rec = {"model": model, "walls_y": cfg.endswith(("_wall", "_nee")), "uw": "_uw" in cfg,
               "bc": "nee" if cfg.endswith("_nee") else ("bb" if cfg.endswith("_wall") else "periodic"),
               "nx": 128, "ny": 128,
               "source": "reference Ark, unmodified sources, OMP_NUM_THREADS=1, %s"
                         % ("SCMP_FlatInterface" if cfg == "pp_flat" else
                            "SCMP_Drop" if cfg.startswith("pp") else "AC_Drop"),
               "steps": {}}
        if cfg in SHIFT:
            rec["shift"] = SHIFT[cfg]
        if cfg == "pp_flat":
            rec["start"] = "flat"
        for k in steps:
            d = load_cells(os.path.join(out, "cells_step%d.bin" % k))
            r = state_digests(d)
            r["sample"] = {name: [float(x) for x in d[REF_KEYS[name]][64, 10:30]]
                           for name in ("phi", "rho", "p", "u", "v")}
            rec["steps"][str(k)] = r
        path = os.path.join(HERE, "ref_%s.json" % cfg)
        with open(path, "w") as fh:
            json.dump(rec, fh, indent=1)
        print("wrote", path)


if __name__ == "__main__":
    main(sys.argv[1:] or None)
