"""Build and run the reference solver (oracle/_ref) and read its state dumps.

Test infrastructure only: used in this container to pin the oracle and to
write golden fixtures; never imported by the product, and never needed on a
machine without the reference sources.

A run directory is laid out the way the reference expects
(Ark/Cpp/Mesh/MeshConstruct.cpp:74-77 reads $HOME/Mesh/<name>.cas; outputs go
to ../FlowField/{global,Convergence} relative to the working directory,
Ark/Header/ZeroMacros.h:24).  stdin is /dev/null so the reference's
getchar() pauses return at once.
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_ROOT = "/root/reference/Ark"
MESH_NAME = "128_128_Car_Periodic_Square_LBM"  # Ark/Constant/ZeroFlip.h:47-49

# name -> (model translation unit, extra compile flags, walls in y, extra -I dir)
CFGDIR = os.path.join(HERE, "refcfg")
CONFIGS = {
    "hcz": ("D2Q9HCZ", "", False, ""),
    "acp": ("D2Q9AC", "", False, ""),
    "hcz_wall": ("D2Q9HCZ", "-D_Wall_3_BCs_FLIP=3 -D_Wall_3_BCs_BB", True, ""),
    "hcz_nee": ("D2Q9HCZ", "-D_Wall_3_BCs_FLIP=3 -D_Wall_3_BCs_NEE", True, ""),
    "acp_nee": ("D2Q9AC", "-D_Wall_3_BCs_FLIP=3 -D_Wall_3_BCs_NEE", True, ""),
    "acp_wall": ("D2Q9AC", "-D_Wall_3_BCs_FLIP=3 -D_Wall_3_BCs_BB", True, ""),
    "hcz_uw": ("D2Q9HCZ", "-include " + os.path.join(CFGDIR, "uw", "flags.h"), False, ""),
    "acp_uw": ("D2Q9AC", "-include " + os.path.join(CFGDIR, "uw", "flags.h"), False, ""),
    "hcz_uw_wall": ("D2Q9HCZ", "-include " + os.path.join(CFGDIR, "uw", "flags.h") +
                    " -D_Wall_3_BCs_FLIP=3 -D_Wall_3_BCs_BB", True, ""),
    "hcz_uw_nee": ("D2Q9HCZ", "-include " + os.path.join(CFGDIR, "uw", "flags.h") +
                   " -D_Wall_3_BCs_FLIP=3 -D_Wall_3_BCs_NEE", True, ""),
    "acp_uw_nee": ("D2Q9AC", "-include " + os.path.join(CFGDIR, "uw", "flags.h") +
                   " -D_Wall_3_BCs_FLIP=3 -D_Wall_3_BCs_NEE", True, ""),
    "acp_uw_wall": ("D2Q9AC", "-include " + os.path.join(CFGDIR, "uw", "flags.h") +
                    " -D_Wall_3_BCs_FLIP=3 -D_Wall_3_BCs_BB", True, ""),
    "pp": ("D2Q9PP", "-include " + os.path.join(CFGDIR, "pp", "flags.h"), False,
           os.path.join(CFGDIR, "pp")),
    # pseudopotential with bounce-back walls in y (the reference cannot run it: with walls
    # on, CreatMesh.h:11-13 skips CarfaceFacesConstruct, so the face force stencil is null)
    "pp_wall": ("D2Q9PP", "-include " + os.path.join(CFGDIR, "pp", "flags.h") +
                " -D_Wall_3_BCs_FLIP=3 -D_Wall_3_BCs_BB", True, os.path.join(CFGDIR, "pp")),
    # SCMP_Drop shifted by (60, 70) cells: the drop crosses both periodic lines
    "pp_shift": ("D2Q9PP", "-include " + os.path.join(CFGDIR, "pp", "flags.h"), False,
                 os.path.join(CFGDIR, "pp_shift")),
    # pseudopotential with UW faces (SCMP_Drop, the -I dir of "pp" selects it)
    "pp_uw": ("D2Q9PP", "-include " + os.path.join(CFGDIR, "pp_uw", "flags.h"), False,
              os.path.join(CFGDIR, "pp")),
    # SCMP_FlatInterface: a liquid band across the periodic x line
    "pp_flat": ("D2Q9PP", "-include " + os.path.join(CFGDIR, "pp", "flags.h"), False,
                os.path.join(CFGDIR, "pp_flat")),
}

CELL_SCALARS = ["Phi", "Rho", "U", "V", "p", "Phi_x", "Phi_y", "Rho_x", "Rho_y",
                "Fx", "Fy", "laplacianPhi", "prevPhiU", "prevPhiV", "Psi", "scG",
                "Mu", "ftau", "faBP", "fbBP", "fcBP", "haBP", "hbBP", "hcBP", "xc", "yc"]


# Configurations on another mesh: the reference's mesh size is compile-time
# (const Nx, Ny, NL, ChLength, X_End, Y_End in Ark/Constant/ZeroReference.h).
# build() writes a copy of that header with these constants replaced into
# oracle/_ref/<cfg>/inc/ and force-includes it; its include guard then keeps
# the original out.  Rayleigh-Taylor size, bounce-back walls in y, no gravity
# (the reference has none, Ark/Cpp/SourceM.cpp:32-33): the CPU baseline of the
# headline workload.
MESH = {"hcz_rt": (256, 1024)}
CONFIGS["hcz_rt"] = ("D2Q9HCZ", "-D_Wall_3_BCs_FLIP=3 -D_Wall_3_BCs_BB", True, "")


def mesh_of(cfg):
    return MESH.get(cfg, (128, 128))


def _mesh_header(cfg):
    """The reference's ZeroReference.h with the mesh constants of `cfg`
    (unit cells: ChLength = NL = nx, so MinL = 1), written under oracle/_ref."""
    import re
    nx, ny = MESH[cfg]
    with open(os.path.join(REF_ROOT, "Constant", "ZeroReference.h")) as fh:
        text = fh.read()
    subs = [(r"\bNx = \d+", "Nx = %d" % nx), (r"\bNy = \d+", "Ny = %d" % ny), (r"\bNL = \d+", "NL = %d" % nx),
            (r"\bChLength = [0-9.]+", "ChLength = %d.0" % nx), (r"\bX_End = [0-9.]+", "X_End = %d" % nx),
            (r"\bY_End = [0-9.]+", "Y_End = %d.0" % ny)]
    for pat, rep in subs:
        text, n = re.subn(pat, rep, text, count=1)
        assert n == 1, pat
    d = os.path.join(HERE, "_ref", cfg, "inc")
    os.makedirs(d, exist_ok=True)
    path = os.path.join(d, "ZeroReference.h")
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, "w") as fh:
            fh.write(text)
    return path


def available():
    return os.path.isdir(REF_ROOT)


def build(cfg):
    model, defs, _, idir = CONFIGS[cfg]
    if cfg in MESH:
        defs = defs + " -include " + _mesh_header(cfg)
    cmd = ["make", "-s", "-f", os.path.join(HERE, "Makefile.ref"), "-j8",
           "CFG=" + cfg, "MODEL=" + model, "DEFS=" + defs, "IQUOTE=" + ("-I" + idir if idir else "")]
    subprocess.run(cmd, check=True, cwd=os.path.dirname(HERE))
    return os.path.join(HERE, "_ref", cfg, "ark")


def run(cfg, steps, max_step, outdir=None, flowfield=False, timeout=3600, writers=False, cell_order=None):
    """Run config `cfg`, dumping the cell state after each step in `steps`."""
    exe = build(cfg)
    base = os.path.join(HERE, "_ref", cfg, "run")
    home = os.path.join(base, "home")
    work = os.path.join(base, "work")
    for d in (os.path.join(home, "Mesh"), work, os.path.join(base, "FlowField", "global"),
              os.path.join(base, "FlowField", "Convergence")):
        os.makedirs(d, exist_ok=True)
    sys.path.insert(0, HERE)
    from casmesh import write_cas
    write_cas(os.path.join(home, "Mesh", MESH_NAME + ".cas"), *mesh_of(cfg),
              walls_y=CONFIGS[cfg][2], cell_order=cell_order)
    outdir = outdir or os.path.join(base, "dump")
    if writers:  # fresh output tree: the convergence files are appended to
        import shutil
        shutil.rmtree(os.path.join(base, "FlowField"), ignore_errors=True)
        for d in ("global", "Convergence"):
            os.makedirs(os.path.join(base, "FlowField", d), exist_ok=True)
    os.makedirs(outdir, exist_ok=True)
    env = dict(os.environ)
    env.update({"HOME": home, "OMP_NUM_THREADS": "1",
                "MPD_REF_DUMP_STEPS": ",".join(str(s) for s in steps),
                "MPD_REF_MAX_STEP": str(max_step), "MPD_REF_DUMP_DIR": outdir,
                "MPD_REF_FLOWFIELD": "1" if flowfield else "0",
                "MPD_REF_WRITERS": "1" if writers else "0"})
    with open(os.devnull) as dn, open(os.path.join(base, "stdout.log"), "w") as log:
        subprocess.run([exe], cwd=work, env=env, stdin=dn, stdout=log,
                       stderr=subprocess.STDOUT, check=True, timeout=timeout)
    return outdir


def load_cells(path):
    """Return dict of (ny, nx) float64 arrays from a cells_step*.bin dump."""
    with open(path, "rb") as fh:
        nx, ny, nf = np.frombuffer(fh.read(12), dtype=np.int32)
        data = np.frombuffer(fh.read(), dtype=np.float64).reshape(nf, ny, nx)
    out = {}
    k = 0
    nq = 9
    if nf == 2 * nq + len(CELL_SCALARS):
        out["h"] = data[0:nq].copy()
        k = nq
    out["f"] = data[k:k + nq].copy()
    k += nq
    for name in CELL_SCALARS:
        out[name] = data[k].copy()
        k += 1
    return out


if __name__ == "__main__":
    cfg = sys.argv[1]
    steps = [int(s) for s in sys.argv[2].split(",")]
    print(run(cfg, steps, max(steps)))


def time_prebuilt(cfg="hcz", steps=2000, threads=1, timeout=120):
    """Time the already-built reference binary of `cfg` on its own case (no
    rebuild, so it also runs where the reference sources are absent): returns
    [(step, ms for the preceding 1000 steps)] from its own progress lines
    (DUGKS2DSolver.cpp:872-876)."""
    import re
    import tempfile
    exe = os.path.join(HERE, "_ref", cfg, "ark")
    if not os.path.exists(exe):
        return None
    base = tempfile.mkdtemp(prefix="mpd_refrun_")
    home = os.path.join(base, "home")
    work = os.path.join(base, "work")
    for d in (os.path.join(home, "Mesh"), work, os.path.join(base, "FlowField", "global"),
              os.path.join(base, "FlowField", "Convergence")):
        os.makedirs(d, exist_ok=True)
    sys.path.insert(0, HERE)
    from casmesh import write_cas
    write_cas(os.path.join(home, "Mesh", MESH_NAME + ".cas"), *mesh_of(cfg), walls_y=CONFIGS[cfg][2])
    env = dict(os.environ)
    env.update({"HOME": home, "OMP_NUM_THREADS": str(threads), "MPD_REF_DUMP_STEPS": "",
                "MPD_REF_MAX_STEP": str(steps + 1), "MPD_REF_DUMP_DIR": base, "MPD_REF_FLOWFIELD": "0",
                "MPD_REF_WRITERS": "0"})
    with open(os.devnull) as dn:
        r = subprocess.run([exe], cwd=work, env=env, stdin=dn, stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, timeout=timeout)
    import shutil
    shutil.rmtree(base, ignore_errors=True)
    out = []
    for line in r.stdout.decode(errors="replace").split("\n"):
        m = re.match(r"^(\d+)\s+\S+\s+\S+\s+\S+\s+\S+\s+(\d+)\s*$", line)
        if m:
            out.append((int(m.group(1)), int(m.group(2))))
    return out


def time_window(cfg="hcz_rt", first=10, last=110, threads=1, timeout=300):
    """Time steps first..last-1 of the already-built reference binary of `cfg`
    (no rebuild) from the harness's step-start timestamps, excluding mesh
    reading and set-up: returns (steps, seconds) or None."""
    import re
    import shutil
    import tempfile
    exe = os.path.join(HERE, "_ref", cfg, "ark")
    if not os.path.exists(exe):
        return None
    base = tempfile.mkdtemp(prefix="mpd_refrun_")
    try:
        home = os.path.join(base, "home")
        work = os.path.join(base, "work")
        for d in (os.path.join(home, "Mesh"), work, os.path.join(base, "FlowField", "global"),
                  os.path.join(base, "FlowField", "Convergence")):
            os.makedirs(d, exist_ok=True)
        sys.path.insert(0, HERE)
        from casmesh import write_cas
        write_cas(os.path.join(home, "Mesh", MESH_NAME + ".cas"), *mesh_of(cfg), walls_y=CONFIGS[cfg][2])
        env = dict(os.environ)
        env.update({"HOME": home, "OMP_NUM_THREADS": str(threads), "MPD_REF_DUMP_STEPS": "",
                    "MPD_REF_TIME_STEPS": "%d,%d" % (first, last), "MPD_REF_MAX_STEP": str(last),
                    "MPD_REF_DUMP_DIR": base, "MPD_REF_FLOWFIELD": "0", "MPD_REF_WRITERS": "0"})
        with open(os.devnull) as dn:
            r = subprocess.run([exe], cwd=work, env=env, stdin=dn, stdout=subprocess.DEVNULL,
                               stderr=subprocess.PIPE, timeout=timeout)
        t = {int(m.group(1)): float(m.group(2))
             for m in re.finditer(r"\[ref_harness\] step (\d+) t ([0-9.]+)", r.stderr.decode(errors="replace"))}
    finally:
        shutil.rmtree(base, ignore_errors=True)
    if first not in t or last not in t:
        return None
    return last - first, t[last] - t[first]
