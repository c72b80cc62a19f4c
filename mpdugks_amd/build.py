"""Compile the HIP sources of the step library for gfx950 into
mpdugks_amd/_lib/libmpdugks_hip.so (in-tree, so it travels with the tree)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT, "libmpdugks_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MPD_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I" + CSRC, "-Wall", "-Wno-unused-function"]
# per-source floating-point contraction: the phase path and the host ABI keep
# the reference's operation order exactly; the fused path may contract.
SOURCES = {
    "capi.hip": ["-ffp-contract=off"],
    "phase_kernels.hip": ["-ffp-contract=off"],
    "fused_kernels.hip": ["-ffp-contract=off"],
    "fused_fp32.hip": ["-ffp-contract=fast"],
    # no contraction: the fused multiply-adds are written out, so every inlined
    # copy of a face computes it alike (a face two tiles share, bit for bit)
    "pp32.hip": ["-ffp-contract=off"],
    "devinit.hip": ["-ffp-contract=off"],
    "hostinit.cpp": ["-ffp-contract=off"],
    "writers.cpp": ["-ffp-contract=off"],
}


def build(verbose=False, jobs=4):
    os.makedirs(OUT, exist_ok=True)
    objs, procs = [], []
    for src, extra in SOURCES.items():
        obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
        flags = COMMON if src.endswith(".hip") else [f for f in COMMON if not f.startswith("--offload")
                                                     and "packed-fp32" not in f and f not in ("-Xclang", "-target-feature")]
        cmd = [HIPCC] + flags + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def _wait(item):
    src, p = item
    out, _ = p.communicate()
    if p.returncode != 0:
        sys.stderr.write(out.decode(errors="replace"))
        raise RuntimeError("hipcc failed on " + src)
    text = out.decode(errors="replace").strip()
    if text:
        sys.stderr.write(text + "\n")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
