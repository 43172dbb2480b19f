"""Build step of the gfx950 backend: compiles binder_amd/csrc/*.hip with hipcc
into the in-tree shared library binder_amd/libbinder_hip.so.

Usage: python binder_amd/build.py [-v]
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbinder_hip.so")

ARCH = os.environ.get("BINDER_OFFLOAD_ARCH", "gfx950")
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # keep hipcc from emitting packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
# per-source extra flags: the Jacobian/topology code must not fuse
# multiply-adds so its FP64 results match the CPU expression trees bit for bit
SOURCES = {
    "estep.hip": [],
    "smoothing.hip": [],
    "registration.hip": [],
    "topology.hip": ["-ffp-contract=off"],
    # FP64 resampling: same operation order as the host reference restatement
    "resample.hip": ["-ffp-contract=off"],
    "warmstart.hip": [],
    "autotune.hip": [],
    "models.hip": [],
}


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


def build(verbose=False):
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(OUT_DIR, exist_ok=True)
    cc = hipcc()
    objs, cmds = [], []
    for src, extra in SOURCES.items():
        obj = os.path.join(OUT_DIR, src.replace(".hip", ".o"))
        cmds.append([cc, *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    # the translation units compile independently: one hipcc process each
    jobs = max(1, min(len(cmds), os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(jobs) as pool:
        for cmd, r in zip(cmds, pool.map(
                lambda c: subprocess.run(c, capture_output=True, text=True), cmds)):
            if verbose:
                print(" ".join(cmd), flush=True)
            if r.stderr:
                print(r.stderr, end="", file=sys.stderr, flush=True)
            if r.returncode != 0:
                raise subprocess.CalledProcessError(r.returncode, cmd, r.stdout, r.stderr)
    tmp = LIB + ".tmp"
    cmd = [cc, *COMMON, "-shared", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
