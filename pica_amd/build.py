"""Builds libpica_amd.so (hand-written HIP for gfx950) in-tree with hipcc.

    python -m pica_amd.build        (or __graft_entry__.build())
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libpica_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# (source, extra flags).  The FP64 parity kernels are built without FP contraction so that
# their arithmetic is pica's operation for operation; the FP32 forms may contract.
SOURCES = [
    ("pk_api.hip", ["-ffp-contract=off"]),
    ("pk_field.hip", ["-ffp-contract=off"]),
    ("pk_particles_f64.hip", ["-ffp-contract=off"]),
    ("pk_particles_split.hip", ["-ffp-contract=fast"]),
    ("pk_ensemble.hip", ["-ffp-contract=off"]),
    ("pk_diag.hip", ["-ffp-contract=off"]),
    ("pk_batch.hip", ["-ffp-contract=off"]),
    ("pk_denoise.hip", ["-ffp-contract=fast"]),
    ("pk_project.hip", ["-ffp-contract=off"]),
    ("pk_snapshot.hip", ["-ffp-contract=off"]),
]

COMMON = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-munsafe-fp-atomics",
    # packed FP32 forms stay off: some of their forms have given silent wrong results on MI355X
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _compile(src, flags, objdir):
    obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + flags + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    return obj


def build(verbose=True, jobs=None):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda sf: _compile(sf[0], sf[1], objdir), SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr[-4000:])
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
