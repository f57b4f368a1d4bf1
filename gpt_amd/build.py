"""Builds the HIP library of gpt_amd for gfx950 (MI355X) in-tree.

``python -m gpt_amd.build`` or ``gpt_amd.build.build()`` compiles
``csrc/*.hip`` with hipcc into ``gpt_amd/libgpt_amd_wc.so``.  The object
files go to ``gpt_amd/_build/`` (git-ignored) and are rebuilt every call.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libgpt_amd_wc.so")
SOURCES = ["wc_kernels.hip", "wc_hop_f64.hip", "wc_hop_f32.hip", "wc_api.hip", "wc_batch.hip", "wc_block.hip",
           "wc_halo.hip", "wc_probe.hip", "wc_noalloc.hip", "wc_alloc.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # keep the compiler from emitting packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # fused multiply-adds only where the source writes fma(): same roundings in every instantiation
    "-ffp-contract=off",
    "-Wall", "-Wno-unused-result",
]


def build(verbose=False):
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(bdir, src.replace(".hip", ".o"))
        cmd = [HIPCC] + FLAGS + ["-c", os.path.join(SRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT), src))
        objs.append(obj)
    failed = False
    for p, src in procs:
        out = p.communicate()[0].decode()
        lines = [l for l in out.splitlines()
                 if "packed-fp32-ops" not in l and "unknown target feature" not in l.lower()]
        if p.returncode != 0 or (verbose and lines):
            sys.stderr.write("\n".join(lines) + "\n")
        if p.returncode != 0:
            failed = True
    if failed:
        raise RuntimeError("hipcc failed building gpt_amd")
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
