"""Builds libaxisym_hip.so (gfx950) in-tree with hipcc.

    python -m pyaxisymflow_amd.csrc.build        (or __graft_entry__.build())

Each .hip source compiles to an object in parallel, then one shared library
is linked next to the package.  Packed-FP32 instructions are disabled and
FMA contraction is off by default (kernels that restate the reference keep
its rounding; the transform kernels opt back in locally).
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libaxisym_hip.so")
ARCH = os.environ.get("AXS_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
    "-Wno-unused-result",
]


def sources():
    return sorted(f for f in os.listdir(HERE) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(HERE, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, r.stdout + r.stderr


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    errors = []
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        for src, obj, rc, out in ex.map(_compile, sources()):
            # the host pass does not know the packed-fp32 feature: harmless
            lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l]
            if rc != 0:
                errors.append(f"--- {src}\n" + "\n".join(lines))
            elif verbose and lines:
                print(f"--- {src}\n" + "\n".join(lines))
            objs.append(obj)
    if errors:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errors))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
