"""Compiles libmpm_hip.so for gfx950 in-tree with hipcc.

Flags: -ffp-contract=off keeps every FP32 operation a separate IEEE rounding (the kernels
restate the reference arithmetic term by term); correctly rounded FP32 divide/sqrt; packed
FP32 instructions disabled.
"""
import os
import pathlib
import shutil
import subprocess
import sys

PKG_DIR = pathlib.Path(__file__).resolve().parent
SRC = PKG_DIR / "csrc" / "mpm_hip.hip"
OUT = PKG_DIR / "libmpm_hip.so"

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-value", "-Wno-unused-result",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and pathlib.Path(c).exists():
            return c
    raise FileNotFoundError("hipcc not found")


def build(verbose=False):
    deps = [SRC] + sorted((PKG_DIR / "csrc").glob("*.h")) + [PKG_DIR.parent / "include" / "mpm_hip.h"]
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc()] + HIP_FLAGS + ["-I", str(PKG_DIR.parent / "include"), "-o", str(tmp), str(SRC)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for libmpm_hip.so")
    if verbose:
        sys.stderr.write("\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l))
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
