"""Build libpawpy_amd.so (HIP, gfx950) in-tree.

    python -m pawpyseed_amd.build

Every .hip source under pawpyseed_amd/csrc is compiled with hipcc for
gfx950 and linked into pawpyseed_amd/libpawpy_amd.so.  Packed-FP32 VALU
instructions are disabled (-packed-fp32-ops); the host pass warns that it
does not know that feature, which is harmless.
"""

from __future__ import annotations

import glob
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpawpy_amd.so")
ARCH = os.environ.get("PAWPY_AMD_ARCH", "gfx950")

FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


def build(verbose: bool = False, jobs: int = 8) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = sources()
    with tempfile.TemporaryDirectory(prefix="pawpy_amd_build_") as tmp:
        objs, procs = [], []
        for s in srcs:
            o = os.path.join(tmp, os.path.basename(s) + ".o")
            objs.append(o)
            cmd = [hipcc, *FLAGS, "-c", s, "-o", o]
            procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
            if len(procs) >= jobs:
                _drain(procs, verbose)
        _drain(procs, verbose)
        tmp_lib = LIB + ".tmp"
        cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp_lib]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n" + r.stdout)
        os.replace(tmp_lib, LIB)
    return LIB


def _drain(procs, verbose):
    while procs:
        cmd, p = procs.pop(0)
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        lines = [ln for ln in text.splitlines() if "packed-fp32-ops" not in ln]
        if p.returncode != 0:
            raise RuntimeError("compile failed: " + " ".join(cmd) + "\n" + "\n".join(lines))
        if verbose and lines:
            print("\n".join(lines))


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
