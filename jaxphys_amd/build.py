"""Build the HIP libraries of jaxphys_amd in-tree for gfx950.

``python -m jaxphys_amd.build`` (or ``build()``) compiles every ``.hip``
library listed in ``LIBS`` with hipcc into ``jaxphys_amd/``.  It always
recompiles; the loaders only use what is already built.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ARCH = os.environ.get("JAXPHYS_AMD_ARCH", "gfx950")

# -ffp-contract=off keeps the float32 rounding sequence of the reference's
# expressions; packed fp32 ops are disabled for correctness on this target.
HIPFLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-mcode-object-version=5",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]

LIBS = {
    "libyee_hip.so": ["csrc/yee.hip", "csrc/yee_fused.hip", "csrc/yee_adjoint.hip"],
    "libconv3d_hip.so": ["csrc/conv3d.hip"],
    "libtm2d_hip.so": ["csrc/tm2d.hip"],
}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False, only=None) -> list:
    built = []
    for lib, srcs in LIBS.items():
        if only and lib not in only:
            continue
        out = os.path.join(HERE, lib)
        tmp = out + ".tmp"
        cmd = [hipcc(), *HIPFLAGS, *[os.path.join(HERE, s) for s in srcs],
               "-I", os.path.join(ROOT, "include"), "-o", tmp]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"hipcc failed for {lib}")
        if verbose and res.stderr.strip():
            sys.stderr.write(res.stderr)
        os.replace(tmp, out)
        built.append(out)
    return built


if __name__ == "__main__":
    for p in build(verbose=True):
        print("built", p)
