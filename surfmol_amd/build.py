"""Build libsurfmol_relax.so (HIP for gfx950 + host C++) in-tree with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsurfmol_relax.so")
SOURCES = ["relax_kernels.hip", "energy_kernels.hip", "order_kernels.hip", "gridff_kernels.hip", "server_kernels.hip", "layer_kernels.hip", "raff_kernels.hip", "plain_noalloc.hip", "raff.cpp", "raff_server.cpp", "host_raff.cpp", "capi.cpp", "launch_plan.cpp", "host_topology.cpp", "server.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SMR_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 (v_pk_*_f32) is kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wall", "-Wno-unused-result", "-Wno-unused-variable",
    "-I", os.path.join(HERE, "..", "include"),
    "-pthread",
]


def build(verbose: bool = True) -> str:
    """Compile every HIP/C++ source into one shared library (always rebuilds)."""
    tmp = OUT + ".tmp"
    objs = []
    jobs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, "_build", os.path.splitext(src)[0] + ".o")
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        cmd = [HIPCC] + [f for f in FLAGS if f != "-shared"] + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = cmd[:1] + ["-x", "hip"] + cmd[1:]
        jobs.append((src, cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = False
    for src, cmd, p in jobs:
        out, _ = p.communicate()
        txt = out.decode(errors="replace")
        # hipcc's host pass does not know the gfx feature flag: harmless warning
        lines = [l for l in txt.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l]
        if p.returncode != 0:
            failed = True
            sys.stderr.write(f"--- {src} failed ---\n" + "\n".join(lines) + "\n")
        elif verbose and any(lines):
            sys.stderr.write(f"--- {src} ---\n" + "\n".join(lines) + "\n")
    if failed:
        raise RuntimeError("hipcc compilation failed")
    link = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread"] + objs + ["-o", tmp]
    subprocess.run(link, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())
