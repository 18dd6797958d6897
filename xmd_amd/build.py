"""Compile the HIP engine for gfx950: xmd_amd/libxmdhip.so.

    python -m xmd_amd.build

Two translation units.  engine.hip (lists, QA, integrator, noise, C ABI) is
compiled with -ffp-contract=off: its FP64 list/QA decision code and the
deterministic Box-Muller math evaluate in source order, as on the host.
forces.hip (the force kernel) is compiled with -ffp-contract=fast, so the
force arithmetic may use fused multiply-adds.  Packed FP32 instructions
are disabled in both.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libxmdhip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# forces.hip fuses multiply-adds except where a block turns contraction off
# with a pragma (the exact state tests of the wall pairs)
UNITS = [("engine.hip", ["-ffp-contract=off"]), ("forces.hip", ["-ffp-contract=fast-honor-pragmas"])]


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    procs, objs = [], []
    for src, extra in UNITS:
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        cmd = [HIPCC, *FLAGS, *extra, "-c", "-o", obj, os.path.join(CSRC, src)]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        objs.append(obj)
    for cmd, pr in procs:
        out, err = pr.communicate()
        if pr.returncode != 0:
            sys.stderr.write(out + err)
            raise RuntimeError("hipcc failed: " + " ".join(cmd))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc link failed: " + " ".join(cmd))
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
