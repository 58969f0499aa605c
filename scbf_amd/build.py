"""Builds scbf_amd/libscb.so (HIP kernels + C ABI) for gfx950 with hipcc.

Run as a script or through build(); the library is written in-tree so that it
travels with the repository snapshot to the GPU host.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libscb.so")
ROOT = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ROCM = os.path.dirname(os.path.dirname(os.path.realpath(HIPCC)))
ARCH = os.environ.get("SCB_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
    f"-I{os.path.join(ROOT, 'include')}",
]
# Host-only sources that call the HIP runtime API.
HOST = ["-x", "c++", "-D__HIP_PLATFORM_AMD__", f"-I{os.path.join(ROCM, 'include')}"]
# Packed-FP32 instructions are disabled for all device code.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build(verbose=False):
    objs = []
    tmpdir = os.path.join(HERE, "_build")
    os.makedirs(tmpdir, exist_ok=True)
    for src, dev in (("scb_kernels.hip", True), ("scb_api.cpp", False), ("scb_serve.cpp", False),
                     ("scb_fastx.cpp", False)):
        obj = os.path.join(tmpdir, src + ".o")
        cmd = [HIPCC] + COMMON + (DEVICE if dev else HOST) + ["-c", os.path.join(CSRC, src), "-o", obj]
        out = _run(cmd)
        if verbose and out:
            print(out)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lz", "-ldl"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
