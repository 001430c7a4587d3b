"""Compile the gfx950 library and the data generator in-tree.

libprzm_gpu.so    HIP kernels + host runtime behind include/przm_gpu.h
libprzm_datagen.so seeded synthetic inputs (host C)
libprzm_decode.so  CPU decoder of the frames (host C, include/przm_decode.h)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build failed: " + " ".join(cmd))
    return out


def build_gpu():
    out = os.path.join(HERE, "libprzm_gpu.so")
    _run([HIPCC] + HIP_FLAGS + [os.path.join(CSRC, "przm_runtime.hip"), "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_datagen():
    out = os.path.join(HERE, "libprzm_datagen.so")
    _run(["gcc", "-O2", "-std=c11", "-D_GNU_SOURCE", "-fPIC", "-shared", "-pthread",
          os.path.join(CSRC, "przm_datagen.c"), "-lm", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_decoder():
    out = os.path.join(HERE, "libprzm_decode.so")
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-Wall", "-Wextra",
          os.path.join(CSRC, "przm_decode.c"), "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build():
    return [build_gpu(), build_datagen(), build_decoder()]


if __name__ == "__main__":
    for p in build():
        print(p)
