"""Builds libcablegen.so (gfx950 HIP kernels + host C ABI) in-tree, plus the
oracle libraries used by the tests (oracle/liborc.so, and oracle/_ref when the
reference sources are present).

    python -m cablegen_amd.build
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libcablegen.so")
BUILD = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# -ffp-contract=off keeps the FP64 spawn average bit-identical to a
# sequential evaluation; packed-fp32 generation is disabled on this pool.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-Wall", "-Wno-unused-function", "-pthread"]


def _run(cmd, cwd=ROOT):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return out


def build_lib():
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in ("kernels.hip", "sort.hip", "engine.hip"):
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    obj = os.path.join(BUILD, "host.o")
    _run(["gcc"] + C_FLAGS + ["-c", os.path.join(CSRC, "host.c"), "-o", obj])
    objs.append(obj)
    tmp = OUT + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs +
         ["-lpthread"])
    os.replace(tmp, OUT)
    return OUT


NATIVE = os.path.join(ROOT, "tests", "native")
CALLER = os.path.join(NATIVE, "_build", "dropin_caller")


def build_caller():
    """tests/native/dropin_caller: a C caller of the drop-in ABI linked to
    libcablegen.so (GPU tests run it; it needs no reference file)"""
    os.makedirs(os.path.dirname(CALLER), exist_ok=True)
    _run(["gcc", "-O2", "-std=gnu11", "-Wall", "-I", os.path.join(ROOT, "include"), "-o", CALLER,
          os.path.join(NATIVE, "dropin_caller.c"), "-L", HERE, "-lcablegen",
          "-Wl,-rpath,$ORIGIN/../../../cablegen_amd"])
    return CALLER


def build_oracle():
    _run(["make", "-s", "-B", "-f", "oracle/Makefile"])
    if os.path.isdir("/root/reference/src"):
        _run(["make", "-s", "-j8", "-f", "oracle/Makefile.ref", "all"])
        _run(["make", "-s", "-B", "-f", "oracle/Makefile.ref", "cli"])


def build_all():
    build_lib()
    build_caller()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", OUT)
