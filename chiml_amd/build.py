"""Builds the in-tree shared libraries.

libchiml_amd.so : the product (HIP kernels for gfx950 + host manager + C ABI)
oracle/liboracle.so : the FP64 oracle used by the tests (test infrastructure)

Invoked by __graft_entry__.build(); it always recompiles.
"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "chiml_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed-FP32 instructions stay off (a flag for the device compiler only).
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, cwd=ROOT):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=cwd)


def build_product(jobs=None):
    out = os.path.join(ROOT, "chiml_amd", "libchiml_amd.so")
    objdir = os.path.join(ROOT, "chiml_amd", "_build")
    os.makedirs(objdir, exist_ok=True)
    dev_obj = os.path.join(objdir, "device.o")
    _run([HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", *DEVICE_FLAGS,
          "-c", os.path.join(CSRC, "device.hip"), "-o", dev_obj])
    host_objs = []
    for src in ("manager.cpp", "setup.cpp", "server.cpp", "writers.cpp"):
        obj = os.path.join(objdir, src.replace(".cpp", ".o"))
        _run([HIPCC, "-O2", "-std=c++17", "-fPIC", "-pthread", "-x", "c++", "-D__HIP_PLATFORM_AMD__",
              "-c", os.path.join(CSRC, src), "-o", obj])
        host_objs.append(obj)
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", out, dev_obj, *host_objs])
    return out


def build_oracle():
    _run(["make", "-B", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle.so")


def main():
    build_product()
    build_oracle()


if __name__ == "__main__":
    sys.exit(main())
