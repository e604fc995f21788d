"""Build recipe of the native libraries (in-tree, no JIT cache).

    python -m naer_amd.build            # libnaer.so (gfx950) + the C++ driver
    python -m naer_amd.build --oracle   # also the CPU oracle (test infrastructure)

Outputs:
    naer_amd/lib/libnaer.so   HIP kernels + C ABI (include/naer/naer.h)
    naer_amd/bin/naer         C++ driver with the reference's CLI (src/main.cpp)
"""
import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "naer_amd")
CSRC = os.path.join(PKG, "csrc")
LIB_DIR = os.path.join(PKG, "lib")
BIN_DIR = os.path.join(PKG, "bin")
OBJ_DIR = os.path.join(PKG, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off: every fp32/fp64 operation of the step is rounded on its
# own, in the order written, so the CPU oracle can reproduce it bit for bit.
# -packed-fp32-ops off: no v_pk_*_f32 (known to misbehave on these boxes).
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wall", "-Wno-unused-result",
]

HIP_SOURCES = ["kernels.hip", "naer_runtime.hip", "naer_model.hip", "kernels_event.hip"]
CXX_DRIVER = ["Network.cpp", "main.cpp"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")
    return r.stdout


def _compile_hip(src):
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
    _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build_lib():
    os.makedirs(OBJ_DIR, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(HIP_SOURCES)) as ex:
        objs = list(ex.map(_compile_hip, HIP_SOURCES))
    out = os.path.join(LIB_DIR, "libnaer.so")
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lhiprtc"])
    os.replace(tmp, out)
    return out


def build_driver():
    srcs = [os.path.join(CSRC, s) for s in CXX_DRIVER if os.path.exists(os.path.join(CSRC, s))]
    if len(srcs) != len(CXX_DRIVER):
        return None
    os.makedirs(BIN_DIR, exist_ok=True)
    out = os.path.join(BIN_DIR, "naer")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "include", "naer"), *srcs,
          "-L", LIB_DIR, "-lnaer", "-Wl,-rpath,$ORIGIN/../lib", "-o", out])
    return out


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle.so")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--oracle", action="store_true")
    ap.add_argument("--clean", action="store_true")
    a = ap.parse_args(argv)
    if a.clean:
        shutil.rmtree(OBJ_DIR, ignore_errors=True)
    print(build_lib())
    d = build_driver()
    if d:
        print(d)
    if a.oracle:
        print(build_oracle())


if __name__ == "__main__":
    main()
