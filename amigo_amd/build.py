"""Build libamigo_amd.so (host symbolic C++ + gfx950 HIP kernels) in-tree.

    python -m amigo_amd.build            # compile every source, link the .so

Each source is compiled with hipcc for gfx950 only; the object files go to
amigo_amd/_build/ and the library to amigo_amd/libamigo_amd.so.
"""

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libamigo_amd.so")

SOURCES = [
    "symbolic.cpp",
    "ldl_numeric.hip",
    "ipm_kernels.hip",
    "assembly.hip",
    "capi.cpp",
    "driver_kernels.hip",
    "ipm_driver.cpp",
]

# Host decisions of the driver compare values computed with the same
# operations as the reference's Python driver: no contraction of a*b+c.
EXTRA_FLAGS = {"ipm_driver.cpp": ["-ffp-contract=off"]}

ARCH = "gfx950"


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found: the HIP toolchain is required")
    return exe


def common_flags():
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-fPIC",
        "-std=c++17",
        "-Wall",
        "-Wno-unused-function",
        "-Wno-unused-variable",
        "-Wno-unknown-pragmas",
        # Packed FP32 instructions are disabled for every device pass.
        "-Xclang",
        "-target-feature",
        "-Xclang",
        "-packed-fp32-ops",
        "-I" + os.path.join(os.path.dirname(HERE), "include"),
        "-I" + CSRC,
    ]


def compile_one(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    extra = EXTRA_FLAGS.get(src, [])
    cmd = [hipcc()] + common_flags() + extra + ["-c", path, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [hipcc(), "-x", "hip"] + common_flags() + extra + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stdout}\n{r.stderr}")
    # hipcc's host pass warns that it does not know the device-only feature.
    msgs = [
        l
        for l in (r.stderr or "").splitlines()
        if "packed-fp32-ops" not in l and l.strip() and "warning generated" not in l
    ]
    return obj, msgs


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        for obj, msgs in ex.map(compile_one, SOURCES):
            objs.append(obj)
            if verbose and msgs:
                print("\n".join(msgs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
