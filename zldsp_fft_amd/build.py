"""Build recipe for libzlfft.so (hipcc, gfx950).

Compiles every kernel unit in parallel and links ``zldsp_fft_amd/libzlfft.so``
in-tree, plus the drop-in C++ example/test program.  Run as
``python -m zldsp_fft_amd.build`` or through ``__graft_entry__.build()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libzlfft.so")
MAX_LDS_ORDER = 14
MIN_MACRO_LOGL, MAX_MACRO_LOGL = 6, 10

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fvisibility=hidden",  # only the zlfft_* C-ABI (include/zlfft.h) is exported
    "-ffp-contract=off",
    # packed FP32 (v_pk_*_f32) is kept off for every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", INCLUDE,
    "-Wall",
]


def _run(cmd: list[str]) -> tuple[list[str], int, str]:
    p = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns that it does not know the packed-fp32 feature
    err = "\n".join(l for l in (p.stdout + p.stderr).splitlines() if "packed-fp32-ops" not in l)
    return cmd, p.returncode, err


# The host layer (C-ABI, plans, host pipeline, ragged planning, tuning-file
# parser) again with AddressSanitizer and UndefinedBehaviorSanitizer on the
# host side only (each -fsanitize after -Xarch_host: the device code is the
# normal build's), linked with the normal kernel objects into
# _build/asan/libzlfft.so, and a CPU driver that walks its error paths
# (tests/test_sanitize_cpu.py).
ASAN_DIR = os.path.join(OBJ, "asan")
ASAN_LIB = os.path.join(ASAN_DIR, "libzlfft.so")
ASAN_UNITS = ("zlfft_api.hip", "zlfft_ragged_api.hip")
SAN_FLAGS = ["-O1", "-g", "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-omit-frame-pointer"]
DRIVER_SRC = os.path.join(ROOT, "tests", "sanitize", "host_abi_driver.cpp")
DRIVER_BIN = os.path.join(ASAN_DIR, "host_abi_driver")


def _clang_rt_dir() -> str:
    base = "/opt/rocm/llvm/lib/clang"
    for v in sorted(os.listdir(base)) if os.path.isdir(base) else []:
        d = os.path.join(base, v, "lib", "linux")
        if os.path.exists(os.path.join(d, "libclang_rt.asan-x86_64.so")):
            return d
    raise RuntimeError("the clang AddressSanitizer runtime is not installed")


def link_sanitized() -> str:
    """Links the sanitized host layer and the driver (objects compiled by build())."""
    objs = [obj for _, obj, _ in _units() if os.path.basename(obj) not in ("api.o", "ragged_api.o")]
    objs += [os.path.join(ASAN_DIR, u.replace("zlfft_", "").replace(".hip", ".o")) for u in ASAN_UNITS]
    vmap = os.path.join(OBJ, "exports.map")
    _, rc, err = _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Xarch_host", "-fsanitize=address",
                       "-Xarch_host", "-fsanitize=undefined", "-Xarch_host", "-shared-libasan",
                       f"-Wl,--version-script={vmap}", "-o", ASAN_LIB, *objs])
    if rc:
        print(err, file=sys.stderr)
        raise RuntimeError("link of the sanitized libzlfft.so failed")
    clangxx = "/opt/rocm/llvm/bin/clang++"
    rt = _clang_rt_dir()
    _, rc, err = _run([clangxx, "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-shared-libasan",
                       "-fno-omit-frame-pointer", "-I", INCLUDE, DRIVER_SRC, "-o", DRIVER_BIN, "-L", ASAN_DIR, "-lzlfft",
                       f"-Wl,-rpath,{ASAN_DIR}", f"-Wl,-rpath,{rt}"])
    if rc:
        print(err, file=sys.stderr)
        raise RuntimeError("sanitizer driver build failed")
    return DRIVER_BIN


def _units() -> list[tuple[str, str, list[str]]]:
    """(source, object, extra flags) of every compilation unit of libzlfft.so."""
    units = []
    for n in range(MAX_LDS_ORDER + 1):
        units.append((os.path.join(CSRC, "zlfft_order.hip"), os.path.join(OBJ, f"order{n}.o"), [f"-DZL_LOGN={n}"]))
    for l in range(MIN_MACRO_LOGL, MAX_MACRO_LOGL + 1):
        units.append((os.path.join(CSRC, "zlfft_macro.hip"), os.path.join(OBJ, f"macro{l}.o"), [f"-DZL_LOGL={l}"]))
    units.append((os.path.join(CSRC, "zlfft_mfma.hip"), os.path.join(OBJ, "mfma.o"), []))
    units.append((os.path.join(CSRC, "zlfft_ragged_k.hip"), os.path.join(OBJ, "ragged_k.o"), []))
    units.append((os.path.join(CSRC, "zlfft_ragged_api.hip"), os.path.join(OBJ, "ragged_api.o"), []))
    units.append((os.path.join(CSRC, "zlfft_api.hip"), os.path.join(OBJ, "api.o"), []))
    return units


def build(jobs: int | None = None, verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(ASAN_DIR, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    units = _units()
    cmds = [[HIPCC, *HIP_FLAGS, *extra, "-c", src, "-o", obj] for src, obj, extra in units]
    for u in ASAN_UNITS:
        cmds.append([HIPCC, *HIP_FLAGS, *SAN_FLAGS, "-c", os.path.join(CSRC, u),
                     "-o", os.path.join(ASAN_DIR, u.replace("zlfft_", "").replace(".hip", ".o"))])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for cmd, rc, err in ex.map(_run, cmds):
            if verbose or rc:
                print(" ".join(cmd), file=sys.stderr)
                if err.strip():
                    print(err, file=sys.stderr)
            if rc:
                raise RuntimeError(f"hipcc failed ({rc}) for {cmd[-3]}")
    objs = [obj for _, obj, _ in units]
    # export the C-ABI only: the kernels' host-side handles stay local too
    vmap = os.path.join(OBJ, "exports.map")
    with open(vmap, "w") as f:
        f.write("{ global: zlfft_*; local: *; };\n")
    tmp = LIB + ".tmp"
    cmd, rc, err = _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", f"-Wl,--version-script={vmap}", "-o", tmp,
                         *objs])
    if rc:
        print(err, file=sys.stderr)
        raise RuntimeError("link of libzlfft.so failed")
    os.replace(tmp, LIB)
    build_probe()
    build_dropin_example()
    link_sanitized()
    return LIB


PROBE_SRC = os.path.join(CSRC, "zlfft_probe.hip")
PROBE_LIB = os.path.join(HERE, "libzlprobe.so")


def build_probe() -> str:
    """libzlprobe.so: the device-copy reference kernels bench.py times beside
    each line (not part of the transform library)."""
    tmp = PROBE_LIB + ".tmp"
    vmap = os.path.join(OBJ, "probe_exports.map")
    with open(vmap, "w") as f:
        f.write("{ global: zlprobe_*; local: *; };\n")
    _, rc, err = _run([HIPCC, *HIP_FLAGS, "-shared", f"-Wl,--version-script={vmap}", PROBE_SRC, "-o", tmp])
    if rc:
        print(err, file=sys.stderr)
        raise RuntimeError("build of libzlprobe.so failed")
    os.replace(tmp, PROBE_LIB)
    return PROBE_LIB


DROPIN_SRC = os.path.join(ROOT, "tests", "dropin", "dropin_example.cpp")
DROPIN_BIN = os.path.join(ROOT, "tests", "dropin", "_build", "dropin_example")


def build_dropin_example() -> str:
    """README-style C++ program against the drop-in headers, linked to libzlfft.so."""
    os.makedirs(os.path.dirname(DROPIN_BIN), exist_ok=True)
    cxx = shutil.which("g++") or "c++"
    cmd = [cxx, "-std=c++20", "-O2", "-Wall", "-Wextra", "-I", INCLUDE, DROPIN_SRC, "-o", DROPIN_BIN,
           "-L", HERE, "-lzlfft", "-Wl,-rpath,$ORIGIN/../../../zldsp_fft_amd"]
    _, rc, err = _run(cmd)
    if rc:
        print(err, file=sys.stderr)
        raise RuntimeError("drop-in example build failed")
    return DROPIN_BIN


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
