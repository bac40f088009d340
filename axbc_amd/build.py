"""In-tree build of the engine's native libraries (gfx950).

    python -m axbc_amd.build

produces
    axbc_amd/lib/libaxbc_hip.so    HIP kernels + C ABI (include/axbc.h)
    axbc_amd/lib/libaxbc_host.so   graph tooling (C, OpenMP)
    axbc_amd/bin/axbc2d            command-line driver with AxBC's flags
and the checker builds under oracle/ (the CPU oracle always; the reference
generator under oracle/_ref only when the reference sources are present).

    python -m axbc_amd.build --sanitize

builds the host-side C (graph tooling, driver, legacy layer) and the CPU
oracle with AddressSanitizer and UndefinedBehaviorSanitizer into _build_san/;
the CPU test suite runs against them with

    AXBC_SAN_DIR=$PWD/_build_san LD_PRELOAD=$(gcc -print-file-name=libasan.so) \
        ASAN_OPTIONS=detect_leaks=0 UBSAN_OPTIONS=halt_on_error=1:print_stacktrace=1 \
        python -m pytest tests -m "not gpu"

(the HIP library stays the ordinary build: the CPU suite only loads it).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "axbc_amd")
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")
HIP_SRC = os.path.join(PKG, "csrc", "hip")
HOST_SRC = os.path.join(PKG, "csrc", "host")
INCLUDE = os.path.join(ROOT, "include")
REFERENCE = "/root/reference"

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
HIP_SOURCES = ["engine.hip", "batch.hip", "wgsweep.hip", "sampler.hip", "mlp.hip", "ingest.hip", "kronecker.hip",
               "noalloc.hip"]
HOST_SOURCES = ["kronecker.c", "graph.c"]


def _run(cmd: list[str], cwd: str = ROOT) -> None:
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")


def _hip_objects() -> list[str]:
    return [os.path.join(HIP_SRC, s) for s in HIP_SOURCES if os.path.exists(os.path.join(HIP_SRC, s))]


def build_hip() -> str:
    out = os.path.join(LIB, "libaxbc_hip.so")
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    objs = []
    with ThreadPoolExecutor(max_workers=4) as ex:
        futs = []
        for src in _hip_objects():
            obj = os.path.join(LIB, os.path.basename(src) + ".o")
            objs.append(obj)
            futs.append(ex.submit(_run, [hipcc, *HIP_FLAGS, "-I", INCLUDE, "-c", src, "-o", obj]))
        for f in futs:
            f.result()
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    for o in objs:
        os.remove(o)
    return out


def build_host() -> str:
    out = os.path.join(LIB, "libaxbc_host.so")
    srcs = [os.path.join(HOST_SRC, s) for s in HOST_SOURCES]
    _run(["gcc", "-O3", "-std=gnu11", "-fopenmp", "-fPIC", "-shared", "-Wall", "-Wextra",
          "-Wno-unused-parameter", *srcs, "-o", out + ".tmp", "-lm", "-lpthread"])
    os.replace(out + ".tmp", out)
    return out


def build_driver() -> str | None:
    src = os.path.join(HOST_SRC, "axbc2d.c")
    if not os.path.exists(src):
        return None
    out = os.path.join(BIN, "axbc2d")
    _run(["gcc", "-O3", "-std=gnu11", "-fopenmp", "-Wall", "-Wextra", "-Wno-unused-parameter",
          "-I", INCLUDE, "-I", HOST_SRC, src, "-o", out + ".tmp",
          "-L", LIB, "-laxbc_host", "-laxbc_hip", "-Wl,-rpath,$ORIGIN/../lib", "-lm"])
    os.replace(out + ".tmp", out)
    return out


def build_legacy() -> str:
    """libaxbc_legacy.so: the reference's device entry points; its globals
    N and heuristic resolve against the host program that loads it."""
    out = os.path.join(LIB, "libaxbc_legacy.so")
    _run(["gcc", "-O2", "-std=gnu11", "-fPIC", "-shared", "-Wall", "-I", INCLUDE,
          os.path.join(HOST_SRC, "legacy.c"), "-o", out + ".tmp", "-L", LIB, "-laxbc_hip",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(out + ".tmp", out)
    return out


def build_legacy_test_host() -> str:
    """Test host of the legacy layer (tests/legacy), built in-tree so the GPU
    tests only run it."""
    src = os.path.join(ROOT, "tests", "legacy", "legacy_host.c")
    if not os.path.exists(src):
        return ""
    bindir = os.path.join(ROOT, "tests", "legacy", "bin")
    os.makedirs(bindir, exist_ok=True)
    out = os.path.join(bindir, "legacy_host")
    _run(["gcc", "-O2", "-std=gnu11", "-Wall", "-I", INCLUDE, src, "-o", out + ".tmp", "-L", LIB,
          "-laxbc_legacy", "-laxbc_hip", "-Wl,-rpath," + os.path.relpath(LIB, bindir).join(["$ORIGIN/", ""])])
    os.replace(out + ".tmp", out)
    return out


def build_oracle() -> None:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    if os.path.isdir(os.path.join(REFERENCE, "generator")):
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", f"REF={REFERENCE}"])


SAN = os.path.join(ROOT, "_build_san")
SAN_FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
             "-fno-sanitize-recover=undefined"]


def build_sanitized() -> str:
    """Host libraries, driver and oracle with ASan + UBSan, into _build_san/."""
    os.makedirs(SAN, exist_ok=True)
    srcs = [os.path.join(HOST_SRC, s) for s in HOST_SOURCES]
    _run(["gcc", *SAN_FLAGS, "-std=gnu11", "-fopenmp", "-fPIC", "-shared", *srcs,
          "-o", os.path.join(SAN, "libaxbc_host.so"), "-lm", "-lpthread"])
    _run(["gcc", *SAN_FLAGS, "-std=gnu11", "-fopenmp", "-fPIC", "-shared",
          os.path.join(ROOT, "oracle", "brandes_oracle.c"), "-o", os.path.join(SAN, "liboracle.so"), "-lm"])
    _run(["gcc", *SAN_FLAGS, "-std=gnu11", "-fopenmp", "-I", INCLUDE, "-I", HOST_SRC,
          os.path.join(HOST_SRC, "axbc2d.c"), "-o", os.path.join(SAN, "axbc2d"),
          "-L", SAN, "-L", LIB, "-laxbc_host", "-laxbc_hip", f"-Wl,-rpath,{SAN}:{LIB}", "-lm"])
    _run(["gcc", *SAN_FLAGS, "-std=gnu11", "-fPIC", "-shared", "-I", INCLUDE, os.path.join(HOST_SRC, "legacy.c"),
          "-o", os.path.join(SAN, "libaxbc_legacy.so"), "-L", LIB, "-laxbc_hip", f"-Wl,-rpath,{LIB}"])
    return SAN


def build_all() -> None:
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    with ThreadPoolExecutor(max_workers=3) as ex:
        fh = ex.submit(build_hip)
        fo = ex.submit(build_host)
        fr = ex.submit(build_oracle)
        fh.result()
        fo.result()
        fr.result()
    build_driver()
    build_legacy()
    build_legacy_test_host()


if __name__ == "__main__":
    if "--sanitize" in sys.argv[1:]:
        print("built", build_sanitized())
    else:
        build_all()
        print("built", LIB, BIN)
