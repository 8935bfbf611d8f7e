"""Builds the engine library heu_amd/libheu_paillier.so for gfx950 with hipcc.

Sources: csrc/ops_{2048,3072}.hip (kernels and launch tables per key size) and
csrc/worklist_{,dec_}{2048,3072}.hip (persistent work-list kernels), all built from
csrc/paillier_device.hpp; csrc/engine.cpp (contexts, keys, one-op entry
points), csrc/batch.cpp (multi-task batches, graphs), csrc/server.cpp (request
server), csrc/tune.cpp (autotuner), csrc/keygen.cpp, csrc/noalloc.hip
(heu_sum_noalloc and its fixed device scratch) and csrc/wire.cpp (HEU's
serialized form of ciphertexts, matrices and keys).
The library links OpenSSL's libcrypto for per-key host constants.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libheu_paillier.so")
ARCH = os.environ.get("HEU_OFFLOAD_ARCH", "gfx950")

SOURCES = ["ops_2048.hip", "ops_3072.hip", "worklist_2048.hip", "worklist_3072.hip", "worklist_dec_2048.hip",
           "worklist_dec_3072.hip", "engine.cpp", "batch.cpp", "server.cpp", "tune.cpp",
           "keygen.cpp", "noalloc.hip", "wire.cpp"]
# Packed FP32 instructions stay off in device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the engine needs ROCm's hipcc to build")
    return path


def _compile(src: str, obj: str) -> None:
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-pass-failed",
           "-Wno-unused-function", "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".hip"):
        cmd[1:1] = DEVICE_FLAGS
    res = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in res.stderr.splitlines() if "is not a recognized feature" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{err}")


def build(force: bool = True) -> str:
    """Compiles every source and links the shared library; returns its path."""
    newest = max(os.path.getmtime(os.path.join(CSRC, f)) for f in os.listdir(CSRC))
    if not force and os.path.exists(LIB) and os.path.getmtime(LIB) >= newest:
        return LIB
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    objs = [os.path.join(obj_dir, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        list(ex.map(lambda a: _compile(*a), zip(SOURCES, objs)))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lcrypto"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stderr}")
    os.replace(tmp, LIB)
    build_spi()
    return LIB


SPI_LIB = os.path.join(HERE, "libheu_paillier_spi.so")
SPI_TEST = os.path.join(os.path.dirname(HERE), "tests", "cpp", "build", "paillier_hip_test")
SPAN_BENCH = os.path.join(HERE, "spi", "build", "span_bench")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def build_spi() -> str:
    """The C++ vectorized-SPI classes (host code over the C ABI), their test
    binary and the span-rate program; all host-only, so g++ builds them."""
    include = os.path.join(os.path.dirname(HERE), "include")
    spi = os.path.join(HERE, "spi")
    common = ["g++", "-std=c++20", "-O2", "-fPIC", "-Wall", "-pthread", f"-I{include}", f"-I{spi}",
              f"-I{ROCM}/include", "-D__HIP_PLATFORM_AMD__"]
    cmd = [*common, "-shared", "-o", SPI_LIB + ".tmp", os.path.join(spi, "paillier_hip.cc"), f"-L{HERE}",
           "-lheu_paillier", f"-L{ROCM}/lib", "-lamdhip64", "-lcrypto", "-Wl,-rpath,$ORIGIN",
           f"-Wl,-rpath,{ROCM}/lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"SPI build failed: {' '.join(cmd)}\n{res.stderr}")
    os.replace(SPI_LIB + ".tmp", SPI_LIB)
    os.makedirs(os.path.dirname(SPI_TEST), exist_ok=True)
    src = os.path.join(os.path.dirname(os.path.dirname(SPI_TEST)), "paillier_hip_test.cc")
    cmd = [*common, "-o", SPI_TEST + ".tmp", src, f"-L{HERE}", "-lheu_paillier_spi", "-lheu_paillier",
           f"-L{ROCM}/lib", "-lamdhip64", "-lcrypto", "-Wl,-rpath,$ORIGIN/../../../heu_amd", f"-Wl,-rpath,{ROCM}/lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"SPI test build failed: {' '.join(cmd)}\n{res.stderr}")
    os.replace(SPI_TEST + ".tmp", SPI_TEST)
    # the span-rate program bench.py runs (SPI with host spans beside device-resident calls)
    os.makedirs(os.path.dirname(SPAN_BENCH), exist_ok=True)
    cmd = [*common, "-o", SPAN_BENCH + ".tmp", os.path.join(spi, "span_bench.cc"), f"-L{HERE}", "-lheu_paillier_spi",
           "-lheu_paillier", f"-L{ROCM}/lib", "-lamdhip64", "-lcrypto", "-Wl,-rpath,$ORIGIN/../..",
           f"-Wl,-rpath,{ROCM}/lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"span bench build failed: {' '.join(cmd)}\n{res.stderr}")
    os.replace(SPAN_BENCH + ".tmp", SPAN_BENCH)
    return SPI_LIB


if __name__ == "__main__":
    print(build())
