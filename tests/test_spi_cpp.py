"""The C++ vectorized-SPI layer (heu_amd/spi/paillier_hip.{h,cc}).

CPU: the SPI library loads (its engine and HIP runtime dependencies resolve)
and the C++ test binary is built.  GPU: the C++ suite, which restates
paillier_zahlen/paillier_test.cc:26-233 and checks the span forms against
closed forms, passes; its seeded key equals the oracle's key for seed 0.
"""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SPI_LIB = os.path.join(ROOT, "heu_amd", "libheu_paillier_spi.so")
SPI_TEST = os.path.join(ROOT, "tests", "cpp", "build", "paillier_hip_test")


def test_spi_library_loads():
    assert os.path.exists(SPI_LIB), "build with `python -m heu_amd.build`"
    ctypes.CDLL(SPI_LIB)
    assert os.access(SPI_TEST, os.X_OK)


@pytest.mark.gpu
def test_spi_cpp_suite(oracle_mod):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    print("running C++ SPI suite", flush=True)
    res = subprocess.run([SPI_TEST], capture_output=True, text=True, timeout=300)
    print(res.stdout[-4000:], res.stderr[-2000:], flush=True)
    assert res.returncode == 0, res.stdout[-4000:] + res.stderr[-2000:]
    assert "ALL PASSED" in res.stdout
    m = re.search(r"seeded n = ([0-9A-F]+)", res.stdout)
    assert m and int(m.group(1), 16) == oracle_mod.keygen(2048, 0)["n"]
