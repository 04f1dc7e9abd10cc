"""Builds and runs the C++ consumer of include/frost_amd.hpp: host-only
checks on the CPU, and a device search under the gpu marker."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "frost_amd" / "_lib" / "consumer"


def build_consumer():
    lib = ROOT / "frost_amd" / "_lib"
    subprocess.run(["g++", "-std=c++20", "-O1", f"-I{ROOT / 'include'}",
                    str(ROOT / "tests" / "cpp" / "consumer.cpp"), "-o", str(EXE),
                    f"-L{lib}", "-lfrost_amd", f"-Wl,-rpath,{lib}"], check=True)


def test_consumer_host():
    build_consumer()
    r = subprocess.run([str(EXE)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "consumer ok" in r.stdout


@pytest.mark.gpu
def test_consumer_gpu():
    if not EXE.exists():
        build_consumer()
    r = subprocess.run([str(EXE), "gpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
