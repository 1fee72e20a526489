"""The solver's bitboard primitives (gpuothello2_amd/csrc/bitboard.hpp),
built for the host with hipcc and checked against the oracle's ray walks and
against each other (tests/bitboard_check.cpp).  CPU only: the device build
of the same code is checked by every GPU score test."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_bitboard_primitives_host(tmp_path):
    lib = os.path.join(ROOT, "oracle")
    assert os.path.exists(os.path.join(lib, "liboracle.so")), "build the oracle first (make -C oracle)"
    out = str(tmp_path / "bitboard_check")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950",
                    "-I", os.path.join(ROOT, "gpuothello2_amd", "csrc"),
                    os.path.join(HERE, "bitboard_check.cpp"), "-L", lib, "-loracle",
                    f"-Wl,-rpath,{lib}", "-o", out], check=True, capture_output=True)
    r = subprocess.run([out], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
