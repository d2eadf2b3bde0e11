"""The upload's host encoders (louvaingpu_amd/csrc/engine/encode.cpp) on the
CPU: tests/encode_check.cpp is compiled with it by g++ and run.  It checks the
row offsets of sorted sources against counting (parts cut 1 to 17 ways; skewed
runs, empty rows, one huge row, sparse sources; corrupted lists refused), the
weight packing at 2, 4 and 8 bits (AVX2 against scalar, decoded back to
min(w, 2^b - 1), escapes and their cap), and the width choice."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ENC = os.path.join(os.path.dirname(HERE), "louvaingpu_amd", "csrc", "engine", "encode.cpp")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_encoders(tmp_path):
    exe = str(tmp_path / "encode_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "encode_check.cpp"),
                    ENC], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("(0 failures)")
