"""store_run (the query kernel's packed estimate stores) compiled for the host
and checked byte by byte for every alignment and length, under the address and
undefined-behaviour sanitizers."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_store_run_writes_exactly_the_run(tmp_path):
    exe = tmp_path / "store_run_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    f"-I{ROOT}/scbf_amd/csrc", os.path.join(ROOT, "tests", "cpp", "store_run_check.cpp"),
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "exact" in out.stdout
