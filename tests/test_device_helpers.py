"""Device helpers with bit-level tricks, compiled for the host and checked
exhaustively under the address and undefined-behaviour sanitizers:
store_run (the query kernel's packed estimate stores) byte by byte for every
alignment and length, pack4_bases (the scan's four-at-a-time base encoding)
against the per-base rule for every byte value at every position, and the
rolling canonical s-mer / k-mer streams of the bin kernels against a direct
evaluation for every k <= 32 and s <= k, and the short form of the counter
address hash against the full SplitMix64 output function."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run_check(tmp_path, name):
    exe = tmp_path / name
    subprocess.run(["g++", "-std=c++17", "-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    f"-I{ROOT}/scbf_amd/csrc", os.path.join(ROOT, "tests", "cpp", name + ".cpp"),
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "exact" in out.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_store_run_writes_exactly_the_run(tmp_path):
    _run_check(tmp_path, "store_run_check")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_pack4_bases_matches_the_per_base_rule(tmp_path):
    _run_check(tmp_path, "pack4_check")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_smer_and_kmer_streams_match_direct_evaluation(tmp_path):
    _run_check(tmp_path, "smer_stream_check")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_address_hash_short_form_matches_splitmix64(tmp_path):
    _run_check(tmp_path, "hash_check")
