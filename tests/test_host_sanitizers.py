"""CPU: the host code (deck and lattice parsers, host loaders, the oracle's
run loop) built with the address and undefined-behaviour sanitizers and run
on the example deck, malformed decks and lattices and small runs of both
field solvers.  No GPU code is built or run."""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "genesis4_cuda_port_amd" / "csrc"
CLANG = Path("/opt/rocm/lib/llvm/bin/clang++")
EXAMPLE = ROOT / "tests" / "golden" / "reference_example"


@pytest.mark.skipif(not CLANG.exists(), reason="ROCm clang++ not installed")
def test_host_code_under_asan_and_ubsan(tmp_path):
    exe = tmp_path / "host_sanitize_main"
    cmd = [str(CLANG), "-x", "c++", "-O1", "-g", "-std=c++17", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include", f"-I{CSRC}", f"-I{ROOT / 'oracle'}", "-fopenmp",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer",
           str(ROOT / "tests" / "host_sanitize_main.cpp"), str(CSRC / "deck.cpp"),
           str(CSRC / "frontend.cpp"), str(ROOT / "oracle" / "g4oracle.cpp"), "-o", str(exe)]
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, res.stderr
    # the OpenMP runtime keeps one allocation of its own until exit
    supp = tmp_path / "lsan.supp"
    supp.write_text("leak:libomp.so\n")
    env = dict(os.environ, OMP_NUM_THREADS="4", LSAN_OPTIONS=f"suppressions={supp}",
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=0")
    res = subprocess.run([str(exe), str(EXAMPLE / "fel.in"), str(EXAMPLE / "fel.lat")],
                         capture_output=True, text=True, timeout=600, env=env)
    out = res.stdout + res.stderr
    assert res.returncode == 0, out[-4000:]
    assert "done" in res.stdout
    assert "runtime error" not in out and "AddressSanitizer" not in out
