"""CPU check of the fused H_eff plan's T1 / T3 layout (planner.cpp: heff_plan,
step1_layout, mpo_layout): tests/heff_layout_check.cpp is compiled with the
planner's source by g++ and run on seeded random two-site bonds, padded and
unpadded (MPSH_HEFF_PADR=0).  It asserts that the MPO pass reads only T1
elements step 1 writes, that every T3 element step 4 reads (padding included)
is written by the MPO pass, the aligned unit-stride rows the pass's vector runs
need in padded plans, and that every MPO work list covers each group's runs
exactly once."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_heff_layout(tmp_path):
    exe = str(tmp_path / "heff_layout_check")
    subprocess.run(["g++", "-std=c++17", "-O1", os.path.join(ROOT, "tests", "heff_layout_check.cpp"),
                    os.path.join(ROOT, "mps_amd", "csrc", "planner.cpp"), "-o", exe], check=True)
    env = {k: v for k, v in os.environ.items() if k != "MPSH_HEFF_PADR"}
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    lines = r.stdout.strip().splitlines()
    assert lines[-1] == "OK 24 cases"
    assert sum(" padded," in ln for ln in lines) == 12
