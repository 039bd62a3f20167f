"""CPU check of the stream-K step-4 work division (planner.cpp: use_streamk,
streamk_schedule) and of the DMA kernel's per-tile records (tile_records):
tests/streamk_schedule_check.cpp is compiled with the planner's source by g++
and run; it asserts chunk coverage, minimum piece length, the partial-tile
bound, the per-work-group balance, and every record field of stream-K and
split-K tile lists against the plan (first chunk found by enumeration)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_streamk_schedule(tmp_path):
    exe = str(tmp_path / "streamk_check")
    subprocess.run(["g++", "-std=c++17", "-O1", os.path.join(ROOT, "tests", "streamk_schedule_check.cpp"),
                    os.path.join(ROOT, "mps_amd", "csrc", "planner.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert r.stdout.strip().splitlines()[-1].startswith("OK")
