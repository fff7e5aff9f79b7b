"""The lane kernel's four-byte-group bit readers, compiled for the host,
against the oracle's byte-wise MEL / MagSgn / VLC readers (readers_check.cpp)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_group_readers_match_bytewise_readers(tmp_path):
    exe = tmp_path / "readers_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-w", "-I", os.path.join(ROOT, "oracle"),
                    "-o", str(exe), os.path.join(HERE, "readers_check.cpp")], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout + out.stderr
