"""The certificate's per-pair split sequence (cert_tuv_math.h): the packed
running-ratio builder must equal the level-by-level replay of the split rule
on every input, including ties, powers of two, subnormals, inf and NaN.
Compiled for the host with hipcc and run on the CPU."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not Path(HIPCC).exists(), reason="hipcc is not installed")
def test_split_sequence_matches_replay(tmp_path):
    exe = tmp_path / "split_sequence_check"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", f"-I{ROOT}/p2c-ccd_amd/csrc", f"-I{ROOT}/include",
                    str(ROOT / "tests/native/split_sequence_check.hip"), "-o", str(exe)], check=True,
                   stderr=subprocess.DEVNULL)
    out = subprocess.run([str(exe)], check=False, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "0 mismatches" in out.stdout
