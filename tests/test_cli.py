"""The rollout CLI (crates/engine-cli/src/main.rs) over libj2k: flag handling
on the CPU, a short run on the GPU."""

import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
CLI = ROOT / "jacobi2000_amd" / "rollout"


def run(*args):
    return subprocess.run([str(CLI), *args], capture_output=True, text=True, timeout=120)


def test_usage_without_model_exits_2():
    r = run("--steps", "3")
    assert r.returncode == 2 and "usage: rollout --model" in r.stderr


def test_unknown_flag_and_case_exit_2(fixture_dir):
    assert run("--model", str(fixture_dir), "--bogus").returncode == 2
    assert run("--model", str(fixture_dir), "--backend", "webgpu").returncode == 2
    assert run("--model", str(fixture_dir), "--backend", "cpu").returncode == 2


@pytest.mark.gpu
def test_rollout_runs_and_writes_npy(fixture_dir, tmp_path):
    out = tmp_path / "h.npy"
    r = run("--model", str(fixture_dir), "--steps", "5", "--case", "drop", "--out", str(out))
    assert r.returncode == 0, r.stderr
    lines = [l for l in r.stdout.splitlines() if l.startswith("step")]
    assert len(lines) == 5 and all("corr=" in l for l in lines)
    h = np.load(out)
    assert h.shape == (16, 16) and np.isfinite(h).all()
