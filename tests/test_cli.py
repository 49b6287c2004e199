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
    assert run("--model", str(fixture_dir), "--backend", "cpu").returncode == 2
    r = run("--model", str(fixture_dir), "--backend", "opengl")
    assert r.returncode == 2 and "unrecognized --backend" in r.stderr


def test_webgpu_backend_name_is_the_gpu_engine(fixture_dir):
    """The reference's GPU backend name selects the HIP engine: it is not
    refused as unknown (without a GPU it exits 1, as the reference does when
    no adapter is granted)."""
    r = run("--model", str(fixture_dir), "--backend", "webgpu", "--steps", "1")
    assert r.returncode != 2 and "unrecognized" not in r.stderr


@pytest.mark.gpu
def test_webgpu_backend_runs_on_the_gpu(fixture_dir):
    r = run("--model", str(fixture_dir), "--backend", "webgpu", "--steps", "3")
    assert r.returncode == 0, r.stderr
    assert len([l for l in r.stdout.splitlines() if l.startswith("step")]) == 3


@pytest.mark.gpu
def test_rollout_runs_and_writes_npy(fixture_dir, tmp_path):
    out = tmp_path / "h.npy"
    r = run("--model", str(fixture_dir), "--steps", "5", "--case", "drop", "--out", str(out))
    assert r.returncode == 0, r.stderr
    lines = [l for l in r.stdout.splitlines() if l.startswith("step")]
    assert len(lines) == 5 and all("corr=" in l for l in lines)
    h = np.load(out)
    assert h.shape == (16, 16) and np.isfinite(h).all()


@pytest.mark.gpu
def test_rollout_prints_the_device_loops_trace(fixture_dir):
    """The printed correlations are j2k_rollout_truth's for the `drop` case
    (main.rs:63-78: centre drop, amplitude 0.05, sigma 4)."""
    import jacobi2000_amd as j
    from jacobi2000_amd import checkpoint as ck
    from jacobi2000_amd import solvers as sw

    r = run("--model", str(fixture_dir), "--steps", "4", "--case", "drop")
    assert r.returncode == 0, r.stderr
    printed = [float(l.split("corr=")[1]) for l in r.stdout.splitlines() if l.startswith("step")]
    m = j.Model.load(fixture_dir)
    c = m.cfg
    p = sw.Params(nx=c.nx, ny=c.ny, dx=c.dx, dt=c.dt, g=c.g, depth=c.depth, nu=c.nu)
    st = sw.State.flat(p)
    st.add_drop(p, c.nx / 2.0, c.ny / 2.0, 0.05, 4.0)
    n = c.nx * c.ny
    x0 = np.concatenate([(st.h - np.float32(c.norm_depth) - np.float32(c.norm_mean[0])) / np.float32(c.norm_std[0]),
                         (st.u - np.float32(c.norm_mean[1])) / np.float32(c.norm_std[1]),
                         (st.v - np.float32(c.norm_mean[2])) / np.float32(c.norm_std[2])]).astype(np.float32)
    _, _, _, _, corr = m.rollout_truth(sw.ShallowWater(p), x0, st.h, st.u, st.v,
                                       ck.meta_of(c.dx, c.dt, c.g, c.depth, c.nu), 4)
    assert len(printed) == 4 and x0.size == 3 * n
    assert all(abs(a - float(b)) <= 6e-6 for a, b in zip(printed, corr)), (printed, corr)
