"""bench.py's command-line contract without a GPU: more than one GPU (by
flag or by a torchrun world size) exits with status 2 and a message, before
anything touches the device."""
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(args, env_extra=None):
    env = dict(os.environ)
    env.update(env_extra or {})
    return subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), *args], cwd=REPO, env=env,
                          capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("n", [2, 8])
def test_more_than_one_gpu_exits_with_message(n):
    r = run_bench(["--gpus", str(n), "--steps", "1", "--warmup", "0"])
    assert r.returncode == 2
    assert "one GPU" in r.stderr and f"--gpus {n}" in r.stderr
    assert r.stdout.strip() == ""  # no JSON line


def test_world_size_above_one_exits():
    r = run_bench(["--steps", "1"], {"WORLD_SIZE": "2", "RANK": "0", "LOCAL_RANK": "0"})
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_help_lists_contract_flags():
    r = run_bench(["--help"])
    assert r.returncode == 0
    for flag in ("--gpus", "--steps", "--warmup"):
        assert flag in r.stdout
