"""bench.py's command line without a GPU: a request for more than one GPU is
refused with a message and exit status 2 before anything touches a device."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.parametrize("argv,env", [(["--gpus", "2"], {}), (["--gpus", "8"], {}), ([], {"WORLD_SIZE": "4"})])
def test_bench_refuses_more_than_one_gpu(argv, env):
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), *argv, "--steps", "1"],
                       cwd=ROOT, env={**os.environ, **env}, capture_output=True, text=True, timeout=120)
    assert r.returncode == 2, (r.returncode, r.stdout, r.stderr)
    assert "measures one GPU only" in r.stdout
    assert "{" not in r.stdout  # no JSON result line
