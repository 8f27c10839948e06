"""CPU checks of bench.py's refusals: they exit with status 2 and one JSON
line before any device is touched."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("args,needle", [
    (["--gpus", "2"], "one GPU"),
    (["--workload", "hubbard8x8_slater", "--gram", "factored"], "factored"),
    (["--workload", "hubbard4x4_fpeps", "--sampler", "learned"], "learned"),
])
def test_bench_refuses(args, needle):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2, r.stderr[-2000:]
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert needle in line["error"]
