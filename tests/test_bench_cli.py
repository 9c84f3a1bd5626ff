"""bench.py's command line on the CPU: more than one GPU is refused with an error line and
exit status 2 (the engine runs on one GPU; sharding across GPUs is designed, not built)."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], cwd=ROOT,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    line = [x for x in r.stdout.splitlines() if x.startswith("{")][-1]
    assert "one GPU" in json.loads(line)["error"]
