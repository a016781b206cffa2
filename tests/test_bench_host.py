"""Host logic of bench.py that runs without a GPU: the wall-clock budget that
skips late stages (so the one JSON line is printed in time) and the refusal
of more than one GPU."""
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402  (imports no GPU library at module level)


def test_budget_skips_what_does_not_fit_and_names_it():
    b = bench.Budget(1e6)
    assert b.fits("small stage", 10.0)
    assert b.skipped == []
    b = bench.Budget(0.0)
    assert not b.fits("late stage", 5.0, reserve_s=0.0)
    assert len(b.skipped) == 1 and b.skipped[0].startswith("late stage")
    # the reserve counts against what is left (the budget runs from the
    # module's import)
    b = bench.Budget(time.monotonic() - bench.T_START + 60.0)
    left = b.left()
    assert 0.0 < left <= 60.0
    assert not b.fits("needs the reserve", left - 5.0, reserve_s=10.0)
    assert b.fits("fits with the reserve", max(0.0, left - 30.0), reserve_s=10.0)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""
