"""The bench record's inputs that need no GPU: the element kernel's measured HBM bytes per
launch come from the committed rocprofv3 summaries (FETCH_SIZE + WRITE_SIZE, scaled by a copy
kernel), so a summary that loses its "HBM bytes per launch" line would turn the record's
roofline.traffic into null; and a request for more than one GPU is refused before any GPU
work."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_traffic_from_committed_profile(precision):
    t = bench.traffic_from_profile(precision)
    assert isinstance(t, int)
    # between the algorithmic bytes of one launch (128,000 elements: records written plus
    # per-IP coefficients read) and three times that
    rec = 304 * 27 * (8 if precision == "fp64" else 4)
    alg = 128000 * (rec + 70 * 27 * 8)
    assert alg * 0.9 < t < 3 * alg, (t, alg)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "one GPU" in (r.stdout + r.stderr)
