"""The benchmark's command line contract, checked without a GPU: one GPU only, and the
headline defaults (256 Poisson fits as 8 folds x 32 variances, 20 timed and 5 warm-up
one-cycle steps)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("n", [2, 8])
def test_more_than_one_gpu_is_refused_with_a_message(n):
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode != 0
    assert "one GPU" in p.stderr
    assert p.stdout == ""   # no JSON line for a run that did not happen


def _bench():
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_cli", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    return bench


def test_headline_defaults():
    bench = _bench()
    a = bench.parser().parse_args([])
    assert (a.gpus, a.steps, a.warmup, a.workload) == (1, 20, 5, "poisson")
    assert (a.batch, a.folds, a.scale) == (256, 8, 1.0)
    # the counter-measured traffic figure is keyed by the headline's own shape
    # (fits, folds, fits per team, workgroups per team)
    assert any(k[:2] == (a.batch, a.folds) for k in bench.TRAFFIC_BYTES_PER_CYCLE)
