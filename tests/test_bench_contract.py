"""bench.py on the host: it measures one GPU (any other --gpus is refused
before a device is touched), and the committed rocprofv3 summaries it prices
the sentence kernel from parse at the launch shape the autotuner picks on the
2.2M-vertex corpus, with commands that run committed files only."""

import importlib.util
import os
import re
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROFILES = os.path.join(REPO, "profiles")


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(REPO, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "--gpus must be 1" in r.stderr
    assert r.stdout == ""


def test_profile_parses_at_the_tuned_shape():
    bench = load_bench()
    tuned = {"minibatch": 524288, "block_waves": 4, "pool_split": 16}
    prof = bench.measured_profile(tuned, 128)
    assert prof is not None
    # HBM bytes per kept position: below the algorithmic ~10 KB (hub rows stay
    # in L2), far above the ~2 KB of the output rows' reads alone
    assert 2000 < prof["traffic_per_kept"] < 20000
    assert 0 < prof["mfma_util_pct"] < 100
    # the autotuner's other pick on this corpus (two wavefronts per workgroup)
    other = bench.measured_profile({"minibatch": 524288, "block_waves": 2, "pool_split": 16}, 128)
    assert other is not None and abs(other["traffic_per_kept"] - prof["traffic_per_kept"]) < 0.05 * prof["traffic_per_kept"]
    # another launch shape or dim has no measured traffic: the bench reports null
    assert bench.measured_profile({"minibatch": 131072, "block_waves": 2, "pool_split": 16}, 128) is None
    assert bench.measured_profile(tuned, 256) is None


def test_profile_commands_run_committed_files():
    names = sorted(f for f in os.listdir(PROFILES) if f.endswith((".md", ".txt")))
    assert names
    for name in names:
        text = open(os.path.join(PROFILES, name)).read()
        assert "rocprofv3" in text, name
        for rel in re.findall(r"\$R/([\w./-]+)", text):
            assert os.path.isfile(os.path.join(REPO, rel)), (name, rel)
