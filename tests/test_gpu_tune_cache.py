"""The autotuner's decisions persist in a file (DH_TUNE_CACHE).

A first process that runs an untuned FP16 pass times the wave layouts and
appends its choice as "<device tag> <key> <choice>"; a second process with the
same file finds the choice and appends nothing, and computes the same Y (the
layouts differ only in how waves are arranged).  DH_TUNE_CACHE set to the
empty string keeps no file.  Each process runs one small pass, one after the
other."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = str(Path(__file__).resolve().parents[1])

BODY = r"""
import sys
sys.path.insert(0, sys.argv[1])
import numpy as np
import torch
import darray_optimizer_amd as dh
n, d, l = 8192, 128, 272
X = torch.from_numpy(dh.fill_mixture(0, d, n, 4, 0.3)).cuda()
Om = dh.to_dev(np.asfortranarray(dh.fill_normal(1, n, l)))
Y = torch.empty((l, n), dtype=torch.float64, device="cuda")
h = dh.Handle(0)
h.rbf_project_f16(X, X, Om, 1.0 / d, Y)
torch.cuda.synchronize()
print("checksum %.17g" % float(Y.double().abs().sum()), flush=True)
h.close()
"""


def run(env_cache, home=None):
    env = dict(os.environ)
    env.pop("DH_TUNE", None)
    env.pop("DH_SCHEDULE", None)
    env["DH_TUNE_CACHE"] = env_cache
    if home is not None:  # where the default path would be
        env["HOME"] = str(home)
        env.pop("XDG_CACHE_HOME", None)
    print(f"subprocess with DH_TUNE_CACHE={env_cache!r}", flush=True)
    r = subprocess.run([sys.executable, "-c", BODY, ROOT], env=env, capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    return [ln for ln in r.stdout.splitlines() if ln.startswith("checksum")][-1]


def test_choice_written_once_then_reused(tmp_path):
    cache = tmp_path / "sub" / "tune.txt"
    first = run(str(cache))
    lines = [ln for ln in cache.read_text().splitlines() if ln and not ln.startswith("#")]
    print("cache after the first process:", lines, flush=True)
    assert len(lines) >= 1
    for ln in lines:
        tag, key, choice = ln.split()
        assert tag.startswith("gfx") and int(key, 16) > 0 and int(choice) >= 0
    second = run(str(cache))
    lines2 = [ln for ln in cache.read_text().splitlines() if ln and not ln.startswith("#")]
    assert lines2 == lines  # found on file: nothing re-tuned, nothing appended
    assert second == first


def test_empty_path_keeps_no_file(tmp_path):
    run("", home=tmp_path)  # other libraries may create their own caches under HOME
    assert not (tmp_path / ".cache" / "darray_hip").exists()


def test_default_path_under_home(tmp_path):
    first = run_default(tmp_path)
    assert (tmp_path / ".cache" / "darray_hip" / "tune.txt").is_file()
    assert first.startswith("checksum")


def run_default(home):
    env = dict(os.environ)
    for k in ("DH_TUNE", "DH_SCHEDULE", "DH_TUNE_CACHE", "XDG_CACHE_HOME"):
        env.pop(k, None)
    env["HOME"] = str(home)
    r = subprocess.run([sys.executable, "-c", BODY, ROOT], env=env, capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    return [ln for ln in r.stdout.splitlines() if ln.startswith("checksum")][-1]
