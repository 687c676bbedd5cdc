"""The autotuner's table and its cache file (OCIP_AUTOTUNE_CACHE) never make a call run a
variant it cannot run.

A first child process tunes three calls and writes their keys to the cache file.  The file is
then rewritten so that every entry names the strip kernel: the 128-wide BF16 template cannot
run it (valid only up to 97 columns; before the key held the template width, a 64-wide
template's tuning could be reused for it), split BF16 cannot (its ring holds one operand
plane), and the 64-wide BF16 template can.  A second child runs the same calls from that
file: each map must be within the MFMA bound of the FP64 oracle, and the two calls whose
entries they cannot run must have tuned again (new entries appended).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as orc
from tests import autotune_child as child
from tests.gpu_util import say

pytestmark = pytest.mark.gpu

UNIT = {1: 2.0 ** -7 + 2.0 ** -16, 3: 2.0 ** -14}


def run_child(cache, out):
    env = dict(os.environ, OCIP_AUTOTUNE_CACHE=str(cache))
    env.pop("OCIP_MFMA_VARIANT", None)
    subprocess.run([sys.executable, "-u", child.__file__, str(out)], env=env, check=True,
                   timeout=120)
    with np.load(out) as z:
        return {k: z[k] for k in z.files}


def test_cache_entries_that_a_call_cannot_run_are_tuned_again(tmp_path):
    cache = tmp_path / "tune.txt"
    say("autotune child 1")
    run_child(cache, tmp_path / "a.npz")
    lines = [ln.split() for ln in cache.read_text().splitlines() if ln.strip()]
    single = [k for k, _ in lines if k.startswith("single:")]
    assert len(single) == len(child.CASES)
    bad = [f"{key} 24" for key in single]      # the strip kernel
    cache.write_text("\n".join(bad) + "\n")
    say("autotune child 2")
    maps = run_child(cache, tmp_path / "b.npz")
    src, tpls = child.inputs()
    for prec, tw, th in child.CASES:
        tpl = tpls[(tw, th)]
        f64 = orc.prox_f64(orc.CC, src, tpl)
        sabs = orc.prox_f64(orc.CC, np.abs(src).astype(np.float64),
                            np.abs(tpl).astype(np.float64))
        b = (UNIT[prec] + 2 * (tpl.size + 16) * 2.0 ** -24) * sabs + 1e-30
        err = np.abs(maps[f"p{prec}_{tw}x{th}"].astype(np.float64) - f64)
        assert np.all(err <= b), (prec, tw, th, float(np.max(err / b)))
    # the impossible entries were replaced: the 128-wide and split-BF16 calls tuned again
    after = [ln.split() for ln in cache.read_text().splitlines() if ln.strip()]
    retuned = {k for k, v in after[len(bad):]}
    for key in single:
        if ":p3:" in key or ":tw128:" in key:
            assert key in retuned, key
