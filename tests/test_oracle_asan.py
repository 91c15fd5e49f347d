"""The FP64 CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

`make -C oracle asan` compiles oracle/asan_check.c with the oracle's two C sources (no OpenMP)
and runs every orc_* entry point on small inputs allocated at exactly their documented sizes,
including the edge shapes the estimators reach (Nside 1 and 2, mmax below lmax, spin 2, zero
radial weights, empty L ranges).  Any out-of-bounds access, use after free or undefined
arithmetic aborts the run with a nonzero status.
"""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None,
                    reason="needs gcc and make")
def test_oracle_entry_points_are_memory_safe():
    r = subprocess.run(["make", "-s", "-C", ORACLE, "asan"], capture_output=True, text=True,
                       timeout=300, env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1"))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "asan_check ok" in r.stdout
