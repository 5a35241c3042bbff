"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer (float-to-integer overflow
included): oracle/sanitize_check.c drives every entry point over odd image sizes, crops at
the borders, supersampling, GPS, spatial reuse and temporal modes; any report fails."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ORACLE = Path(__file__).resolve().parents[1] / "oracle"


@pytest.mark.skipif(shutil.which("gcc") is None and shutil.which("cc") is None, reason="no C compiler")
def test_oracle_is_sanitizer_clean():
    r = subprocess.run(["make", "-s", "-C", str(ORACLE), "sanitize"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "cases clean" in r.stdout
