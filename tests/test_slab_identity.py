"""The traversal's slab test against the oracle on the CPU.

bg::dev::slab_entry (bonsai_amd/csrc/dev_geom.h) combines the reference's six
plane distances with max/min instead of the reference's sequence of early exits
and updates (bounds3.rs:246-288). tests/slab_identity.cpp compiles that header
for the host and compares "entry < t_max" with the oracle's line-by-line
restatement (oracle/rt.h, intersect_range_precomputed) on 8 million boxes and
rays built from special values (+-0, +-inf, denormals, origins on planes, flat
boxes, zero direction components), for several t_max per case including the
entry distance itself and the next float above it. Planting any of these
mistakes in a copy of the header makes it fail: dropping the x-axis NaN check,
< for <=, >= 0 for > 0, a missing x padding, a dropped z entry."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(shutil.which(HIPCC) is None and not Path(HIPCC).exists(), reason="hipcc not available")
def test_slab_entry_equals_oracle_slab_test(tmp_path):
    exe = tmp_path / "slab_identity"
    subprocess.run([HIPCC, "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", "-ffp-contract=off",
                    "-fno-fast-math", f"-I{ROOT / 'include'}", "-o", str(exe), str(ROOT / "tests" / "slab_identity.cpp")],
                   check=True, capture_output=True, timeout=300)
    r = subprocess.run([str(exe), "8000000"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    last = r.stdout.strip().splitlines()[-1].split()
    assert last[-1] == "0" and int(last[3]) > 100_000, r.stdout
