"""bench.py's host logic on the CPU: the one-GPU guard and the HBM traffic it reads from the
committed rocprofv3 summary of the intersection workload."""
import re
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode != 0
    assert "one GPU" in r.stdout + r.stderr


def test_traffic_from_the_committed_profile():
    """Per query: FETCH_SIZE KiB x the copy factor + WRITE_SIZE KiB x its factor, in bytes; the dominant
    query's figure equals the one the summary states in GB."""
    text = Path(bench.TRAFFIC_PROFILE).read_text()
    ff = float(re.search(r"FETCH_SIZE \(factor ([\d.]+)\)", text).group(1))
    wf = float(re.search(r"WRITE_SIZE \(factor ([\d.]+)\)", text).group(1))
    for q in ("coherent_closest", "incoherent_closest", "coherent_any", "incoherent_any"):
        label = q.replace("_", " ")
        m = re.search(r"^\| " + label + r" \| ([\d,]+) \| [\d.]+ \| ([\d,]+) ->", text, flags=re.M)
        f, w = (int(m.group(k).replace(",", "")) for k in (1, 2))
        assert bench.traffic_from_profile(bench.TRAFFIC_PROFILE, q) == (f * ff + w * wf) * 1024
    stated = float(re.search(r"Dominant kernel traffic: ([\d.]+) GB", text).group(1))
    got = bench.traffic_from_profile(bench.TRAFFIC_PROFILE, "incoherent_closest") / 1e9
    assert abs(got - stated) < 0.01
    assert bench.traffic_from_profile(str(ROOT / "no_such_profile.md"), "incoherent_closest") is None
