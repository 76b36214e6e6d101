"""bench.py contract pieces that need no GPU: more than one GPU is refused with a
message, and roofline.traffic comes from the committed rocprofv3 counter summary
for exactly the workload it was measured on."""
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)


def test_more_than_one_gpu_is_refused():
    res = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120)
    assert res.returncode == 2
    assert "exactly one GPU" in res.stderr and res.stdout == ""


def test_traffic_comes_from_the_counter_summary():
    import bench
    traffic, shapes = bench.pmc_traffic(80, 68)
    assert isinstance(traffic, int) and traffic > 1_000_000_000
    assert len(shapes.split(";")) == 4
    with open(os.path.join(REPO, bench.PMC_SUMMARY)) as f:
        text = f.read()
    assert f"traffic_bytes_per_launch_avg: {traffic}" in text and "rocprofv3 --pmc FETCH_SIZE" in text
    assert bench.pmc_traffic(64, 68) == (None, None) and bench.pmc_traffic(80, 12) == (None, None)
