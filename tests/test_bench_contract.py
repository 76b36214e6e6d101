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


def test_counter_rows_are_attributed_per_level(tmp_path):
    """The live traffic leg's CSV parsing: the plan run's own sweeps are
    skipped, the last n sweeps of each level are averaged, a resident level's
    single n-iteration launch is divided by n, and the last copy calibrates."""
    import bench
    n = 3
    lines = ["Dispatch_Id,Kernel_Name,Counter_Name,Counter_Value"]
    d = 0

    def add(name, value, counter="FETCH_SIZE"):
        nonlocal d
        d += 1
        lines.append(f'{d},"{name}",{counter},{value}')

    add("__amd_rocclr_copyBuffer", 5.0)                  # an early copy: not the calibration
    for _ in range(4):
        add("void svs::k_sweep<1, true, true>(svs::LevelDesc const*)", 999.0)   # the plan's own run
    for lv, v in enumerate((100.0, 20.0, 4.0)):
        for k in range(n):
            add("void svs::k_sweep<1, true, true>(svs::LevelDesc const*)", v + k)
    add("void svs::k_sweep_resident<1, true, true, 1>(svs::LevelDesc const*)", 30.0)
    add("void svs::k_sweep<1, true, true>(x)", 7.0, counter="WRITE_SIZE")    # another counter: ignored
    add("__amd_rocclr_copyBuffer", 524320.0)
    sub = tmp_path / "host" / "1234"
    sub.mkdir(parents=True)
    (sub / "run_counter_collection.csv").write_text("\n".join(lines) + "\n")
    rows, header = bench._counter_rows(str(tmp_path), "FETCH_SIZE")
    assert "Counter_Value" in header and len(rows) == d - 1
    per_level, copy = bench.per_level_counter(rows, [False, False, False, True], n)
    assert per_level == [101.0, 21.0, 5.0, 10.0] and copy == 524320.0
    assert bench.per_level_counter(rows[:5], [False, False, False, True], n) is None
