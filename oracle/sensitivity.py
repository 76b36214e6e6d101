"""FP32 rounding sensitivity of the TV-L1 solve (test infrastructure).

The reference's FP32 result depends on rounding choices that its sources leave
to the toolchain: whether nvcc fuses multiply-adds inside the operators
(setup.py:43 builds with nvcc -O3, whose --fmad default is on), the error of
CUDA's expf/powf in the anisotropic weight (anisotropicDifferentialOps.cu:113),
cuDNN's summation order in the 5^3 pre-blur (flow_utils.py:110), and the order
of torch's 3-term sums (tvl13d.py:201, 229).  Each variant of the oracle
changes one of them; this script solves the same pair with every variant and
reports how far each flow lies from the default FP32 oracle and from the FP64
twin of the same iteration.

    python oracle/sensitivity.py                  # the 80^3 pair, 4 x 10 x 50
    python oracle/sensitivity.py --size 24 --schedule 2,3,10

Prints one JSON line per variant and a markdown table at the end.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# run as a script, this folder is first on the path and would shadow the package
sys.path = [q for q in sys.path if os.path.abspath(q or ".") != HERE]
sys.path.insert(0, os.path.dirname(HERE))

from oracle import oracle as O  # noqa: E402
from singleventriclesegmentation_amd import phantom  # noqa: E402
from singleventriclesegmentation_amd.flow_utils import gaussian_weights3d  # noqa: E402

VARIANTS = {
    "dot_abc": "torch's 3-term sums as (a + b) + c",
    "dot_a_bc": "torch's 3-term sums as a + (b + c)",
    "fma": "multiply-adds fused inside the operators (-ffp-contract=fast)",
    "libm": "glibc expf/powf in the anisotropic weight",
    "weight_ulp": "anisotropic weight one ulp off at half the voxels",
    "blur_f64": "pre-blur accumulated in double, rounded once",
}


def solve(frames, params, dtype=np.float32):
    f0, f1 = frames
    w = gaussian_weights3d(1.0)
    u, _ = O.solve(O.blur(f1, w, dtype=dtype), O.blur(f0, w, dtype=dtype), params, dtype=dtype)
    return u.astype(np.float64)


def compare(u, ref):
    d = u - ref
    e = np.sqrt((d ** 2).sum(-1))
    return {"max_abs": float(np.abs(d).max()), "mean_epe": float(e.mean()),
            "p99_epe": float(np.percentile(e, 99)), "p99.9_epe": float(np.percentile(e, 99.9)),
            "voxels_above_1e-3": int(np.count_nonzero(np.abs(d).max(-1) > 1e-3)),
            "voxels_above_1e-2": int(np.count_nonzero(np.abs(d).max(-1) > 1e-2))}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--size", type=int, default=80)
    ap.add_argument("--schedule", default="4,10,50", help="levels,warps,iterations")
    ap.add_argument("--variants", default=",".join(VARIANTS))
    args = ap.parse_args(argv)
    S, W, K = (int(v) for v in args.schedule.split(","))
    shape = (args.size,) * 3
    f0, f1, disp = phantom.frame_pair(shape)
    params = O.solve_params(num_scales=S, max_warps=W, max_outer_iterations=K)
    names = args.variants.split(",") if args.variants else []
    subprocess.run(["make", "-s", "-C", HERE, "all"] + [f"_variants/libsvs_oracle_{v}.so" for v in names], check=True)

    O.use_library(None)
    t0 = time.perf_counter()
    base = solve((f0, f1), params)
    t_base = time.perf_counter() - t0
    f64 = solve((f0, f1), params, dtype=np.float64)
    rows = []

    def row(name, what, u, seconds):
        r = {"variant": name, "what": what, "seconds": round(seconds, 2),
             "vs_fp32_oracle": compare(u, base), "vs_fp64": compare(u, f64),
             "mean_epe_vs_true_displacement": float(np.sqrt(((u - disp) ** 2).sum(-1)).mean())}
        print(json.dumps(r), flush=True)
        rows.append(r)

    row("fp32", "default FP32 oracle (the GPU path's arithmetic)", base, t_base)
    row("fp64", "FP64 twin of the same iteration", f64, 0.0)
    for name in names:
        O.use_library(os.path.join(HERE, "_variants", f"libsvs_oracle_{name}.so"))
        t0 = time.perf_counter()
        u = solve((f0, f1), params)
        row(name, VARIANTS[name], u, time.perf_counter() - t0)
    O.use_library(None)

    print(f"\n{shape[0]}^3 pair, {S} x {W} x {K}: flow differences in voxels")
    print("| variant | vs FP32 oracle: max | mean EPE | p99.9 EPE | voxels > 1e-3 / > 1e-2 "
          "| vs FP64: max | mean EPE | EPE vs true d |")
    print("|---|---|---|---|---|---|---|---|")
    for r in rows:
        a, b = r["vs_fp32_oracle"], r["vs_fp64"]
        print(f"| {r['variant']} ({r['what']}) | {a['max_abs']:.2e} | {a['mean_epe']:.2e} | {a['p99.9_epe']:.2e} "
              f"| {a['voxels_above_1e-3']} / {a['voxels_above_1e-2']} | {b['max_abs']:.2e} | {b['mean_epe']:.2e} "
              f"| {r['mean_epe_vs_true_displacement']:.4f} |")
    return rows


if __name__ == "__main__":
    main()
