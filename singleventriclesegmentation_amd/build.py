"""Builds libsvsflow.so (every HIP source under csrc/) for gfx950 with hipcc.

Flags: -ffp-contract=off keeps the reference's rounding order (no implicit
multiply-add); packed FP32 instructions are disabled.

``python -m singleventriclesegmentation_amd.build --host-sanitize[=address|undefined]``
builds libsvsflow_host_sanitize.so instead: the same device code, with the host
code (plan building, tile tables, staging ring, argument checks) instrumented
by AddressSanitizer and UndefinedBehaviorSanitizer (default) or by one of them.
Load it with SVS_FLOW_LIB pointing at it and clang's runtime in LD_PRELOAD
(/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so, or
libclang_rt.ubsan_standalone-x86_64.so for the undefined-only build), with
ASAN_OPTIONS=detect_leaks=0.  Next to the HIP runtime on a GPU, ASan's
allocator cannot start (it reports out of memory at the runtime's first
allocation), so GPU runs use the undefined-only build; the CPU tests of the
C ABI run under both.
"""
import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_PATH = os.path.join(PKG_DIR, "libsvsflow.so")
SANITIZE_LIB_PATH = os.path.join(PKG_DIR, "libsvsflow_host_sanitize.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-mcode-object-version=5",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


# Host-only sanitizers: each -fsanitize= applies to the host pass alone, so the
# device code is the product's.
def host_sanitize_flags(which=("address", "undefined")):
    out = ["-g", "-Xarch_host", "-fno-omit-frame-pointer"]
    for name in which:
        out += ["-Xarch_host", f"-fsanitize={name}"]
    return out


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def build(verbose=False, host_sanitize=None):
    """Compiles every HIP source into LIB_PATH, or with ``host_sanitize`` (a
    tuple of sanitizer names) into SANITIZE_LIB_PATH; raises on failure."""
    out = SANITIZE_LIB_PATH if host_sanitize else LIB_PATH
    flags = FLAGS + (host_sanitize_flags(host_sanitize) if host_sanitize else [])
    cmd = [HIPCC, *flags, "-I", os.path.join(REPO_DIR, "include"), "-I", CSRC, "-o", out + ".tmp", *sources()]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed building {os.path.basename(out)}")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    which = None
    for a in sys.argv[1:]:
        if a == "--host-sanitize":
            which = ("address", "undefined")
        elif a.startswith("--host-sanitize="):
            which = tuple(a.split("=", 1)[1].split(","))
    print(build(verbose=True, host_sanitize=which))
