"""Build libxmp_amd.so (HIP kernels + C ABI) in-tree for gfx950.

Usage: python -m xmp_amd.build   (or xmp_amd.build.build())

Every translation unit is compiled with hipcc for --offload-arch=gfx950 and
linked into xmp_amd/libxmp_amd.so.  Packed-FP32 instructions are disabled
(-mno-packed-fp32-ops equivalent) for every HIP unit.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(ROOT, "_build", "obj")
LIB = os.path.join(HERE, "libxmp_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = [
    "api.cpp",
    "ops.cpp",
    "tuner.cpp",
    "batch.cpp",
    "serve.cpp",
    "ext.cpp",
    "logic.cpp",
    "k_xfer.hip",
    "k_arith.hip",
    "k_div.hip",
    "k_logic.hip",
    "k_powm.hip",
    "k_powm_latency.hip",
    "k_powm_wide.hip",
    "k_probe.hip",
    "k_noalloc.hip",
]

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden", "-Wall", "-Wno-unused-result",
    "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]

# Per-file device-only options.  The lane-layout Montgomery kernels are
# scheduled by LLVM's ILP-oriented iterative GCN scheduler: measured A/B on one
# MI355X, the 65,536 x 2048-bit powm launch took 48.5-48.7 ms against
# 49.4-49.7 ms with the default occupancy-oriented scheduler (the kernel runs
# at its register-limited occupancy of two waves per SIMD either way).  The
# few-digit latency layouts (k_powm_latency.hip) keep the default scheduler,
# and so does k_arith.hip, whose build the ILP scheduler crashes (the ROCm 7.2
# compiler ends with a segmentation fault).
DEVICE_OPTS = {
    "k_powm.hip": ["-Xarch_device", "-mllvm=-misched=gcn-iterative-ilp"],
}


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    lang = ["-x", "hip"]
    cmd = [HIPCC] + COMMON + DEVICE_OPTS.get(src, []) + lang + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    return src, r.returncode, msgs


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    sources = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = min(len(sources), max(1, (os.cpu_count() or 4)))
    failed = []
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        for src, rc, msgs in ex.map(_compile, sources):
            if rc != 0:
                failed.append((src, msgs))
            elif verbose and msgs.strip():
                print(msgs)
    if failed:
        for src, msgs in failed:
            sys.stderr.write("==== %s\n%s\n" % (src, msgs))
        raise RuntimeError("HIP build failed: " + ", ".join(s for s, _ in failed))
    objs = [os.path.join(OBJ, s + ".o") for s in sources]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread", "-lrocprofiler-sdk-roctx"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
