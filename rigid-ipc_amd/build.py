"""Build librigc.so (HIP, gfx950) in-tree.

    python rigid-ipc_amd/build.py [--jobs N]

Every .hip source under csrc/ is compiled with hipcc for gfx950 and linked
into rigid-ipc_amd/librigc.so. Floating-point contraction is off so that
every product and sum rounds once, as on the host; packed FP32 is disabled.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "librigc.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def compile_one(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj, r.stderr


HOST = os.path.join(HERE, "host")
SIM_OUT = os.path.join(HERE, "librigc_sim.so")


def build_host(verbose=False):
    """The host time stepper (host/simulation.cpp, host/batch.cpp) as librigc_sim.so,
    linked against librigc.so (found next to it at run time)."""
    tmp = SIM_OUT + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-shared", "-pthread", os.path.join(HOST, "simulation.cpp"),
                             os.path.join(HOST, "batch.cpp"), "-L" + HERE, "-lrigc", "-Wl,-rpath,$ORIGIN",
                             "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for the host stepper:\n{r.stderr}")
    if verbose and r.stderr.strip():
        print(r.stderr)
    os.replace(tmp, SIM_OUT)
    return SIM_OUT


def build(jobs=8, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(compile_one, srcs))
    objs = [o for o, _ in results]
    if verbose:
        for _, err in results:
            if err.strip():
                print(err)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    build_host(verbose)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("-v", "--verbose", action="store_true")
    args = ap.parse_args()
    print(build(args.jobs, args.verbose))
    sys.exit(0)
