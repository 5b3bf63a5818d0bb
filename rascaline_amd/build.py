"""Build librascaline_amd.so (HIP for gfx950) in-tree with hipcc.

Usage: python -m rascaline_amd.build   (or rascaline_amd.build.build())
"""

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBRARY = os.path.join(HERE, "librascaline_amd.so")

# (source, extra flags). The neighbour list must not contract a*b+c into FMAs so
# that pair vectors are bit-identical to the reference's.
SOURCES = [
    ("kernels_nl.hip", ["-ffp-contract=off"]),
    ("neighbors.cpp", ["-ffp-contract=off"]),
    # matrix-core accumulators in ordinary VGPRs (none of these kernels needs the extra
    # accumulation registers; the copies out of them cost occupancy)
    ("kernels_soap.hip", ["-mllvm", "-amdgpu-mfma-vgpr-form"]),
    ("kernels_lode.hip", []),
    ("kernels_mlp.hip", []),
    ("surrogate.cpp", []),
    ("scan.hip", []),
    ("calculator.cpp", []),
    ("lode_host.cpp", []),
    ("neighbor_list_host.cpp", []),
    ("by_pair_host.cpp", []),
    ("hypers.cpp", []),
    ("radial.cpp", []),
    ("tensormap.cpp", []),
    ("selection.cpp", []),
    ("host_transfer.cpp", []),
    ("batch_graph.cpp", []),
    ("capi.cpp", []),
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RASCALINE_AMD_ARCH", "gfx950")
COMMON = [
    "-x", "hip", f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-command-line-argument",
]


def _compile(source, flags):
    src = os.path.join(CSRC, source)
    obj = os.path.join(BUILD, source + ".o")
    cmd = [HIPCC] + COMMON + flags + ["-c", src, "-o", obj]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"compilation of {source} failed:\n{' '.join(cmd)}\n{result.stdout}\n{result.stderr}")
    return obj


def build(jobs=None, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(SOURCES), int(os.environ.get("MAX_JOBS", "8")))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        futures = [pool.submit(_compile, s, f) for s, f in SOURCES]
        objects = [f.result() for f in futures]
    tmp = LIBRARY + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objects + ["-lpthread"]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{result.stdout}\n{result.stderr}")
    os.replace(tmp, LIBRARY)
    if verbose:
        print(f"built {LIBRARY}")
    return LIBRARY


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
