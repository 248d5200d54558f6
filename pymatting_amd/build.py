"""Builds libpymatting_amd.so for gfx950 in-tree with hipcc.

The library links against the HIP runtime that ships inside PyTorch
(torch/lib/libamdhip64.so) so that a process importing torch holds a single
HIP runtime and torch's device pointers and streams are valid in the library.

Usage: python -m pymatting_amd.build   (or __graft_entry__.build())
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpymatting_amd.so")
ROOT = os.path.dirname(HERE)

SOURCES = {
    # file: extra flags
    "cf_exact.hip": ["-ffp-contract=off"],
    "grid.hip": [],
    "solver.hip": [],
    "fg.hip": ["-ffp-contract=off"],
    "scan.hip": [],
    "adjoint.hip": [],
    "batch.hip": [],
    "ichol.hip": ["-ffp-contract=off"],
    "sparse.hip": [],
    "fg_cf.hip": ["-ffp-contract=off"],
    "lkm.hip": [],
    "sync_probe.hip": [],
}

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CLANG = "/opt/rocm/llvm/bin/clang++"
COMMON = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
    "-Wno-unused-variable", "-Wno-unused-but-set-variable",
    # Keep packed-FP32 instructions out of the generated code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]


def torch_lib_dir():
    import importlib.util
    spec = importlib.util.find_spec("torch")
    if spec is None or spec.origin is None:
        raise RuntimeError("PyTorch (ROCm) is required: the library links its HIP runtime")
    return os.path.join(os.path.dirname(spec.origin), "lib")


def compile_one(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    extra = os.environ.get("HM_EXTRA_FLAGS", "").split()  # e.g. tuning -D overrides
    cmd = [HIPCC, *COMMON, *SOURCES[src], *extra, "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr))
    return out, r.stderr


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(compile_one, sorted(SOURCES)))
    objs = [o for o, _ in results]
    if verbose:
        for (_, err) in results:
            if err.strip():
                sys.stderr.write(err)
    tl = torch_lib_dir()
    tmp = LIB + ".tmp"
    cmd = [CLANG, "-shared", "-fPIC", "-o", tmp, *objs, "-L" + tl, "-lamdhip64",
           "-Wl,-rpath," + tl, "-Wl,--disable-new-dtags", "-Wl,--no-undefined"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
