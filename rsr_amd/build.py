"""Build librsr_amd.so (HIP kernels for gfx950 + C ABI runtime) in-tree with hipcc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librsr_amd.so")
SOURCES = ["kernels.hip", "lower.cpp", "runtime.cpp", "hostcheck.hip", "optim.hip", "jobs.cpp", "fit.cpp", "serve.cpp", "select.cpp", "simplify.cpp", "island.cpp", "ingest.cpp", "noalloc.hip", "variation.hip", "engine.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = [
    "--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC",
    # every f32 operation stays a separate IEEE operation (no FMA contraction, no fast math)
    "-ffp-contract=off", "-fno-fast-math",
    # correctly rounded f32 divide and sqrt on the device
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    # never emit packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # keep wave-uniform control flow (the interpreter's dispatch on the scalar opcode) as
    # plain scalar branches instead of structurized exec-mask regions
    "-mllvm", "-structurizecfg-skip-uniform-regions=true",
    "-Wno-unused-command-line-argument",
]


def build(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + COMMON + ["-I", CSRC, "-c", path, "-o", obj]
        if src.endswith(".hip"):
            cmd[1:1] = ["-x", "hip"]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, stderr=None if verbose else subprocess.PIPE, text=True)
        if r.returncode != 0:
            raise RuntimeError("compile of %s failed:\n%s" % (src, r.stderr or ""))
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
