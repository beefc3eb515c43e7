"""Build recipe for libbwg.so (HIP, gfx950) and libbwg_workload.so (host C++).

Invoked by ``__graft_entry__.build()``; compiles in-tree so the shared objects travel with the
repository snapshot.  Every HIP translation unit is compiled with packed-FP32 code generation
disabled (the kernels are integer-only, so this only guards against the compiler introducing
packed float ops anywhere).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["plain_kernels.hip", "ext_batched.hip", "ksw_batched.hip", "rescue_windows.hip", "global_batched.hip",
               "noalloc.hip", "bwg_batched.cpp", "bwg_runtime.cpp", "bwg_api.cpp", "bwg_server.cpp", "bwg_tuner.cpp"]
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
# Per-source flags.  ext_batched.hip: without the load-store vectorizer the group kernel's score
# bytes stay single ds_read_i8 loads; merged into wide loads they each cost a bit-field extract.
# Without the SLP vectorizer a slot's two 16-bit halves (H of one column, E of the next) stay two
# ds_write_b16; merged, each slot's store costs a v_perm_b32 in a kernel bound by VALU issue.
SOURCE_FLAGS = {"ext_batched.hip": ["-mllvm", "-amdgpu-load-store-vectorizer=false", "-fno-slp-vectorize"]}
LIB = os.path.join(HERE, "libbwg.so")
WORKLOAD_LIB = os.path.join(HERE, "libbwg_workload.so")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        return _run([HIPCC] + HIP_FLAGS + SOURCE_FLAGS.get(src, []) + lang + ["-c", os.path.join(CSRC, src), "-o", obj]), obj

    with ThreadPoolExecutor(max_workers=min(8, len(HIP_SOURCES))) as ex:
        results = list(ex.map(compile_one, HIP_SOURCES))
    objs = [o for _, o in results]
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    tmpw = WORKLOAD_LIB + ".tmp"
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread",
          os.path.join(CSRC, "workload.cpp"), os.path.join(CSRC, "serve_driver.cpp"), "-o", tmpw,
          "-L" + HERE, "-lbwg", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmpw, WORKLOAD_LIB)
    if verbose:
        for out, _ in results:
            if out.strip():
                print(out)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print("built", LIB, WORKLOAD_LIB)
    sys.exit(0)
