"""Build libtsem_amd.so (host C++ + HIP kernels for gfx950) in-tree.

    python -m tsem_amd.build          # or tsem_amd.build.build()

Every source is compiled with hipcc for --offload-arch=gfx950; packed FP32
instructions are disabled (-packed-fp32-ops) for all device code.  The C++
example programs (examples/*.cpp, host code over include/tsem.hpp) are then
linked against the library into examples/bin/.
"""
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtsem_amd.so")
SOURCES = ["refel.cpp", "space.cpp", "slab.cpp", "capi.cpp", "serve.cpp", "eigen.cpp", "quad_api.cpp", "kernels.hip", "apply_mfma.hip",
           "apply_mfma_h.hip", "pmg.hip", "block.hip", "quad.hip", "noalloc.hip", "source.hip", "bc.hip"]
EXAMPLES = ["poisson_mixbc", "laplace_eig"]
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-result", "-Wno-unused-variable"]


def _compile(src, obj):
    lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
    cmd = [HIPCC] + FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
    subprocess.check_call(cmd)


def build(jobs=None, verbose=True):
    """Compile all sources (in parallel) and link the shared library."""
    from concurrent.futures import ThreadPoolExecutor
    tmp = tempfile.mkdtemp(prefix="tsem_build_")
    objs = [os.path.join(tmp, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    with ThreadPoolExecutor(max_workers=jobs or min(8, len(SOURCES))) as ex:
        list(ex.map(lambda a: _compile(*a), zip(SOURCES, objs)))
    out_tmp = LIB + ".tmp"
    subprocess.check_call([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", out_tmp] + objs + ["-ldl"])
    os.replace(out_tmp, LIB)
    for o in objs:
        os.remove(o)
    os.rmdir(tmp)
    build_examples()
    if verbose:
        print("built", LIB)
    return LIB


def build_examples():
    """Link the example programs against the in-tree library (rpath to it)."""
    out = os.path.join(HERE, "examples", "bin")
    os.makedirs(out, exist_ok=True)
    for name in EXAMPLES:
        exe = os.path.join(out, name)
        subprocess.check_call(["g++", "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__",
                               "-I/opt/rocm/include", "-I" + INCLUDE, os.path.join(HERE, "examples", name + ".cpp"),
                               "-o", exe + ".tmp", "-L" + HERE, "-ltsem_amd", "-L/opt/rocm/lib", "-lamdhip64",
                               "-Wl,-rpath,$ORIGIN/../..:/opt/rocm/lib"])
        os.replace(exe + ".tmp", exe)


if __name__ == "__main__":
    build()
    sys.exit(0)
