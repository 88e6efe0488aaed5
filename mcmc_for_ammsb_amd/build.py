"""Build libammsb.so (gfx950) in-tree with hipcc.

Every HIP/C++ source under csrc/ is compiled to an object in parallel and
linked into mcmc_for_ammsb_amd/libammsb.so, which the Python front end loads
with ctypes.  Run as ``python mcmc_for_ammsb_amd/build.py`` (a script, so a
stale library never has to be imported to rebuild it).
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libammsb.so")
DRIVER = os.path.join(HERE, "ammsb_mcmc")  # command-line driver executable
OBJDIR = os.path.join(HERE, "_build")

ARCH = os.environ.get("AMMSB_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", shutil.which("g++") or "g++")

# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 in generated code.
COMMON = ["-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas"]

SOURCES = ["kernels.hip", "capi.hip", "host_graph.cpp", "host_options.cpp"]


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):  # host-only translation unit
        cmd = [CXX, "-std=c++17", "-O3", "-fPIC", "-Wall", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    # the host pass does not know the packed-fp32 target feature: harmless
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{err}")
    return obj, err


def build(verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        results = list(ex.map(_compile, SOURCES))
    objs = [o for o, _ in results]
    if verbose:
        for _, e in results:
            if e.strip():
                print(e, file=sys.stderr)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp] + objs + ["-lpthread", "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stderr}")
    os.replace(tmp, OUT)
    # the command-line driver (the reference's main), linked against the library
    drv_tmp = DRIVER + ".tmp"
    cmd = [CXX, "-std=c++17", "-O2", "-Wall", os.path.join(CSRC, "ammsb_main.cpp"), "-o", drv_tmp,
           "-L" + HERE, "-lammsb", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver link failed: {' '.join(cmd)}\n{r.stderr}")
    os.replace(drv_tmp, DRIVER)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
