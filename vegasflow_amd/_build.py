"""Build recipe of libmivegas.so (gfx950), used by ``__graft_entry__.build``.

Every ``csrc/*.hip`` source is compiled by hipcc for gfx950 only and linked
into ``vegasflow_amd/lib/libmivegas.so`` next to this file, so the library
travels with the source tree. ``csrc/probe/issue_cost.hip``, the measurement
of VALU issue costs that bench.py reports, becomes ``lib/libmvprobe.so`` of
its own (``build_probe``). Flags:

- ``-ffp-contract=off``: no implicit FMA contraction, so the index arithmetic
  (xn = 50 * (1 - u), the bin cast) follows the reference's rounding; FMAs
  are written explicitly where they are wanted;
- ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32
  (v_pk_*_f32) instructions are generated.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIBDIR = os.path.join(HERE, "lib")
LIBNAME = "libmivegas.so"
PROBE_SRC = os.path.join(CSRC, "probe", "issue_cost.hip")
PROBE_LIB = "libmvprobe.so"
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-I", INCLUDE,
]


def hipcc():
    exe = os.environ.get("HIPCC") or shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError(f"hipcc not found ({exe}); the HIP toolchain is required")
    return exe


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def lib_path():
    return os.path.join(LIBDIR, LIBNAME)


def build(verbose=False, jobs=None):
    """Compile all sources and link the shared library; returns its path."""
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()
    srcs = sources()

    def compile_one(src):
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        cmd = [cc, *HIP_FLAGS, "-c", src, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
        if verbose and r.stderr.strip():
            print(r.stderr, file=sys.stderr)
        return obj

    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 2) // 2), 8)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    out = lib_path()
    tmp = out + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link of {LIBNAME} failed:\n{r.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


def probe_path():
    return os.path.join(LIBDIR, PROBE_LIB)


def build_probe(verbose=False):
    """Compile and link the issue-cost probe library; returns its path."""
    os.makedirs(LIBDIR, exist_ok=True)
    out = probe_path()
    tmp = out + ".tmp"
    cmd = [hipcc(), *HIP_FLAGS, "-shared", PROBE_SRC, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {PROBE_SRC}:\n{r.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    build(verbose=True)
    build_probe(verbose=True)
