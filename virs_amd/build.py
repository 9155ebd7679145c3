"""Build recipe for the native libraries (in-tree, gfx950 only).

    python -m virs_amd.build

produces
    virs_amd/lib/libvirs_hip.so   HIP kernels + C ABI   (include/virs_fdtd.h)
    virs_amd/lib/libvirs_host.so  host grid pipeline     (include/virs_host.h)
and, for the tests, the dispatch adapter (include/virs_dispatch.hpp) compiled
against a stand-in VectorSpace:
    tests/adapter/_build/libdispatch_driver.so

    python -m virs_amd.build --sanitize DIR

builds the host library with AddressSanitizer and UndefinedBehaviorSanitizer
(host code only) into DIR, next to an ordinary build of the HIP library.
"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "virs_amd")
LIB = os.path.join(PKG, "lib")
INC = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    # packed FP32 instructions stay off (see the kernels' header)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INC}",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-pthread", f"-I{INC}"]
SANITIZE_FLAGS = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd)}")


def build(verbose=False, out_dir=None, defines=(), sanitize=False):
    """out_dir / defines build a variant (e.g. VIRS_TAIL_BATCH=2) elsewhere;
    sanitize builds only the host library, with ASan + UBSan."""
    lib = out_dir or LIB
    os.makedirs(lib, exist_ok=True)
    hip_src = os.path.join(PKG, "csrc", "fdtd", "virs_fdtd.hip")
    host_src = os.path.join(PKG, "csrc", "host", "virs_host.cpp")
    hip_out = os.path.join(lib, "libvirs_hip.so")
    host_out = os.path.join(lib, "libvirs_host.so")
    extra = SANITIZE_FLAGS if sanitize else []
    _run(["g++"] + HOST_FLAGS + extra + [host_src, "-o", host_out + ".tmp", "-lz"], verbose)
    os.replace(host_out + ".tmp", host_out)
    if sanitize:  # host code only; the HIP library is built the ordinary way
        return None, host_out
    _run([HIPCC] + HIP_FLAGS + [f"-D{d}" for d in defines] + [hip_src, "-o", hip_out + ".tmp"], verbose)
    os.replace(hip_out + ".tmp", hip_out)
    return hip_out, host_out


def build_dispatch_driver(verbose=False):
    src = os.path.join(ROOT, "tests", "adapter", "dispatch_driver.cpp")
    out_dir = os.path.join(ROOT, "tests", "adapter", "_build")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "libdispatch_driver.so")
    _run(["g++"] + HOST_FLAGS + ["-Wall", "-Wextra", src, "-o", out + ".tmp", f"-L{LIB}",
                                 "-lvirs_hip", "-lvirs_host", "-Wl,-rpath,$ORIGIN/../../../virs_amd/lib"],
         verbose)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    if len(sys.argv) == 3 and sys.argv[1] == "--sanitize":
        build(verbose=True, out_dir=sys.argv[2], sanitize=True)
    else:
        build(verbose=True)
        build_dispatch_driver(verbose=True)
