"""Compile the HIP/C++ sources into gammaray_amd/libgammaray_amd.so (gfx950).

Run as ``python -m gammaray_amd.build``.  Every translation unit is built with
-ffp-contract=off (bit-exact f32 traversal: the reference never fuses
multiply-adds) and with the packed-FP32 instructions disabled.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgammaray_amd.so")
ARCH = os.environ.get("GR_OFFLOAD_ARCH", "gfx950")

SOURCES = ["gr_bvh.cpp", "gr_api.cpp", "gr_exr.cpp", "gr_obj.cpp", "gr_traverse.hip", "gr_sort.hip", "gr_guide.hip", "gr_render.hip",
           "gr_noalloc.hip"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-I" + os.path.join(HERE, "..", "include")]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found")
    return path


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + COMMON + DEVICE + ["-x", "hip"]
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    # hipcc's host pass does not know the packed-fp32 feature; that warning is
    # expected and not echoed.
    return obj, r.stderr


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    objs = [o for o, _ in results]
    if verbose:
        for _, err in results:
            lines = [l for l in err.splitlines() if "packed-fp32-ops" not in l and l.strip()]
            if lines:
                print("\n".join(lines), file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    build_cli()
    return LIB


CLI_SRC = os.path.join(CSRC, "gr_cli.cpp")
CLI = os.path.join(HERE, "gammaray")


def build_cli():
    """The progressive render driver (the reference binary's loop), linked
    against the library next to it."""
    gxx = shutil.which("g++") or "g++"
    tmp = CLI + ".tmp"
    cmd = [gxx, "-O2", "-std=c++17", "-Wall", "-pthread", "-o", tmp, CLI_SRC, "-L" + HERE, "-lgammaray_amd",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("cli build failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, CLI)
    return CLI


if __name__ == "__main__":
    print(build(verbose=True))
