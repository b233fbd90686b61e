"""Build the MI355X QZ library in-tree.

  libgqz.so          HIP kernels (gfx950) + host driver + C ABI (include/gqz.h)
  libpdhgeqz_f.so    Fortran ISO_C_BINDING shim exporting pdhgeqz_/pdgghrd_/kkqzconf_/pilaenvx_

Run directly (python gevp-pdhgeqz_amd/build.py) or through __graft_entry__.build().
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
OBJ = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("GQZ_FC", "/opt/rocm/bin/amdflang")
ARCH = os.environ.get("GQZ_ARCH", "gfx950")

HIP_SOURCES = ["k_wgqz.hip", "k_aed.hip", "k_misc.hip", "k_gemm.hip", "k_chase.hip", "k_ht.hip", "k_batched.hip",
               "k_mlp.hip", "k_ht_blk.hip"]
CXX_SOURCES = ["qz_host.cpp", "qz_desc.cpp", "mm_io.cpp"]

HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + CSRC, "-I" + os.path.join(ROOT, "include")]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + ["-c", path, "-o", obj]
    else:
        cmd = [HIPCC] + HIP_FLAGS + ["-x", "hip", "-c", path, "-o", obj]
    _run(cmd)
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in HIP_SOURCES + CXX_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, srcs))
    lib = os.path.join(HERE, "libgqz.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs +
         ["-Wl,-soname,libgqz.so"])
    os.replace(tmp, lib)
    if verbose:
        print("built", lib)
    fsh = os.path.join(FSRC, "pdhgeqz_gpu.f90")
    if os.path.exists(fsh) and shutil.which(FC):
        flib = os.path.join(HERE, "libpdhgeqz_f.so")
        ftmp = flib + ".tmp"
        _run([FC, "-O2", "-fPIC", "-shared", "-J", OBJ, "-o", ftmp, fsh, "-L" + HERE, "-lgqz",
              "-Wl,-rpath,$ORIGIN"])
        os.replace(ftmp, flib)
        if verbose:
            print("built", flib)
    return lib


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
