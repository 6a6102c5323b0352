"""Build libketsv.so (gfx950) in-tree with hipcc.

The library is the product: every HIP kernel and the C ABI declared in
include/ketsv.h.  Called by __graft_entry__.build(); also runnable as
``python -m ket_amd.build``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libketsv.so")
SOURCES = ["capi.hip", "k_pass.hip", "k_rpass.hip", "k_bpass.hip", "k_reduce.hip", "k_adjoint.hip", "k_qaoa.hip", "k_shots.hip", "k_dm.hip", "schedule.cpp", "adjplan.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions are disabled on purpose for every HIP build
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


# Per-source flags.  k_bpass.hip: MFMA accumulators in VGPRs, so the block
# epilogue reads them without v_accvgpr_read moves (VALU work that does not
# overlap the FP64 MFMA work).
SOURCE_FLAGS = {"k_bpass.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def build(verbose=False):
    """Compile every source and link libketsv.so; raises on failure."""
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = ([HIPCC] + FLAGS + SOURCE_FLAGS.get(src, []) + ["-I", os.path.join(ROOT, "include")] + lang +
               ["-c", os.path.join(CSRC, src), "-o", obj])
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((src, out))
        elif verbose and out.strip():
            print(out)
    if failed:
        msg = "\n".join(f"--- {s} ---\n{o}" for s, o in failed)
        raise RuntimeError("libketsv build failed:\n" + msg)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("libketsv link failed:\n" + r.stdout)
    os.replace(tmp, LIB)
    build_napi(verbose)
    return LIB


NAPI_SRC = os.path.join(HERE, "napi", "ketsv_napi.c")
NAPI_OUT = os.path.join(LIBDIR, "ketsv.node")
NODE_INCLUDE = "/usr/include/node"


def build_napi(verbose=False):
    """The Node-API addon (plain C over libketsv.so, no HIP code), when Node's
    headers are installed; returns its path or None."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        return None
    tmp = NAPI_OUT + ".tmp"
    cmd = ["gcc", "-O2", "-shared", "-fPIC", "-std=c11", "-Wall", "-I", NODE_INCLUDE, "-I", os.path.join(ROOT, "include"),
           NAPI_SRC, "-L", LIBDIR, "-lketsv", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("ketsv.node build failed:\n" + r.stdout)
    os.replace(tmp, NAPI_OUT)
    return NAPI_OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
