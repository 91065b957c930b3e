"""Build the gfx950 shared library of the SWE patch path in-tree.

    python -m peanoclaw_amd.build

produces peanoclaw_amd/libpeanoclaw_amd.so (hipcc, --offload-arch=gfx950).
FP64 arithmetic is compiled with -ffp-contract=off so no multiply-add is
fused behind the source's back; packed FP32 instructions are disabled.
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
SRC = os.path.join(PKG, "csrc", "pc_abi.hip")
DEPS = [os.path.join(PKG, "csrc", f) for f in
        ("pc_abi.hip", "pc_kernels.hpp", "pc_patch_kernels.hpp", "pc_server.hpp",
         "pc_amr.hpp", "pc_vtu.hpp", "swe_device.hpp")] + \
       [os.path.join(ROOT, "include", "peanoclaw_amd.h")]
LIB = os.path.join(PKG, "libpeanoclaw_amd.so")
ARCH = "gfx950"


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flags():
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-ffp-contract=off", "-fno-fast-math",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-result",
    ]


def source_hash():
    """SHA-256 over the library's sources and compile flags: identifies the
    build a committed profile was measured on."""
    import hashlib
    h = hashlib.sha256()
    for f in DEPS:
        with open(f, "rb") as fh:
            h.update(os.path.basename(f).encode() + b"\0" + fh.read())
    h.update(" ".join(flags()).encode())
    return h.hexdigest()


def build(verbose=False, out=None, defines=()):
    """Compile the library; `defines` are extra -D tuning macros (e.g.
    PC_UPDATE_WAVES_PER_EU=4) and `out` another output path."""
    lib = out or LIB
    tmp = lib + ".tmp"
    cmd = [hipcc()] + flags() + [f"-D{d}" for d in defines] + ["-o", tmp, SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for " + SRC)
    os.replace(tmp, lib)
    return lib


REPLAY_SRC = os.path.join(ROOT, "tests", "cpp", "numerics_replay.cpp")
REPLAY = os.path.join(ROOT, "tests", "cpp", "numerics_replay")


def build_replay(verbose=False):
    """Compile the C++ Numerics-adapter replay program of the tests (host
    code only, g++, linked against the library built by `build`)."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-Werror",
           "-I", os.path.join(ROOT, "include"), "-o", REPLAY + ".tmp", REPLAY_SRC,
           "-L", PKG, "-lpeanoclaw_amd", "-Wl,-rpath,$ORIGIN/../../peanoclaw_amd",
           "-Wl,-rpath-link,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed for " + REPLAY_SRC)
    os.replace(REPLAY + ".tmp", REPLAY)
    return REPLAY


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_replay(verbose=True))
