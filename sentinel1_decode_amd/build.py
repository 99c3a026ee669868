"""Build recipe for the gfx950 library (libs1gpu.so), the C++ drop-in surface
over it (libs1dropin.so), and the synthetic-scene library (synth/libs1synth.so).  Run ``python -m sentinel1_decode_amd.build``."""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libs1gpu.so")
ARCH = os.environ.get("S1G_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 in generated code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


def _headers():
    return glob.glob(os.path.join(CSRC, "*.h")) + [os.path.join(HERE, "..", "include", "s1gpu.h")]


def build_gpu(verbose=False, jobs=8, force=True):
    """Compile every .hip source (force=False skips objects newer than their
    source and all headers) and link libs1gpu.so."""
    os.makedirs(os.path.join(HERE, "_obj"), exist_ok=True)
    objs, procs = [], []
    newest_header = max(os.path.getmtime(h) for h in _headers())
    for src in sources():
        obj = os.path.join(HERE, "_obj", os.path.basename(src) + ".o")
        objs.append(obj)
        if (not force and os.path.exists(obj) and os.path.getmtime(obj) > os.path.getmtime(src)
                and os.path.getmtime(obj) > newest_header):
            continue
        cmd = [HIPCC, *HIP_FLAGS, "-I", os.path.join(HERE, "..", "include"), "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len([p for _, p in procs if p.poll() is None]) >= jobs:
            procs[0][1].wait()
    failed = False
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed = True
            sys.stderr.write(f"--- {os.path.basename(src)} failed\n{text}\n")
        elif verbose and text.strip():
            sys.stderr.write(text)
    if failed:
        raise RuntimeError("HIP build failed")
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs], check=True)
    os.replace(tmp, LIB)
    return LIB


DROPIN = os.path.join(HERE, "libs1dropin.so")
CPP_TEST = os.path.join(HERE, "..", "tests", "cpp", "test_dropin")


def build_dropin(verbose=False):
    """The C++ drop-in surface (include/s1/*.h) over libs1gpu.so, and the
    C++ test program that exercises it."""
    inc = os.path.join(HERE, "..", "include")
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    # host-only C++: g++ with the HIP runtime headers and library
    host = ["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-I", inc]
    tmp = DROPIN + ".tmp"
    cmd = [*host, "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(rocm, "include"), "-shared", "-o", tmp,
           os.path.join(HERE, "cpp", "dropin.cpp"), "-L", HERE, "-ls1gpu", "-L", os.path.join(rocm, "lib"),
           "-lamdhip64", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(rocm, "lib")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, DROPIN)
    src = os.path.join(HERE, "..", "tests", "cpp", "test_dropin.cpp")
    if os.path.exists(src):
        cmd = [*host, "-o", CPP_TEST, src, "-L", HERE, "-ls1dropin", "-ls1gpu",
               "-Wl,-rpath,$ORIGIN/../../sentinel1_decode_amd"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return DROPIN


def build_synth():
    from .synth import build as b
    return b(force=True)


def main():
    build_gpu(verbose="-v" in sys.argv, force="--force" in sys.argv)
    build_dropin(verbose="-v" in sys.argv)
    build_synth()
    print(LIB)


if __name__ == "__main__":
    main()
