"""Build the gfx950 shared library ifos_amd/lib/libifos2d_hip.so.

HIP sources are compiled by hipcc for gfx950 only, host C by gcc; both with
-ffp-contract=off so that the FP32 arithmetic matches the reference's CPU
build (gcc -O3 -std=c99 on x86-64, no FMA) bit for bit.  Device code keeps
IEEE denormals and is built without packed-FP32 instructions.

Usage: python -m ifos_amd.build  (or build() from __graft_entry__)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libifos2d_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value", "-Wno-unused-result",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]
C_FLAGS = ["-O2", "-g", "-std=gnu11", "-fPIC", "-ffp-contract=off", "-Wall",
           "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]

HIP_SOURCES = ["ifos_kernels.hip", "ifos_step.hip", "ifos_engine.hip", "ifos_dropin.hip", "ifos_opt.hip", "ifos_fp64.hip"]
C_SOURCES = ["host/ifos_setup.c", "host/ifos_formats.c", "host/ifos_taper.c", "host/ifos_filter.c", "host/ifos_stf.c"]
CXX_SOURCES = ["ifos_serve.cpp"]
CXX_FLAGS = ["-O2", "-g", "-std=c++17", "-fPIC", "-pthread", "-Wall", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return out


def sources():
    """Every source file the library is linked from, relative to csrc/."""
    return list(C_SOURCES) + list(CXX_SOURCES) + [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]


def build(verbose=False, obj_dir=None, lib_path=None):
    """Compile every source and link the library.  obj_dir / lib_path default
    to the in-tree ifos_amd/_build and ifos_amd/lib/libifos2d_hip.so; every
    object is compiled on every call (no timestamp shortcuts)."""
    obj_dir = obj_dir or OBJ
    lib_path = lib_path or LIB
    os.makedirs(obj_dir, exist_ok=True)
    os.makedirs(os.path.dirname(lib_path), exist_ok=True)
    objs = []
    jobs = []
    for src in C_SOURCES:
        o = os.path.join(obj_dir, os.path.basename(src).replace(".c", ".o"))
        jobs.append(["gcc", *C_FLAGS, "-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    for src in CXX_SOURCES:
        o = os.path.join(obj_dir, os.path.basename(src).replace(".cpp", ".o"))
        jobs.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        o = os.path.join(obj_dir, os.path.basename(src).replace(".hip", ".o"))
        jobs.append([HIPCC, *HIP_FLAGS, "-c", path, "-o", o])
        objs.append(o)
    procs = [(j, subprocess.Popen(j, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)) for j in jobs]
    failed = False
    for j, p in procs:
        out, _ = p.communicate()
        out = "\n".join(l for l in out.splitlines()
                        if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
        if p.returncode != 0:
            sys.stderr.write(out + "\n")
            failed = True
        elif verbose and out.strip():
            print(out)
    if failed:
        raise RuntimeError("compilation failed")
    tmp = lib_path + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs,
          "-Wl,--no-undefined", "-lm", "-lpthread"])
    os.replace(tmp, lib_path)
    return lib_path


if __name__ == "__main__":
    print(build(verbose=True))
