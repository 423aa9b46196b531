"""Build every native artefact of the package, in tree.

    python -m listctrecons_amd.build          # liblct.so, C++ drop-in, CLI, oracle

Outputs (git-ignored, shipped with the tree):
  listctrecons_amd/liblct.so            HIP kernels + C ABI (include/lct.h), gfx950
  listctrecons_amd/libListCTRecons.so   C++ drop-in classes over the C ABI
  listctrecons_amd/bin/LISTCTRecons     the reference CLI (LISTCTRecons <prm>)
  listctrecons_amd/bin/issue_rates      issue cost of VALU / SALU / LDS instruction classes
  oracle/liblct_oracle.so               FP64 CPU oracle (test infrastructure)
  oracle/_ref/*                         reference sources that compile alone, and the
                                        reference's main.cpp linked to the package
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
CPP = os.path.join(PKG, "cpp")
INC = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LCT_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 ops disabled for every device function
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    f"-I{INC}", f"-I{CSRC}",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{INC}", f"-I{os.path.join(INC, 'listctrecons')}"]

LIB = os.path.join(PKG, "liblct.so")
CPPLIB = os.path.join(PKG, "libListCTRecons.so")
CLI = os.path.join(PKG, "bin", "LISTCTRecons")


def _run(cmd, cwd=ROOT):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}")
    return r.stdout


def _objects(sources, compiler, flags, jobs):
    os.makedirs(BUILD, exist_ok=True)
    outs = []
    cmds = []
    for src in sources:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        outs.append(obj)
        cmds.append([compiler, *flags, "-c", src, "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if out.strip():
                sys.stderr.write(out)
    return outs


def build_lct(jobs: int = 8) -> str:
    hip_srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpp_srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".cpp"))
    objs = _objects(hip_srcs, HIPCC, HIP_FLAGS, jobs)
    objs += _objects(cpp_srcs, "g++", CXX_FLAGS, jobs)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_cpp(jobs: int = 8) -> str:
    if not os.path.isdir(CPP):
        return ""
    srcs = sorted(os.path.join(CPP, f) for f in os.listdir(CPP)
                  if f.endswith(".cpp") and f != "main.cpp")
    if not srcs:
        return ""
    objs = _objects(srcs, "g++", CXX_FLAGS, jobs)
    tmp = CPPLIB + ".tmp"
    _run(["g++", "-shared", "-fPIC", "-o", tmp, *objs, f"-L{PKG}", "-llct",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, CPPLIB)
    main = os.path.join(CPP, "main.cpp")
    if os.path.exists(main):
        os.makedirs(os.path.dirname(CLI), exist_ok=True)
        _run(["g++", *CXX_FLAGS, "-o", CLI + ".tmp", main, f"-L{PKG}", "-lListCTRecons",
              "-llct", "-Wl,-rpath,$ORIGIN/.."])
        os.replace(CLI + ".tmp", CLI)
    return CPPLIB


def build_oracle() -> str:
    oracle_dir = os.path.join(ROOT, "oracle")
    _run(["make", "-B", "-s", "-C", oracle_dir])
    ref_mk = os.path.join(oracle_dir, "ref.mk")
    if os.path.isdir("/root/reference") and os.path.exists(ref_mk):
        _run(["make", "-B", "-s", "-C", oracle_dir, "-f", "ref.mk"])
    return os.path.join(oracle_dir, "liblct_oracle.so")


NATIVE_TESTS = os.path.join(ROOT, "tests", "native")


def build_native_tests() -> None:
    """C++ test programs of the drop-in headers (tests/native/*.cpp) into
    tests/native/bin/, linked against the built libraries."""
    if not os.path.isdir(NATIVE_TESTS):
        return
    out_dir = os.path.join(NATIVE_TESTS, "bin")
    os.makedirs(out_dir, exist_ok=True)
    for f in sorted(os.listdir(NATIVE_TESTS)):
        if not f.endswith(".cpp"):
            continue
        exe = os.path.join(out_dir, f[:-4])
        _run(["g++", *CXX_FLAGS, "-o", exe + ".tmp", os.path.join(NATIVE_TESTS, f), f"-L{PKG}",
              "-lListCTRecons", "-llct", "-Wl,-rpath,$ORIGIN/../../../listctrecons_amd"])
        os.replace(exe + ".tmp", exe)


def build_ref_main() -> None:
    """The reference's Src/main.cpp compiled unchanged against the drop-in
    headers and linked to the package (oracle/_ref/ref_main), when the
    reference sources are present."""
    oracle_dir = os.path.join(ROOT, "oracle")
    if os.path.isdir("/root/reference"):
        _run(["make", "-B", "-s", "-C", oracle_dir, "-f", "ref.mk", "_ref/ref_main"])


TOOLS = os.path.join(PKG, "tools")


def build_tools() -> None:
    """Measurement programs (tools/*.hip) into bin/: issue_rates, the issue
    cost of the instruction classes the backprojector uses."""
    if not os.path.isdir(TOOLS):
        return
    out_dir = os.path.join(PKG, "bin")
    os.makedirs(out_dir, exist_ok=True)
    for f in sorted(os.listdir(TOOLS)):
        if not f.endswith(".hip"):
            continue
        exe = os.path.join(out_dir, f[:-4])
        _run([HIPCC, "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-Xclang", "-target-feature",
              "-Xclang", "-packed-fp32-ops", "-o", exe + ".tmp", os.path.join(TOOLS, f)])
        os.replace(exe + ".tmp", exe)


def build_all(jobs: int = 8) -> None:
    build_oracle()
    build_lct(jobs)
    build_cpp(jobs)
    build_native_tests()
    build_ref_main()
    build_tools()


if __name__ == "__main__":
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        sys.exit(f"hipcc not found at {HIPCC}")
    build_all()
    print("built:", LIB, CPPLIB, CLI)
