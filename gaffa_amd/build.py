"""Build the native parts of gaffa_amd in-tree.

* ``libgaffa_hip.so`` - the HIP kernels and the C ABI (include/gaffa_hip.h),
  compiled by hipcc for gfx950 only.
* ``_core*.so``       - the C++ host layer (gaffa's plan / result types and
  GPU programs) plus its pybind11 module, compiled by g++ and linked against
  libgaffa_hip.so through an $ORIGIN rpath.

Run ``python -m gaffa_amd.build`` (or ``__graft_entry__.build()``).

``python gaffa_amd/build.py --sanitize`` rebuilds ``_core`` in place with
AddressSanitizer and UndefinedBehaviorSanitizer on the host code (g++ only:
the HIP library is unchanged, GPU code is never built with them), and
``make -C oracle SANITIZE=1 -B oracle`` does the same for the CPU oracle.  Run
the CPU tests with the runtimes preloaded::

    LD_PRELOAD="$(g++ -print-file-name=libasan.so) $(g++ -print-file-name=libubsan.so)" \
    ASAN_OPTIONS=detect_leaks=0 UBSAN_OPTIONS=halt_on_error=1:print_stacktrace=1 \
    python -m pytest tests -m "not gpu"

then ``python gaffa_amd/build.py`` and ``make -C oracle -B oracle`` again for
the normal build (build.py run as a file: importing the package would load the
sanitized module first).
"""

from __future__ import annotations

import os
import shlex
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_SOURCES = ["runtime.hip", "dedisperse.hip", "preprocess.hip", "ffa.hip", "synth.hip",
               "fold.hip", "scorer.hip", "candidate.hip"]
HOST_SOURCES = ["host/dedispersion.cpp", "host/preprocessing.cpp", "host/ffa_plan.cpp",
                "host/ffa_gpu.cpp", "host/hip_runtime.cpp", "host/synth.cpp", "host/pipeline.cpp",
                "host/ffa_batch.cpp", "host/candidate.cpp", "host/folding.cpp",
                "host/filterbank.cpp", "host/scorer.cpp",
                "host/bindings.cpp"]

# gfx950 only.  -ffp-contract=off keeps every FP32/FP64 operation separately
# rounded, as in gaffa's no-FMA x86-64 CPU build; packed FP32 is disabled.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", f"-I{INCLUDE}", f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(" ".join(shlex.quote(c) for c in cmd) + "\n")
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ({res.returncode})")


def _newer(target: str, deps: list[str]) -> bool:
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def hip_library(force: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    headers = [os.path.join(INCLUDE, "gaffa_hip.h"), os.path.join(CSRC, "common.hpp")]
    objs, jobs = [], []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        if force or _newer(obj, [path] + headers):
            jobs.append([HIPCC, *HIP_FLAGS, "-c", path, "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        list(ex.map(_run, jobs))
    lib = os.path.join(PKG, "libgaffa_hip.so")
    if force or jobs or not os.path.exists(lib):
        _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", lib, *objs])
    return lib


SANITIZE_FLAGS = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer",
                  "-fno-sanitize-recover=undefined"]


def core_module(force: bool = True, sanitize: bool = False) -> str:
    import pybind11

    os.makedirs(BUILD, exist_ok=True)
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    target = os.path.join(PKG, "_core" + suffix)
    srcs = [os.path.join(CSRC, s) for s in HOST_SOURCES]
    hdrs = [os.path.join(CSRC, "host", h) for h in os.listdir(os.path.join(CSRC, "host"))
            if h.endswith(".h")]
    if not force and not _newer(target, srcs + hdrs + [os.path.join(INCLUDE, "gaffa_hip.h")]):
        return target
    py_inc = sysconfig.get_paths()["include"]
    flags = ["-std=c++20", "-O2", "-fPIC", "-fvisibility=hidden", "-fopenmp",
             "-ffp-contract=off", "-Wall", "-Wextra", "-Wno-unused-parameter",
             f"-I{INCLUDE}", f"-I{os.path.join(CSRC, 'host')}", f"-I{pybind11.get_include()}",
             f"-I{py_inc}"]
    san = SANITIZE_FLAGS if sanitize else []
    objs, jobs = [], []
    for s in srcs:
        o = os.path.join(BUILD, ("san_" if sanitize else "host_") +
                         os.path.basename(s).replace(".cpp", ".o"))
        objs.append(o)
        jobs.append([CXX, *flags, *san, "-c", s, "-o", o])
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        list(ex.map(_run, jobs))
    _run([CXX, "-shared", "-fopenmp", *san, "-o", target, *objs, f"-L{PKG}", "-lgaffa_hip",
          "-Wl,-rpath,$ORIGIN"])
    return target


def build(force: bool = True, sanitize: bool = False) -> None:
    hip_library(force and not sanitize)
    if os.path.isdir(os.path.join(CSRC, "host")):
        core_module(force or sanitize, sanitize)


if __name__ == "__main__":
    build(force="--incremental" not in sys.argv, sanitize="--sanitize" in sys.argv)
    print("built", os.path.join(PKG, "libgaffa_hip.so"))
