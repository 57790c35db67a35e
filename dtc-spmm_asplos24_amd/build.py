"""Build libdtc.so (hand-written HIP for gfx950) and the DTCSpMM extension.

    python build.py            # rebuild what is out of date
    python build.py --force    # rebuild everything

Outputs land next to this file: libdtc.so and DTCSpMM.<abi>.so.  The
extension links libdtc.so through an $ORIGIN rpath, and both resolve the HIP
runtime that torch ships (same SONAME as ROCm's), so one HIP runtime serves
the whole process.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
INCLUDE = os.path.join(REPO, "include")
HIP_SOURCES = ["primitives.hip", "metcf.hip", "spmm.hip", "backward.hip", "selector.hip",
               "reorder.hip", "gcn.hip"]
HEADERS = [os.path.join(CSRC, "dtc_internal.h"), os.path.join(CSRC, "seg_sort.h"),
           os.path.join(INCLUDE, "dtc.h")]
LIB = os.path.join(HERE, "libdtc.so")
# measurement aid beside the library (not part of include/dtc.h)
PROBE_SRC = "probe_gather.hip"
PROBE_LIB = os.path.join(HERE, "libdtc_probe.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("DTC_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-mcode-object-version=5",
    # Packed FP32 (v_pk_*_f32) stays off: the compiler must not emit it.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", f"-I{INCLUDE}", f"-I{CSRC}",
]
# Per-source extra flags.  The ME-TCF window kernels keep 16 entries per lane
# in register arrays updated under per-slot branches; turned into vectors by
# the backend's alloca promotion, every branch copies the whole array (234
# VGPRs and 580 bytes of scratch per lane instead of 120 and none).
FILE_FLAGS = {"metcf.hip": ["-mllvm", "-disable-promote-alloca-to-vector"]}


def _torch_paths():
    import torch
    import torch.utils.cpp_extension as ce
    lib = os.path.join(os.path.dirname(torch.__file__), "lib")
    return ce.include_paths(), lib, int(torch._C._GLIBCXX_USE_CXX11_ABI)


def ext_path() -> str:
    return os.path.join(HERE, "DTCSpMM" + sysconfig.get_config_var("EXT_SUFFIX"))


def _stale(target: str, deps: list[str]) -> bool:
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps if os.path.exists(d))


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in (r.stderr or "").splitlines() if noise not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)


def build_lib(force: bool = False, verbose: bool = False, jobs: int = 4) -> str:
    os.makedirs(OBJ, exist_ok=True)
    objs, todo = [], []
    for src in HIP_SOURCES:
        s = os.path.join(CSRC, src)
        o = os.path.join(OBJ, src.replace(".hip", ".o"))
        objs.append(o)
        if force or _stale(o, [s] + HEADERS):
            todo.append([HIPCC, *HIP_FLAGS, *FILE_FLAGS.get(src, []), "-c", s, "-o", o])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for f in [ex.submit(_run, c, verbose) for c in todo]:
            f.result()
    if force or todo or _stale(LIB, objs):
        _, torch_lib, _ = _torch_paths()
        rpath = f"-Wl,-rpath,{torch_lib}:{ROCM}/lib"
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", LIB + ".tmp",
              rpath, "-Wl,-soname,libdtc.so"], verbose)
        os.replace(LIB + ".tmp", LIB)
    return LIB


def build_probe(force: bool = False, verbose: bool = False) -> str:
    src = os.path.join(CSRC, PROBE_SRC)
    if force or _stale(PROBE_LIB, [src]):
        _, torch_lib, _ = _torch_paths()
        _run([HIPCC, *HIP_FLAGS, "-shared", src, "-o", PROBE_LIB + ".tmp",
              f"-Wl,-rpath,{torch_lib}:{ROCM}/lib"], verbose)
        os.replace(PROBE_LIB + ".tmp", PROBE_LIB)
    return PROBE_LIB


def build_ext(force: bool = False, verbose: bool = False) -> str:
    out = ext_path()
    src = os.path.join(CSRC, "DTCSpMM_ext.cpp")
    if not (force or _stale(out, [src, LIB] + HEADERS)):
        return out
    incs, torch_lib, abi = _torch_paths()
    py_inc = sysconfig.get_paths()["include"]
    cmd = ["g++", "-O2", "-fPIC", "-shared", "-std=c++17", f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
           "-DUSE_ROCM=1", "-D__HIP_PLATFORM_AMD__=1", "-DTORCH_EXTENSION_NAME=DTCSpMM",
           "-DTORCH_API_INCLUDE_EXTENSION_H", "-Wno-deprecated-declarations",
           *[f"-I{p}" for p in incs], f"-I{py_inc}", f"-I{INCLUDE}", f"-I{ROCM}/include",
           src, "-o", out + ".tmp", f"-L{HERE}", "-ldtc", f"-L{torch_lib}", "-lc10", "-lc10_hip",
           "-ltorch", "-ltorch_cpu", "-ltorch_hip", "-ltorch_python", "-lamdhip64",
           f"-Wl,-rpath,$ORIGIN:{torch_lib}:{ROCM}/lib"]
    _run(cmd, verbose)
    os.replace(out + ".tmp", out)
    return out


# Host code of the library under AddressSanitizer and UndefinedBehaviorSanitizer
# (device code is compiled as usual: -Xarch_host keeps the sanitizers off the
# gfx950 pass).  The sanitizer runtime is not linked in; the process preloads
# ROCm clang's ASan runtime, as tests/run_sanitizers.sh does.  CPU use only.
SAN_DIR = os.path.join(HERE, "_build_san")
SAN_LIB = os.path.join(SAN_DIR, "libdtc.so")
SAN_FLAGS = [
    "-O1", "-g", "-fno-omit-frame-pointer",
    "-Xarch_host", "-fsanitize=address,undefined",
    "-Xarch_host", "-fno-sanitize=vptr",
    "-Xarch_host", "-fno-sanitize-recover=undefined",
]


def build_sanitized(verbose: bool = False, jobs: int = 4) -> str:
    os.makedirs(SAN_DIR, exist_ok=True)
    flags = [f for f in HIP_FLAGS if f != "-O3"] + SAN_FLAGS
    objs, todo = [], []
    for src in HIP_SOURCES:
        s = os.path.join(CSRC, src)
        o = os.path.join(SAN_DIR, src.replace(".hip", ".o"))
        objs.append(o)
        if _stale(o, [s] + HEADERS):
            todo.append([HIPCC, *flags, *FILE_FLAGS.get(src, []), "-c", s, "-o", o])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for f in [ex.submit(_run, c, verbose) for c in todo]:
            f.result()
    _, torch_lib, _ = _torch_paths()
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", SAN_LIB + ".tmp",
          f"-Wl,-rpath,{torch_lib}:{ROCM}/lib", "-Wl,-soname,libdtc.so"], verbose)
    os.replace(SAN_LIB + ".tmp", SAN_LIB)
    return SAN_LIB


def build(force: bool = False, verbose: bool = False) -> None:
    build_lib(force=force, verbose=verbose)
    build_ext(force=force, verbose=verbose)
    build_probe(force=force, verbose=verbose)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    ap.add_argument("--lib-only", action="store_true")
    ap.add_argument("--sanitize", action="store_true",
                    help="build only the ASan/UBSan host build of the library (_build_san/libdtc.so)")
    a = ap.parse_args()
    if a.sanitize:
        print("built", build_sanitized(verbose=a.verbose))
        sys.exit(0)
    build_lib(force=a.force, verbose=a.verbose)
    if not a.lib_only:
        build_ext(force=a.force, verbose=a.verbose)
        build_probe(force=a.force, verbose=a.verbose)
    print("built", LIB, "" if a.lib_only else f"{ext_path()} {PROBE_LIB}")
