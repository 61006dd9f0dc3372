"""Build the HIP library libcuckoo_amd.so for gfx950 (in-tree).

    python cuckoo-filter_amd/build.py [--force] [--jobs N]

Compiles csrc/capi_inst.hip once per (policy, key width), csrc/capi.hip, and
links them into cuckoo-filter_amd/libcuckoo_amd.so.  Without --force, objects
newer than every source are reused.
"""
from __future__ import annotations

import argparse
import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcuckoo_amd.so")
RLIB = os.path.join(HERE, "libcuckoo_roofline.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CUCKOO_OFFLOAD_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC", "-Wall", "-Wno-unused-function",
    # keep the compiler from forming packed-FP32 instructions (none are wanted here)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-DCUCKOO_FILTER_NO_THRUST", f"-I{CSRC}",
]

UNITS = [("capi", "capi.hip", []), ("server", "server.hip", []), ("kmer", "kmer.hip", [])] + [
    (f"inst_p{p}_k{k}", "capi_inst.hip", [f"-DCF_INST_POLICY={p}", f"-DCF_INST_KEY={k}"])
    for p in (0, 1, 2) for k in (4, 8)
]


def _sources_mtime() -> float:
    files = glob.glob(os.path.join(CSRC, "**", "*.h*"), recursive=True)
    files += glob.glob(os.path.join(CSRC, "*.hip"))
    files += [os.path.join(HERE, "..", "include", "cuckoo_filter.h"), os.path.abspath(__file__)]
    return max(os.path.getmtime(f) for f in files)


def _compile(unit, force, newest):
    name, src, defs = unit
    obj = os.path.join(OBJ, name + ".o")
    if not force and os.path.exists(obj) and os.path.getmtime(obj) > newest:
        return obj, None
    cmd = [HIPCC, *COMMON, *defs, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return obj, f"$ {' '.join(cmd)}\n{r.stdout}\n{r.stderr}"
    return obj, None


def build(force: bool = False, jobs: int | None = None) -> str:
    os.makedirs(OBJ, exist_ok=True)
    newest = _sources_mtime()
    jobs = jobs or min(len(UNITS), os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda u: _compile(u, force, newest), UNITS))
    errors = [e for _, e in results if e]
    if errors:
        raise RuntimeError("HIP build failed:\n" + "\n".join(errors))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *[o for o, _ in results]]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n$ {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # measurement kernels (random-gather / CAS / copy roofline), a separate library
    robj = os.path.join(OBJ, "roofline.o")
    if force or not os.path.exists(RLIB) or os.path.getmtime(RLIB) <= newest:
        cmd = [HIPCC, *COMMON, "-shared", "-o", RLIB + ".tmp", os.path.join(CSRC, "roofline.hip")]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"roofline build failed:\n$ {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(RLIB + ".tmp", RLIB)
    del robj
    # C++ template-API test program (tests/cpp), run by the GPU test suite
    src = os.path.join(HERE, "..", "tests", "cpp", "test_cuckoo_filter.cpp")
    exe = os.path.splitext(src)[0]
    if os.path.exists(src) and (force or not os.path.exists(exe) or os.path.getmtime(exe) <= max(newest, os.path.getmtime(src))):
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-std=c++20", "-O2", "-Xclang", "-target-feature", "-Xclang",
               "-packed-fp32-ops", f"-I{CSRC}", src, "-o", exe + ".tmp"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"C++ API test build failed:\n$ {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(exe + ".tmp", exe)
    return LIB


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args()
    print(build(force=a.force, jobs=a.jobs))
    return 0


if __name__ == "__main__":
    sys.exit(main())
