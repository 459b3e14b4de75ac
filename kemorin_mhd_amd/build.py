"""Build libkemorin_sht.so (gfx950) in-tree with hipcc.

Usage: python -m kemorin_mhd_amd.build   (or build() from Python)."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libkemorin_sht.so")
SOURCES = ["sht_plan.hip", "sht_legendre.hip", "sht_fft.hip", "sht_reorder.hip",
           "sht_split.hip", "sht_capi.hip", "sht_batch.hip", "sht_pole.hip",
           "sht_nonlinear.hip", "sht_copy.hip", "sht_coriolis.hip", "sht_tune.hip", "sht_dft.hip",
           "sht_files.hip"]
ARCH = "gfx950"
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-fno-gpu-rdc",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-Wno-pass-failed"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


HEADERS = [os.path.join(CSRC, "sht_internal.h"),
           os.path.join(os.path.dirname(HERE), "include", "kemorin_sht.h")]


def _stale(src: str, obj: str) -> bool:
    """An object older than its source or a shared header."""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(f) > t for f in [os.path.join(CSRC, src), *HEADERS])


def build(verbose: bool = False, incremental: bool = False) -> str:
    """Compile every source (incremental: only those whose object is older
    than the source or a shared header) and link the library."""
    os.makedirs(OBJ, exist_ok=True)
    cc = hipcc()
    todo = [s for s in SOURCES
            if not incremental or _stale(s, os.path.join(OBJ, s.replace(".hip", ".o")))]

    def one(src):
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        if src not in todo:
            return obj
        cmd = [cc, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
        if verbose and r.stderr.strip():
            print(r.stderr)
        return obj

    def waits():
        # the split GEMMs count their own memory waits, and they and the
        # split-form FFTs issue LDS DMA as inline assembly: verify the
        # generated code keeps every in-flight load's registers untouched and
        # the wait states the assembly needs, on all paths
        from .isa_waits import check_source
        for src, pattern in (("sht_split.hip", "k_leg_"), ("sht_fft.hip", "k_fft3_")):
            if src not in todo:
                continue
            probs = check_source(cc, FLAGS, os.path.join(CSRC, src), pattern)
            bad = {k: v for k, v in probs.items() if v}
            if not probs or bad:
                raise RuntimeError(f"{src} wait check failed:\n" +
                                   "\n".join(f"{k}: {v[:5]}" for k, v in bad.items()))

    with ThreadPoolExecutor(max_workers=min(7, os.cpu_count() or 1)) as ex:
        fut = ex.submit(waits)
        objs = list(ex.map(one, SOURCES))
        fut.result()
    tmp = LIB + ".tmp"
    r = subprocess.run([cc, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs,
                        "-L/opt/rocm/lib", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="--incremental" in sys.argv))
