"""Build libqms's own HIP sources for the CPU (TEST INFRASTRUCTURE ONLY).

The kernels and the C ABI are compiled with g++ against the emulation header
in ``tests/cpuemu/include`` (every ``<hip/hip_runtime.h>`` include resolves
there), with AddressSanitizer and UBSan, into
``tests/cpuemu/_build/libqms_cpu.so``.  The CPU tests load that library in a
child process (AddressSanitizer's runtime preloaded) and run the same parity
bodies as the GPU tests, so an out-of-bounds index or a missing barrier shows
up on the CPU first.  Usage: ``python tests/cpuemu/build.py [-v]``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "alpineqms_amd", "csrc")
OUTDIR = os.path.join(HERE, "_build")
OUT = os.path.join(OUTDIR, "libqms_cpu.so")
sys.path.insert(0, ROOT)
from alpineqms_amd.build import SOURCES  # noqa: E402  (the same source list as the GPU build)

SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]


def gxx() -> str:
    c = shutil.which("g++")
    if not c:
        raise RuntimeError("g++ not found")
    return c


def flags(sanitize: bool = True):
    f = ["-x", "c++", "-std=c++20", "-O1", "-g", "-fPIC", "-pthread", "-Wno-attributes", "-Wno-unknown-pragmas",
         "-Wno-ignored-attributes", "-I", os.path.join(HERE, "include"), "-I", os.path.join(ROOT, "include"),
         "-I", CSRC]
    return f + (SAN if sanitize else [])


def build(verbose: bool = False, sanitize: bool = True) -> str:
    os.makedirs(OUTDIR, exist_ok=True)
    cc = gxx()
    objs, procs = [], []
    for src in SOURCES:
        obj = os.path.join(OUTDIR, src.replace(".hip", ".o"))
        cmd = [cc, *flags(sanitize), "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    sw = os.path.join(OUTDIR, "qemu_switch.o")
    procs.append(("qemu_switch.S", subprocess.Popen([cc, "-c", os.path.join(HERE, "qemu_switch.S"), "-o", sw],
                                                    stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    objs.append(sw)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((src, out.decode(errors="replace")))
    if failed:
        raise RuntimeError("g++ failed:\n" + "\n\n".join(f"[{s}]\n{t}" for s, t in failed))
    tmp = OUT + ".tmp"
    cmd = [cc, "-shared", "-pthread", *(SAN if sanitize else []), "-o", tmp, *objs]
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


def asan_runtime() -> str:
    """Path of g++'s AddressSanitizer runtime (to preload into python)."""
    return subprocess.run([gxx(), "-print-file-name=libasan.so"], check=True, capture_output=True,
                          text=True).stdout.strip()


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
