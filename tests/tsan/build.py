"""Builds the host ThreadSanitizer harness of the serving scheduler.

csrc/serve.hip (the scheduler) and server_stub.hip (the HIP calls it makes,
on the host, and a deterministic search) are compiled host-only with
-Xarch_host -fsanitize=thread into tests/tsan/_build/server_stub, which a CPU
test runs.  No device code is built here.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from smartcore_amd import build as B  # noqa: E402

OUT_DIR = os.path.join(HERE, "_build")
OUT = os.path.join(OUT_DIR, "server_stub")
SOURCES = [os.path.join(ROOT, "smartcore_amd", "csrc", "serve.hip"),
           os.path.join(HERE, "server_stub.hip")]


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    flags = ["--cuda-host-only", "-std=c++17", "-O1", "-g", "-fPIC", "-Xarch_host", "-fsanitize=thread",
             f"-I{os.path.join(ROOT, 'smartcore_amd', 'csrc')}", f"-I{os.path.join(ROOT, 'include')}",
             "-Wno-unused-command-line-argument"]
    objs = []
    for src in SOURCES:
        obj = os.path.join(OUT_DIR, os.path.basename(src) + ".o")
        r = subprocess.run([B._hipcc(), *flags, "-c", src, "-o", obj], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
        objs.append(obj)
    r = subprocess.run([B._hipcc(), "-Xarch_host", "-fsanitize=thread", "-o", OUT + ".tmp", *objs],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(OUT + ".tmp", OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build(verbose=True)
