"""Build libbmnd_hip.so in-tree with hipcc for gfx950 (sources compiled in
parallel, then linked)."""

from __future__ import annotations

import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
SRCS = [os.path.join(HERE, "csrc", "bmnd_hip.hip"), os.path.join(HERE, "csrc", "bmnd_sigma.hip"),
        os.path.join(HERE, "csrc", "bmnd_conv.hip"), os.path.join(HERE, "csrc", "bmnd_schedule.hip"),
        os.path.join(HERE, "csrc", "bmnd_mad.hip")]
OUT = os.path.join(HERE, "libbmnd_hip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-mcode-object-version=5",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    with tempfile.TemporaryDirectory(prefix="bmnd_build_") as tmpdir:
        objs, procs = [], []
        for src in SRCS:
            obj = os.path.join(tmpdir, os.path.basename(src) + ".o")
            cmd = [hipcc, *HIPCC_FLAGS, "-c", "-o", obj, src]
            procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                stderr=subprocess.PIPE, text=True)))
            objs.append(obj)
        for cmd, proc in procs:
            _, err = proc.communicate()
            if proc.returncode != 0:
                raise RuntimeError(f"hipcc failed:\n{' '.join(cmd)}\n{err}")
            if verbose and err:
                print(err)
        tmp = OUT + ".tmp"
        cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"hipcc link failed:\n{' '.join(cmd)}\n{res.stderr}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
