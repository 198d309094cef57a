"""The ctypes mirror of bmnd_hip_pipeline_desc (bmnd_amd._lib.PipelineDesc)
has the C header's layout: every field's offset and the size, as gcc lays
the header's struct out."""

import ctypes
import os
import shutil
import subprocess

import pytest

from bmnd_amd._lib import PipelineDesc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_descriptor_layout_matches_the_header(tmp_path):
    fields = [name for name, _ in PipelineDesc._fields_]
    src = tmp_path / "layout.c"
    body = "".join(f'  printf("{f} %zu\\n", offsetof(bmnd_hip_pipeline_desc, {f}));\n'
                   for f in fields)
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n"
                   f'#include "{os.path.join(ROOT, "include", "bmnd_hip.h")}"\n'
                   "int main(void) {\n"
                   '  printf("size %zu\\n", sizeof(bmnd_hip_pipeline_desc));\n'
                   f"{body}  return 0;\n}}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True,
                                                        text=True, check=True).stdout.splitlines())
    assert int(out["size"]) == ctypes.sizeof(PipelineDesc)
    for f in fields:
        assert getattr(PipelineDesc, f).offset == int(out[f]), f
