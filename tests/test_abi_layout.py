"""CPU check of the C ABI's struct layouts: every ctypes mirror in the
Python host package has the size and the field offsets that the C compiler
gives the struct of the same name in include/fpmhip.h (a mirror that drops,
adds or reorders a field would make the library read or write the wrong
bytes)."""
import ctypes as ct
import os
import subprocess

import pytest

from fpmhip import abi, runtime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

MIRRORS = {
    "fpmhip_frame_schedule": abi.FrameSchedule,
    "fpmhip_optics": abi.Optics,
    "fpmhip_planar_array": abi.PlanarArray,
    "fpmhip_model_desc": abi.ModelDesc,
    "fpmhip_compiled_shape": abi.CompiledShape,
    "fpmhip_algorithm": abi.Algorithm,
    "fpmhip_capabilities": runtime.Capabilities,
    "fpmhip_camera": abi.Camera,
    "fpmhip_intensity_stats": abi.IntensityStats,
    "fpmhip_frame_comparison": abi.FrameComparison,
    "fpmhip_diagnostics": abi.Diagnostics,
    "fpmhip_batch_schedule": runtime.BatchSchedule,
    "fpmhip_batch_tuning": runtime.BatchTuning,
    "fpmhip_server_config": runtime.ServerConfig,
    "fpmhip_server_statistics": runtime.ServerStatistics,
}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    d = tmp_path_factory.mktemp("layout")
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "fpmhip.h"', "int main(void) {"]
    for name, cls in MIRRORS.items():
        lines.append(f'  printf("{name} %zu\\n", sizeof({name}));')
        for field, *_ in cls._fields_:
            lines.append(f'  printf("{name}.{field} %zu\\n", offsetof({name}, {field}));')
    lines += ["  return 0;", "}"]
    src = d / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = d / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return dict((k, int(v)) for k, v in (line.split() for line in out.splitlines()))


@pytest.mark.parametrize("name", sorted(MIRRORS))
def test_ctypes_mirror_matches_c_layout(c_layout, name):
    cls = MIRRORS[name]
    assert ct.sizeof(cls) == c_layout[name]
    for field, *_ in cls._fields_:
        assert getattr(cls, field).offset == c_layout[f"{name}.{field}"], field


def test_every_struct_of_the_header_has_a_mirror():
    import re

    text = open(os.path.join(ROOT, "include", "fpmhip.h")).read()
    declared = set(re.findall(r"typedef struct (fpmhip_[a-z0-9_]+) \{", text))
    assert declared == set(MIRRORS)
