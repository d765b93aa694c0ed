"""The ctypes mirror of the C ABI against the header itself: a small C
program built with gcc from include/nraps.h prints sizeof and offsetof of
every field of every public struct, and each must equal the ctypes layout
in nraps_amd/_abi.py (a field added on one side only, or a reordered or
resized one, fails here before any GPU call could misread it)."""
import os
import shutil
import subprocess

import pytest

from nraps_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STRUCTS = {
    "nraps_problem": _abi.Problem,
    "nraps_run_opts": _abi.RunOpts,
    "nraps_trace_event": _abi.TraceEvent,
    "nraps_result": _abi.Result,
    "nraps_deck_info": _abi.DeckInfo,
    "nraps_server_opts": _abi.ServerOpts,
    "nraps_diffusion_opts": _abi.DiffusionOpts,
    "nraps_diffusion_result": _abi.DiffusionResult,
}


def c_layouts(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "nraps.h"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in cls._fields_:
            lines.append(f'  printf("{cname} {f[0]} %zu %zu\\n", offsetof({cname}, {f[0]}), '
                         f'sizeof((({cname} *)0)->{f[0]}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        parts = line.split()
        if parts[1] == "size":
            got[(parts[0], None)] = int(parts[2])
        else:
            got[(parts[0], parts[1])] = (int(parts[2]), int(parts[3]))
    return got


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_ctypes_structs_match_the_header(tmp_path):
    got = c_layouts(tmp_path)
    for cname, cls in STRUCTS.items():
        assert got[(cname, None)] == __import__("ctypes").sizeof(cls), cname
        for f in cls._fields_:
            d = getattr(cls, f[0])
            assert got[(cname, f[0])] == (d.offset, d.size), f"{cname}.{f[0]}"
