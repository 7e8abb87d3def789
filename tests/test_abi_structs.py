"""The ctypes mirrors of the C ABI's structs (padenti_amd/_lib.py) have the
size and field offsets of the structs in include/padenti_hip.h, checked
against a C program compiled here: a mirror that is shorter than its C
struct lets the library write past a Python-allocated array (per-level
statistics are written as an array of structs)."""
import ctypes
import os
import shutil
import subprocess

import pytest

from padenti_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PAIRS = [("pd_train_params", L.TrainParams), ("pd_level_stats", L.LevelStats), ("pd_tuning", L.Tuning),
         ("pd_tuning_rule", L.TuningRule), ("pd_adam", L.Adam), ("pd_sampler", L.Sampler),
         ("pd_ingest_stats", L.IngestStats), ("pd_png_info", L.PngInfo)]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs a C compiler")
def test_struct_mirrors_match_the_header(tmp_path):
    lines = ['#include "padenti_hip.h"', "#include <stdio.h>", "#include <stddef.h>", "int main(void) {"]
    for cname, py in PAIRS:
        lines.append(f'  printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "abi.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "abi"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                  text=True).stdout.splitlines())
    for cname, py in PAIRS:
        assert int(got[cname]) == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(py, fname).offset, f"{cname}.{fname}"
