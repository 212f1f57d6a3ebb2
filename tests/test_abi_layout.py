"""The Python ctypes mirrors of the C ABI's structs have the header's layout: a C program compiled with gcc against
include/olfhip.h prints every struct's size and the offset of each field the mirror names, and they must equal the
mirror's (a misnamed field does not compile). CPU only."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from olfgpu import _native as N

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MIRRORS = {
    "olfhip_reactive_limits": N.ReactiveLimits,
    "olfhip_tuning": N.Tuning,
    "olfhip_distributed_slack": N.DistributedSlackDesc,
    "olfhip_solve_inputs": N.SolveInputs,
    "olfhip_solve_outputs": N.SolveOutputs,
    "olfhip_limit_table_desc": N.LimitTableDesc,
}


def c_layout(tmp_path):
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "olfhip.h"', 'int main(void) {']
    for cname, mirror in MIRRORS.items():
        lines.append(f'    printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in mirror._fields_:
            lines.append(f'    printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append('    printf("olfhip_violation %zu\\n", sizeof(olfhip_violation));')
    for fname in N.VIOLATION_DTYPE.names:
        lines.append(f'    printf("olfhip_violation.{fname} %zu\\n", offsetof(olfhip_violation, {fname}));')
    lines += ['    return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True,
                   capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {k: int(v) for k, v in (line.split() for line in out.splitlines())}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_ctypes_mirrors_match_the_header(tmp_path):
    lay = c_layout(tmp_path)
    for cname, mirror in MIRRORS.items():
        assert lay[cname] == C.sizeof(mirror), cname
        for fname, _ in mirror._fields_:
            assert lay[f"{cname}.{fname}"] == getattr(mirror, fname).offset, f"{cname}.{fname}"
    assert lay["olfhip_violation"] == N.VIOLATION_DTYPE.itemsize
    for fname in N.VIOLATION_DTYPE.names:
        assert lay[f"olfhip_violation.{fname}"] == N.VIOLATION_DTYPE.fields[fname][1], fname
