"""The Python mirrors of the C-ABI structs (ctypes structures, the numpy
record dtype of the trace) have exactly the layout include/pcc.h gives them:
a C program built against the header reports sizeof and offsetof of every
field, which must equal the mirror's.  A field added to one side only, or in
a different position, fails here (no GPU needed)."""
import os
import shutil
import subprocess

import pytest

from pcc import abi_types as T

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CTYPES = {
    "pcc_params": T.Params,
    "pcc_bodies": T.BodiesC,
    "pcc_constraints": T.ConstraintsC,
    "pcc_step_result": T.StepResult,
}
# C member names that are Python keywords get another name in the mirror
C_NAME = {("pcc_params", "lam"): "lambda"}


def mirror_layouts():
    out = {}
    for cname, st in CTYPES.items():
        fields = [(f[0], getattr(st, f[0]).offset) for f in st._fields_]
        out[cname] = (C_sizeof(st), fields)
    rec = T.BBPGD_RECORD
    out["pcc_bbpgd_record"] = (rec.itemsize, [(n, rec.fields[n][1]) for n in rec.names])
    return out


def C_sizeof(st):
    import ctypes
    return ctypes.sizeof(st)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no C compiler")
def test_python_mirrors_match_the_header_layout(tmp_path):
    want = mirror_layouts()
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "pcc.h"', "int main(void) {"]
    for cname, (_, fields) in want.items():
        lines.append(f'  printf("%s %zu\\n", "{cname}", sizeof({cname}));')
        for f, _ in fields:
            member = C_NAME.get((cname, f), f)
            lines.append(f'  printf("%s.%s %zu\\n", "{cname}", "{f}", offsetof({cname}, {member}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(REPO, "include"), str(src), "-o",
                    str(exe)], check=True)
    got = dict(line.split() for line in
               subprocess.run([str(exe)], check=True, capture_output=True,
                              text=True).stdout.splitlines())
    for cname, (size, fields) in want.items():
        assert int(got[cname]) == size, cname
        for f, off in fields:
            assert int(got[f"{cname}.{f}"]) == off, (cname, f)
