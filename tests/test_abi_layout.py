"""The ctypes mirrors in angstromcube_amd/gfa.py against the C structs of include/gfa.h: a small C program built with
gcc prints sizeof and offsetof of every field as the C compiler lays them out, and each mirror must match field by
field (name, offset) and in size.  A field added, renamed or reordered on one side only fails here, on the CPU."""
import os
import shutil
import subprocess

import ctypes as C
import pytest

from angstromcube_amd import gfa as G

HERE = os.path.dirname(os.path.abspath(__file__))
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

MIRRORS = {
    "gfa_plan_desc": G.PlanDesc,
    "gfa_sizes": G.Sizes,
    "gfa_solve_opts": G.SolveOpts,
    "gfa_solve_stats": G.SolveStats,
    "gfa_chunked_stats": G.ChunkedStats,
    "gfa_contour_opts": G.ContourOpts,
    "gfa_contour_stats": G.ContourStats,
}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc is not available")
    d = tmp_path_factory.mktemp("abi")
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "gfa.h"', "int main(void) {"]
    for cname, mirror in MIRRORS.items():
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in mirror._fields_:
            lines.append(f'    printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["    return 0;", "}"]
    src = d / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = d / "layout"
    r = subprocess.run(["gcc", "-std=c11", "-I", INCLUDE, "-o", str(exe), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr   # a field the mirror names but the header lacks fails to compile
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    layout = {}
    for line in out.splitlines():
        cname, key, value = line.split()
        layout[(cname, key)] = int(value)
    return layout


@pytest.mark.parametrize("cname", sorted(MIRRORS))
def test_ctypes_mirror_matches_the_header(c_layout, cname):
    mirror = MIRRORS[cname]
    assert C.sizeof(mirror) == c_layout[(cname, "size")], f"{cname}: size differs"
    for fname, _ in mirror._fields_:
        assert getattr(mirror, fname).offset == c_layout[(cname, fname)], f"{cname}.{fname}: offset differs"


def test_header_fields_are_all_mirrored():
    """Every field the header declares in these structs appears in the mirror (no field left out at the end)."""
    import re
    text = open(os.path.join(INCLUDE, "gfa.h")).read()
    for cname, mirror in MIRRORS.items():
        end = re.search(r"\}\s*" + cname + r";", text).start()
        body = text[text.rindex("typedef struct {", 0, end) + len("typedef struct {"):end]
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        declared = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            decl = re.sub(r"\[[^\]]*\]", "", decl)     # array extents
            for part in decl.split(","):               # several declarators of one type
                declared.append(re.search(r"(\w+)\s*$", part).group(1))
        assert declared == [f for f, _ in mirror._fields_], cname
