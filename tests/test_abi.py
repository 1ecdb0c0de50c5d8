"""The C ABI library loads and exports every entry point include/abx.h
declares, and the ctypes record layouts match the C structs.  No compute
calls (this runs without a GPU)."""
import ctypes as C
import os
import re
import subprocess

from abeille_amd import abi
import abeille_amd

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "abx.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(abx_\w+)\s*\(", text, re.M)))


def test_all_declared_symbols_exported():
    names = declared_functions()
    assert len(names) >= 15
    L = C.CDLL(abeille_amd.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_version_string():
    assert b"gfx950" in abeille_amd.lib().abx_version()


STRUCTS = {"abx_surface": abi.Surface, "abx_cell": abi.Cell, "abx_universe": abi.Universe,
           "abx_material": abi.Material, "abx_source": abi.Source,
           "abx_mesh_tally": abi.MeshTally, "abx_settings": abi.Settings,
           "abx_problem_desc": abi.ProblemDesc, "abx_particle": abi.Particle,
           "abx_banked": abi.Banked, "abx_scores": abi.Scores, "abx_diag": abi.Diag,
           "abx_run_stats": abi.RunStats, "abx_trace_rec": abi.TraceRec,
           "abx_cancel_mesh": abi.CancelMesh, "abx_tuning": abi.Tuning}


def test_record_layouts_match_c(tmp_path):
    """sizeof and every field offset of the ctypes mirrors equal the C
    compiler's."""
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(){"]
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    got = dict(line.split() for line in out.splitlines())
    for cname, py in STRUCTS.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)
