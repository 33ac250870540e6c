"""The ctypes mirrors of the ABI structs match the C headers field by field.

A C program compiled from include/*.h prints sizeof and offsetof of every field;
each ctypes Structure in clmsoil._abi must agree. A field added, dropped or
reordered on one side only would otherwise pass pointers into the wrong slots."""
import shutil
import subprocess
from pathlib import Path

import pytest

from clmsoil import _abi

REPO = Path(__file__).resolve().parent.parent
STRUCTS = {
    "clmsoil_levels": _abi.Levels,
    "clmsoil_swm_config": _abi.SwmConfig,
    "clmsoil_mf_args": _abi.MfArgs,
    "clmsoil_heat_args": _abi.HeatArgs,
    "clmsoil_heat_config": _abi.HeatConfig,
    "clmsoil_zd_args": _abi.ZdArgs,
    "clmsoil_ghf_args": _abi.GhfArgs,
    "clmsoil_seb_args": _abi.SebArgs,
    "clmsoil_fractions_args": _abi.FractionsArgs,
    "clmsoil_water_table_args": _abi.WaterTableArgs,
    "clmsoil_theta_wt_args": _abi.ThetaWtArgs,
    "clmsoil_surfhyd_args": _abi.SurfHydArgs,
    "clmsoil_surfhyd_params": _abi.SurfHydParams,
    "clmsoil_rootsink_args": _abi.RootSinkArgs,
    "clmsoil_lake_solve_args": _abi.LakeSolveArgs,
    "clmsoil_lvt_args": _abi.LvtArgs,
    "clmsoil_lvt_params": _abi.LvtParams,
    "clmsoil_ch4_tran_args": _abi.Ch4TranArgs,
    "clmsoil_ch4_tran_params": _abi.Ch4TranParams,
}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = ["#include <stdio.h>", "#include <stddef.h>",
             '#include "clmsoil.h"', '#include "clmsoil_zd09.h"', '#include "clmsoil_ghf.h"',
             '#include "clmsoil_hydro.h"', '#include "clmsoil_surfhyd.h"', '#include "clmsoil_lake.h"', '#include "clmsoil_lvt.h"', '#include "clmsoil_ch4.h"',
             "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["return 0;", "}"]
    d = tmp_path_factory.mktemp("abi")
    src, exe = d / "layout.c", d / "layout"
    src.write_text("\n".join(lines) + "\n")
    subprocess.run([cc, "-std=c11", "-I", str(REPO / "include"), str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    table = {}
    for ln in out.splitlines():
        s, f, v = ln.split()
        table[(s, f)] = int(v)
    return table


@pytest.mark.parametrize("cname", list(STRUCTS))
def test_struct_matches_header(c_layout, cname):
    py = STRUCTS[cname]
    import ctypes as C
    assert C.sizeof(py) == c_layout[(cname, "sizeof")], cname
    for f, _ in py._fields_:
        assert getattr(py, f).offset == c_layout[(cname, f)], (cname, f)
