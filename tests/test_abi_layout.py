"""The ctypes mirror of the C ABI (aieplace_amd/api.py) against the headers:
every struct has the C size, and every field the C offset.  A C program
built with gcc from include/*.h prints sizeof and offsetof; no GPU is used."""
import ctypes
import json
import os
import shutil
import subprocess
import tempfile

import pytest

from aieplace_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# ctypes class -> (C type, header)
STRUCTS = {
    "EpNetlist": ("ep_netlist", "eplace.h"),
    "EpParams": ("ep_params", "eplace.h"),
    "EpStepInfo": ("ep_step_info", "eplace.h"),
    "EpRunInfo": ("ep_run_info", "eplace.h"),
    "EpKernelStat": ("ep_kernel_stat", "eplace.h"),
    "EpGenParams": ("ep_gen_params", "eplace_gen.h"),
    "EpDesign": ("ep_design", "eplace_gen.h"),
    "EpMlpTrainParams": ("ep_mlp_train_params", "eplace_mlp.h"),
    "EpBatchRunInfo": ("ep_batch_run_info", "eplace_batch.h"),
    "EpLefdefInfo": ("ep_lefdef_info", "eplace_io.h"),
    "EpLegalInfo": ("ep_legal_info", "eplace_legal.h"),
}


def c_field(name: str) -> str:
    return name[:-1] if name.endswith("_") else name  # lambda_ -> lambda


@pytest.fixture(scope="module")
def c_layout():
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    lines = ["#include <stddef.h>", "#include <stdio.h>"]
    for h in sorted({h for _, h in STRUCTS.values()}):
        lines.append(f'#include "{h}"')
    lines += ["int main(void) {", '    printf("{");']
    first = True
    for cls, (ctype, _) in STRUCTS.items():
        fields = [c_field(f[0]) for f in getattr(api, cls)._fields_]
        parts = [f'\\"size\\": %zu'] + [f'\\"{f}\\": %zu' for f in fields]
        args = [f"sizeof({ctype})"] + [f"offsetof({ctype}, {f})" for f in fields]
        sep = "" if first else ","
        first = False
        lines.append(f'    printf("{sep}\\"{cls}\\": {{{", ".join(parts)}}}", {", ".join(args)});')
    lines += ['    printf("}\\n");', "    return 0;", "}"]
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "layout.c"), os.path.join(d, "layout")
        with open(src, "w") as f:
            f.write("\n".join(lines) + "\n")
        subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return json.loads(out)


@pytest.mark.parametrize("cls", sorted(STRUCTS))
def test_struct_matches_header(c_layout, cls):
    st = getattr(api, cls)
    lay = c_layout[cls]
    assert ctypes.sizeof(st) == lay["size"], f"{cls}: ctypes {ctypes.sizeof(st)} bytes, C {lay['size']}"
    for name, _ in st._fields_:
        assert getattr(st, name).offset == lay[c_field(name)], f"{cls}.{name}"
