"""The C-ABI library loads and exports every entry point include/*.h declares
(CPU only: no compute calls)."""
import ctypes
import os
import re
import subprocess

import pytest

from simall_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"MI_API\s+[\w\s\*]+?\b(mi_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    names = declared_symbols()
    assert len(names) >= 40
    for n in sorted(names):
        assert hasattr(lib, n), f"{n} declared but not exported"
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (mi_\w+)", out))
    assert names == exported, f"undeclared exports: {exported - names}; missing: {names - exported}"


def test_python_prototypes_cover_the_header():
    assert declared_symbols() == set(capi.PROTOTYPES)


def test_abi_version_and_struct_sizes():
    lib = capi.load()
    assert lib.mi_get_abi_version() == 1
    cfg = capi.default_cfg()
    assert cfg.kind == capi.MI_SOLVER_CG and cfg.tolerance == 1e-8 and cfg.max_iterations == 1000
    assert cfg.restart == 30



def test_struct_layouts_match_the_c_header(tmp_path):
    # ctypes mirrors must have the C compiler's layout
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "simall_amd.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(mi_solver_cfg), '
                   'sizeof(mi_stats), sizeof(mi_mesh_soa), offsetof(mi_solver_cfg, use_graph), '
                   'offsetof(mi_stats, solve_seconds), sizeof(mi_flow_cfg), '
                   'offsetof(mi_flow_cfg, pressure), sizeof(mi_flow_residuals));return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]
    assert got == [ctypes.sizeof(capi.SolverCfg), ctypes.sizeof(capi.Stats),
                   ctypes.sizeof(capi.MeshSoa), capi.SolverCfg.use_graph.offset,
                   capi.Stats.solve_seconds.offset, ctypes.sizeof(capi.FlowCfg),
                   capi.FlowCfg.pressure.offset, ctypes.sizeof(capi.FlowResiduals)]


def test_no_device_is_reported_not_faked():
    lib = capi.load()
    cnt = ctypes.c_int()
    lib.mi_device_count(ctypes.byref(cnt))
    if cnt.value > 0:
        pytest.skip("a GPU is present")
    assert lib.mi_init(0) == capi.MI_ERR_NO_DEVICE
    assert b"no HIP device" in lib.mi_last_error() or b"failed" in lib.mi_last_error()
    p = ctypes.c_void_p()
    assert lib.mi_csr_create(1, 1, None, None, None, ctypes.byref(p)) != capi.MI_OK
