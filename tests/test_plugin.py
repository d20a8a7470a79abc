"""The plugin module and the ILinearSolver adapter (the drop-in boundary).

CPU: the module loads through a reproduction of core::PluginLoader's steps
(ABI 1, CreatePlugin, Solver category, no-throw on_load without a device);
when the SimAll sources are present, the adapter also compiles against the
host's own headers and the interface types have identical layouts.
GPU: solves through solver::ILinearSolver, including the reference's CG
known-answer test and its iteration-limit semantics.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "simall_amd")
MODULE = os.path.join(PKG, "simall_hip_solver.so")
CHECK = os.path.join(PKG, "plugin_check")
REF = "/root/reference"


def run_check(mode):
    return subprocess.run([CHECK, MODULE, mode], capture_output=True, text=True, timeout=300)


def test_module_loads_and_reports_no_device():
    import ctypes
    from simall_amd import capi

    cnt = ctypes.c_int()
    capi.load().mi_device_count(ctypes.byref(cnt))
    if cnt.value:
        pytest.skip("a GPU is present")
    r = run_check("load")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "plugin simall_hip_solver 1.0.0 abi 1 category 0" in r.stdout
    assert "registered factory: no" in r.stdout and "not registered" in r.stderr


LAYOUT = r"""
#include <cstddef>
#include <cstdio>
#include "simall_iface.hpp"
using namespace simall;
int main() {
  std::printf("%zu %zu %zu %zu %zu %zu %zu %zu %d %d\n",
    sizeof(solver::CSRMatrix), offsetof(solver::CSRMatrix, colIdx), offsetof(solver::CSRMatrix, values),
    sizeof(solver::LinearSolverConfig), offsetof(solver::LinearSolverConfig, tolerance),
    offsetof(solver::LinearSolverConfig, restart), sizeof(util::aligned_vector<double>),
    sizeof(plugins::CategorisedPlugin), int(solver::PreconditionerKind::AMG), plugins::kPluginAbiVersion);
}
"""


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src", "solver")),
                    reason="SimAll sources not present")
def test_adapter_matches_the_host_headers(tmp_path):
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(PKG, "plugin")]
    host = ["-DSIMALL_HIP_USE_HOST_HEADERS", "-I", os.path.join(REF, "src"), "-I",
            os.path.join(REF, "plugins"), "-I", REF]
    for f in ("HipLinearSolver.cpp", "SimAllHipPlugin.cpp", "plugin_check.cpp"):
        subprocess.run(["g++", "-std=c++20", "-fsyntax-only", *host, *inc,
                        os.path.join(PKG, "plugin", f)], check=True)
    src = tmp_path / "layout.cpp"
    src.write_text(LAYOUT)
    outs = []
    for extra in ([], host):
        exe = tmp_path / ("own" if not extra else "host")
        subprocess.run(["g++", "-std=c++20", *extra, *inc, str(src), "-o", str(exe)], check=True)
        outs.append(subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout)
    assert outs[0] == outs[1]


@pytest.mark.gpu
def test_ilinearsolver_through_the_plugin(gpu):
    r = run_check("solve")
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "registered factory: yes" in r.stdout and r.stdout.strip().endswith("OK")
