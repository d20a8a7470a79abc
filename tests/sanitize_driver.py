"""Exercises the host code of the library (checkpoint reader and writer,
SA-AMG setup, RCM) built with AddressSanitizer and UBSan
(simall_amd/_build/asan/libsimall_host_asan.so), on valid inputs and on
corrupted checkpoint files and malformed CSR input.  Run by
tests/test_host_sanitizers.py in a child process with the sanitizer runtime
preloaded; any invalid memory access or undefined behaviour aborts it."""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from simall_amd import capi  # noqa: E402
from simall_amd import workload as wl  # noqa: E402


def bind(path):
    # the solver defaults come from the ordinary library (device code is not
    # part of the sanitized build); nothing else is called there
    full = C.CDLL(capi.LIB_PATH)
    full.mi_solver_cfg_default.argtypes = [C.POINTER(capi.SolverCfg)]
    full.mi_solver_cfg_default.restype = None

    def default_cfg(**kw):
        cfg = capi.SolverCfg()
        full.mi_solver_cfg_default(C.byref(cfg))
        for k, v in kw.items():
            setattr(cfg, k, v)
        return cfg

    capi.default_cfg = default_cfg
    lib = C.CDLL(path)
    for name, (res, args) in capi.PROTOTYPES.items():
        fn = getattr(lib, name, None)
        if fn is not None:
            fn.restype = res
            fn.argtypes = args
    capi._lib = lib  # the capi helpers now call the sanitized library
    return lib


def random_spd(n, seed):
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for i in range(n):
        for j in set(rng.integers(0, n, 3).tolist()) - {i}:
            w = rng.uniform(0.1, 2.0)
            rows += [i, j]
            cols += [j, i]
            vals += [-w, -w]
    import scipy.sparse as sp
    A = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()
    A.sum_duplicates()
    d = -np.asarray(A.sum(axis=1)).ravel() + 0.1
    A = (A + sp.diags(d)).tocsr()
    A.sort_indices()
    return A


def main(path, workdir):
    lib = bind(path)
    # SA-AMG setup and level queries on random SPD systems of several sizes
    for n, seed in ((1, 1), (7, 2), (300, 3), (5000, 4)):
        A = random_spd(n, seed)
        cfg = capi.default_cfg()
        cfg.amg_coarse_size = 16
        H = capi.AmgHost(A.indptr, A.indices, A.data, cfg)
        for lev in range(H.levels):
            for which in (H.A, H.P, H.R):
                if lev + 1 == H.levels and which != H.A:
                    continue
                H.scipy(lev, which)
        capi.rcm_permutation(A.indptr, A.indices)
    # malformed CSR input is refused
    bad = [(np.array([0, 2, 1], np.int32), np.array([0, 1], np.int32)),
           (np.array([0, 1, 2], np.int32), np.array([0, 5], np.int32)),
           (np.array([1, 1, 2], np.int32), np.array([0, 1], np.int32)),
           (np.array([0, 1, 2], np.int32), np.array([-1, 1], np.int32))]
    for rp, ci in bad:
        perm = np.empty(rp.size - 1, np.int32)
        bw = C.c_int32()
        rc = lib.mi_rcm_permutation(rp.size - 1, rp.ctypes.data, ci.ctypes.data, perm.ctypes.data,
                                    C.byref(bw))
        assert rc == capi.MI_ERR_BAD_ARG, rc
        v = np.ones(ci.size)
        h = C.c_void_p()
        rc = lib.mi_amg_host_build(rp.size - 1, rp.ctypes.data, ci.ctypes.data, v.ctypes.data,
                                   None, C.byref(h))
        assert rc == capi.MI_ERR_BAD_ARG, rc
    # checkpoint: a valid file with an AMG trailer, then corrupted copies
    m = wl.Mesh(6, 5, 4, perturb=0.2, node_seed=12345)
    rng = np.random.default_rng(0)
    p = rng.standard_normal(m.n_cells)
    U = [rng.standard_normal(m.n_cells) for _ in range(3)]
    good = os.path.join(workdir, "good.simcp")
    capi.checkpoint_write(good, m, {"p": p}, {"U": U}, iteration=4, time=0.5)
    import oracle
    rp, ci = oracle.build_sparsity(m)
    vals, _, _ = oracle.assemble_pressure(m, oracle.Matrix(rp, ci, np.zeros(ci.size)),
                                          [wl.BC_WALL] * 7)
    cfg = capi.default_cfg()
    cfg.amg_coarse_size = 16
    capi.checkpoint_append_amg(good, capi.AmgHost(rp, ci, vals, cfg))
    c = capi.Checkpoint(good)
    assert np.array_equal(c.scalar("p"), p) and c.amg_levels >= 2
    data = open(good, "rb").read()
    cases = [data[:k] for k in range(0, len(data), max(1, len(data) // 400))]
    for t in range(600):
        b = bytearray(data)
        for _ in range(int(rng.integers(1, 4))):
            b[int(rng.integers(0, len(b)))] = int(rng.integers(0, 256))
        cases.append(bytes(b))
    bad_path = os.path.join(workdir, "bad.simcp")
    opened = refused = 0
    for case in cases:
        with open(bad_path, "wb") as f:
            f.write(case)
        try:
            cc = capi.Checkpoint(bad_path)
            opened += 1
            try:
                cc.scalar("p")
            except capi.MiError:
                pass
        except capi.MiError:
            refused += 1
    print(f"sanitized host code: {len(cases)} corrupted files ({opened} opened, {refused} refused)",
          flush=True)


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
