"""mi_face_flux_noalloc (the zone velocity table in a library-owned device
array) against mi_face_flux (a table uploaded into a buffer of its own) on
two calls in a row, the second on a smaller mesh with other inlet
velocities: the fluxes must be identical, bit for bit (same kernel, same
table values).  A mesh with more zones than the table holds is refused."""
import ctypes as C
import math

import numpy as np
import pytest

from simall_amd import capi
from simall_amd import workload as wl

pytestmark = pytest.mark.gpu

BC = [wl.BC_WALL, wl.BC_VELOCITY_INLET, wl.BC_OUTFLOW, wl.BC_SYMMETRY, wl.BC_VELOCITY_INLET,
      wl.BC_PRESSURE, wl.BC_WALL]


def _flux_pair(lib, N, seed, bcv):
    L = 2 * math.pi
    m = wl.Mesh(N, N + 1, N + 2, extent=(L, L, L), perturb=0.1, node_seed=seed, perm_seed=seed + 1)
    print(f"no-alloc face flux: mesh {m.n_cells} cells, {m.n_faces} faces", flush=True)
    U = wl.taylor_green(m)
    dU = [capi.DeviceArray.from_host(u) for u in U]
    dm = capi.DeviceMesh(m, BC)
    a, b = capi.DeviceArray(m.n_faces), capi.DeviceArray(m.n_faces)
    bv = np.ascontiguousarray(bcv, dtype=np.float64)
    capi.check(lib.mi_face_flux(dm.h, 1.3, dU[0].ptr, dU[1].ptr, dU[2].ptr,
                                bv.ctypes.data, a.ptr))
    capi.check(lib.mi_face_flux_noalloc(dm.h, 1.3, dU[0].ptr, dU[1].ptr, dU[2].ptr,
                                        bv.ctypes.data, b.ptr))
    fa, fb = a.download(), b.download()
    return m, dm, dU, fa, fb


def test_face_flux_noalloc_matches_allocating_entry_point(gpu):
    bcv1 = np.zeros((7, 3))
    bcv1[1] = [1.0, -0.5, 0.25]
    bcv1[4] = [0.3, 0.7, -1.1]
    m1, _, _, fa, fb = _flux_pair(gpu, 12, 3, bcv1)
    assert np.array_equal(fa, fb)
    assert np.count_nonzero(fa) > m1.n_faces // 2
    # second call: a smaller mesh and other inlet velocities; the table the
    # first call left behind must not leak into the result
    bcv2 = np.zeros((7, 3))
    bcv2[1] = [-2.0, 0.125, 0.5]
    bcv2[4] = [0.0, -0.25, 3.0]
    m2, dm2, dU2, fa2, fb2 = _flux_pair(gpu, 7, 5, bcv2)
    assert m2.n_faces < m1.n_faces
    assert np.array_equal(fa2, fb2)
    # a null table means zero inlet velocity in both entry points
    a, b = capi.DeviceArray(m2.n_faces), capi.DeviceArray(m2.n_faces)
    capi.check(gpu.mi_face_flux(dm2.h, 1.0, dU2[0].ptr, dU2[1].ptr, dU2[2].ptr, None, a.ptr))
    capi.check(gpu.mi_face_flux_noalloc(dm2.h, 1.0, dU2[0].ptr, dU2[1].ptr, dU2[2].ptr, None,
                                        b.ptr))
    assert np.array_equal(a.download(), b.download())


def test_face_flux_noalloc_refuses_too_many_zones(gpu):
    m = wl.Mesh(4, 4, 4)
    bc = [wl.BC_WALL] * 1100  # zones beyond the 1024-entry table
    dm = capi.DeviceMesh(m, bc)
    dU = [capi.DeviceArray(m.n_cells) for _ in range(3)]
    d = capi.DeviceArray(m.n_faces)
    rc = gpu.mi_face_flux_noalloc(dm.h, 1.0, dU[0].ptr, dU[1].ptr, dU[2].ptr, None, d.ptr)
    assert rc == capi.MI_ERR_BAD_ARG
    assert b"1024" in gpu.mi_last_error()
