"""SIMALLCP checkpoint files (mi_checkpoint_*) against the reference's own
io::Checkpoint::write / read (src/io/Checkpoint.cpp:67-230), built into
oracle/_ref: files written here are read by the reference and the reverse,
and both writers produce the same bytes apart from the timestamp and the
meta struct's padding.  Host-only code: no GPU needed.
"""
import ctypes as C
import os

import numpy as np
import pytest

from simall_amd import capi
from simall_amd import workload as wl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libsimall_ref.so")
need_ref = pytest.mark.skipif(not os.path.exists(REF_LIB), reason="reference library not built")
P = C.c_void_p

CASE = dict(N=(6, 5, 4), origin=(0.0, 0.0, 0.0), extent=(1.0, 1.0, 1.0), perturb=0.2, seed=12345)


@pytest.fixture(scope="module")
def lib():
    try:
        return capi.load()
    except ImportError as e:
        pytest.skip(str(e))


@pytest.fixture(scope="module")
def ref():
    lib = C.CDLL(REF_LIB)
    i, d, u64 = C.c_int, C.c_double, C.c_uint64
    lib.ref_checkpoint_write.argtypes = [C.c_char_p, i, i, i, P, P, d, u64, P, P, i, d, u64]
    lib.ref_checkpoint_read.argtypes = [C.c_char_p, P, P, P, P] + [P] * 9 + [P, P, P]
    for f in (lib.ref_checkpoint_write, lib.ref_checkpoint_read):
        f.restype = C.c_int
    return lib


def _mesh():
    N = CASE["N"]
    return wl.Mesh(*N, origin=CASE["origin"], extent=CASE["extent"], perturb=CASE["perturb"],
                   node_seed=CASE["seed"])


def _fields(n):
    rng = np.random.default_rng(3)
    return rng.standard_normal(n), [rng.standard_normal(n) for _ in range(3)]


def _ref_write(ref, path, p, U, iteration, time, hsh):
    o = np.array(CASE["origin"])
    L = np.array(CASE["extent"])
    Ub = np.concatenate(U)
    rc = ref.ref_checkpoint_write(os.fsencode(path), *CASE["N"], o.ctypes.data, L.ctypes.data,
                                  CASE["perturb"], CASE["seed"], p.ctypes.data, Ub.ctypes.data,
                                  iteration, time, hsh)
    assert rc == 0


def _ref_read(ref, path, nC, nF):
    it, t, h = C.c_int(), C.c_double(), C.c_uint64()
    sizes = np.zeros(3, np.int64)
    out = dict(owner=np.empty(nF, np.uint64), neighbor=np.empty(nF, np.uint64), area=np.empty(3 * nF),
               fc=np.empty(3 * nF), zone=np.empty(nF, np.uint32), vol=np.empty(nC), cc=np.empty(3 * nC),
               cfo=np.empty(nC + 1, np.int32), cfi=np.empty(6 * nC, np.uint64), p=np.empty(nC),
               U=np.empty(3 * nC))
    have = np.zeros(2, np.int32)
    rc = ref.ref_checkpoint_read(os.fsencode(path), C.byref(it), C.byref(t), C.byref(h),
                                 sizes.ctypes.data,
                                 *(out[k].ctypes.data for k in ("owner", "neighbor", "area", "fc",
                                                                "zone", "vol", "cc", "cfo", "cfi",
                                                                "p", "U")),
                                 have.ctypes.data)
    return rc, (it.value, t.value, h.value), sizes, out, have


def test_write_read_round_trip(lib, tmp_path):
    m = _mesh()
    p, U = _fields(m.n_cells)
    path = tmp_path / "state.simcp"
    capi.checkpoint_write(path, m, {"p": p, "k": 2 * p}, {"U": U}, iteration=17, time=0.25,
                          settings_hash=0xDEADBEEF12345678)
    assert not os.path.exists(str(path) + ".tmp")
    c = capi.Checkpoint(path)
    assert (c.iteration, c.time, c.settings_hash) == (17, 0.25, 0xDEADBEEF12345678)
    assert (c.n_cells, c.n_faces, c.n_nodes) == (m.n_cells, m.n_faces, m.nodeX.size)
    assert np.array_equal(c.scalar("p"), p) and np.array_equal(c.scalar("k"), 2 * p)
    assert all(np.array_equal(a, b) for a, b in zip(c.vector("U"), U))
    with pytest.raises(capi.MiError):
        c.scalar("nope")
    with pytest.raises(capi.MiError):
        c.vector("p")


def test_header_layout(lib, tmp_path):
    """Header {magic, version 1, flags 0, timestamp} then meta {i32, pad,
    f64, u64}, then the x node coordinates as a u64 count and doubles."""
    m = _mesh()
    path = tmp_path / "h.simcp"
    capi.checkpoint_write(path, m, iteration=-3, time=1.5, settings_hash=7)
    b = open(path, "rb").read()
    assert b[:8] == b"SIMALLCP"
    assert np.frombuffer(b[8:16], np.uint32).tolist() == [1, 0]
    assert np.frombuffer(b[24:28], np.int32)[0] == -3
    assert np.frombuffer(b[32:40], np.float64)[0] == 1.5
    assert np.frombuffer(b[40:48], np.uint64)[0] == 7
    assert np.frombuffer(b[48:56], np.uint64)[0] == m.nodeX.size
    assert np.array_equal(np.frombuffer(b[56:56 + 8 * m.nodeX.size], np.float64), m.nodeX)
    # mesh section of a mesh with no fields ends with the two zero field counts
    assert np.frombuffer(b[-8:], np.uint32).tolist() == [0, 0]


def test_rejects_bad_files(lib, tmp_path):
    m = _mesh()
    good = tmp_path / "g.simcp"
    capi.checkpoint_write(good, m, {"p": np.zeros(m.n_cells)})
    b = bytearray(open(good, "rb").read())
    bad_magic = tmp_path / "m.simcp"
    bad_magic.write_bytes(b"SIMALLCQ" + bytes(b[8:]))
    newer = tmp_path / "v.simcp"
    newer.write_bytes(bytes(b[:8]) + np.uint32(2).tobytes() + bytes(b[12:]))
    trunc = tmp_path / "t.simcp"
    trunc.write_bytes(bytes(b[:-5]))
    for path, code in ((bad_magic, capi.MI_ERR_BAD_ARG), (newer, capi.MI_ERR_UNSUPPORTED),
                       (trunc, capi.MI_ERR_BAD_ARG), (tmp_path / "absent", capi.MI_ERR_BAD_ARG)):
        with pytest.raises(capi.MiError) as e:
            capi.Checkpoint(path)
        assert e.value.code == code, path


@need_ref
def test_reference_reads_our_file(lib, ref, tmp_path):
    m = _mesh()
    p, U = _fields(m.n_cells)
    path = tmp_path / "ours.simcp"
    capi.checkpoint_write(path, m, {"p": p}, {"U": U}, iteration=42, time=3.5, settings_hash=99)
    rc, meta, sizes, out, have = _ref_read(ref, path, m.n_cells, m.n_faces)
    assert rc == 0
    assert meta == (42, 3.5, 99)
    assert sizes.tolist() == [m.n_cells, m.n_faces, m.nodeX.size]
    assert have.tolist() == [1, 1]
    assert np.array_equal(out["p"], p)
    assert np.array_equal(out["U"], np.concatenate(U))
    assert np.array_equal(out["owner"], m.owner) and np.array_equal(out["neighbor"], m.neighbor)
    assert np.array_equal(out["area"], np.concatenate([m.areaX, m.areaY, m.areaZ]))
    assert np.array_equal(out["cc"], np.concatenate([m.cellCx, m.cellCy, m.cellCz]))
    assert np.array_equal(out["vol"], m.volume) and np.array_equal(out["zone"], m.boundaryZone)
    assert np.array_equal(out["cfo"], m.cellFaceOffsets)
    assert np.array_equal(out["cfi"][:m.cellFaceIndices.size], m.cellFaceIndices)


@need_ref
def test_we_read_reference_file(lib, ref, tmp_path):
    m = _mesh()
    p, U = _fields(m.n_cells)
    path = tmp_path / "theirs.simcp"
    _ref_write(ref, path, p, U, 8, 0.125, 0xFFFFFFFFFFFFFFFF)
    assert not os.path.exists(str(path) + ".tmp")
    c = capi.Checkpoint(path)
    assert (c.iteration, c.time, c.settings_hash) == (8, 0.125, 0xFFFFFFFFFFFFFFFF)
    assert (c.n_cells, c.n_faces, c.n_nodes) == (m.n_cells, m.n_faces, m.nodeX.size)
    assert np.array_equal(c.scalar("p"), p)
    assert all(np.array_equal(a, b) for a, b in zip(c.vector("U"), U))


@need_ref
def test_same_bytes_as_reference(lib, ref, tmp_path):
    """Both writers, same mesh and fields: identical files apart from the
    timestamp (bytes 16-24) and the meta padding (bytes 28-32)."""
    m = _mesh()
    p, U = _fields(m.n_cells)
    ours, theirs = tmp_path / "a.simcp", tmp_path / "b.simcp"
    capi.checkpoint_write(ours, m, {"p": p}, {"U": U}, iteration=5, time=0.5, settings_hash=11)
    _ref_write(ref, theirs, p, U, 5, 0.5, 11)
    a, b = bytearray(open(ours, "rb").read()), bytearray(open(theirs, "rb").read())
    assert len(a) == len(b)
    for lo, hi in ((16, 24), (28, 32)):
        a[lo:hi] = b[lo:hi] = bytes(hi - lo)
    assert a == b


@need_ref
def test_reference_rejects_what_we_reject(lib, ref, tmp_path):
    m = _mesh()
    good = tmp_path / "g.simcp"
    capi.checkpoint_write(good, m, {"p": np.ones(m.n_cells)})
    b = open(good, "rb").read()
    for name, data in (("magic", b"XIMALLCP" + b[8:]),
                       ("version", b[:8] + np.uint32(2).tobytes() + b[12:]),
                       ("truncated", b[:-3])):
        path = tmp_path / name
        path.write_bytes(data)
        rc, *_ = _ref_read(ref, path, m.n_cells, m.n_faces)
        assert rc == 1, name
        with pytest.raises(capi.MiError):
            capi.Checkpoint(path)


def _amg_host(m):
    """A host SA-AMG hierarchy of the mesh's pressure matrix (unit conductances)."""
    import oracle
    rp, ci = oracle.build_sparsity(m)
    vals, _, _ = oracle.assemble_pressure(m, oracle.Matrix(rp, ci, np.zeros(ci.size)),
                                          [wl.BC_WALL] * 7)
    cfg = capi.default_cfg()
    cfg.amg_coarse_size = 16
    return capi.AmgHost(rp, ci, vals, cfg)


def test_amg_trailer_round_trip(lib, tmp_path):
    """The optional AMG trailer after the field section is read back with the
    fields intact; a truncated trailer or foreign trailing data is refused."""
    m = _mesh()
    p, U = _fields(m.n_cells)
    H = _amg_host(m)
    assert H.levels >= 2
    path = tmp_path / "amg.simcp"
    capi.checkpoint_write(path, m, {"p": p}, {"U": U}, iteration=3)
    assert capi.Checkpoint(path).amg_levels == 0
    capi.checkpoint_append_amg(path, H)
    c = capi.Checkpoint(path)
    assert c.amg_levels == H.levels
    assert np.array_equal(c.scalar("p"), p) and c.iteration == 3
    b = open(path, "rb").read()
    for name, data in (("trunc", b[:-5]), ("foreign", b[:-1] + b"\x00" * 9)):
        bad = tmp_path / name
        bad.write_bytes(data)
        with pytest.raises(capi.MiError):
            capi.Checkpoint(bad)


@need_ref
def test_reference_reads_file_with_amg_trailer(lib, ref, tmp_path):
    m = _mesh()
    p, U = _fields(m.n_cells)
    path = tmp_path / "amg.simcp"
    capi.checkpoint_write(path, m, {"p": p}, {"U": U}, iteration=9, time=1.0)
    capi.checkpoint_append_amg(path, _amg_host(m))
    rc, meta, sizes, out, have = _ref_read(ref, path, m.n_cells, m.n_faces)
    assert rc == 0 and meta == (9, 1.0, 0) and have.tolist() == [1, 1]
    assert np.array_equal(out["p"], p) and np.array_equal(out["U"], np.concatenate(U))


def test_rejects_counts_beyond_the_file(lib, tmp_path):
    """An array count larger than what is left of the file is refused before
    anything is allocated for it."""
    m = _mesh()
    path = tmp_path / "c.simcp"
    capi.checkpoint_write(path, m, {"p": np.zeros(m.n_cells)})
    b = bytearray(open(path, "rb").read())
    # the count of the last array (field "p") sits 8 * n_cells + 8 bytes from the end
    pos = len(b) - 8 * m.n_cells - 8
    assert np.frombuffer(bytes(b[pos:pos + 8]), np.uint64)[0] == m.n_cells
    b[pos:pos + 8] = np.uint64(1 << 39).tobytes()
    bad = tmp_path / "huge.simcp"
    bad.write_bytes(bytes(b))
    with pytest.raises(capi.MiError):
        capi.Checkpoint(bad)


def _trailer_fields(b):
    """Byte offsets of each level's {A, P, R} row/column counts in the AMG
    trailer of file bytes b (layout: tag, u32 levels, per level u32 has_P, CSR A,
    f64, i32, [CSR P, CSR R]; a CSR is i32 rows, i32 cols, three u64-counted
    arrays of 4, 4 and 8 bytes per element)."""
    pos = b.rindex(b"SIMAMGH1") + 8
    nl = int(np.frombuffer(b[pos:pos + 4], np.uint32)[0])
    pos += 4
    out = []

    def csr(pos):
        at = pos
        pos += 8
        for width in (4, 4, 8):
            cnt = int(np.frombuffer(b[pos:pos + 8], np.uint64)[0])
            pos += 8 + width * cnt
        return at, pos

    for _ in range(nl):
        has_p = int(np.frombuffer(b[pos:pos + 4], np.uint32)[0])
        pos += 4
        a, pos = csr(pos)
        pos += 12
        lev = {"A": a}
        if has_p:
            lev["P"], pos = csr(pos)
            lev["R"], pos = csr(pos)
        out.append(lev)
    return out


def test_rejects_amg_trailer_whose_shapes_do_not_chain(lib, tmp_path):
    """Each CSR of the trailer is valid on its own, but the V-cycle needs A_l
    square, P_l of shape n_l x n_{l+1} and R_l of n_{l+1} x n_l: a trailer that
    breaks any of these is refused (it would index level vectors out of range)."""
    m = _mesh()
    path = tmp_path / "amg.simcp"
    capi.checkpoint_write(path, m, {"p": np.zeros(m.n_cells)})
    capi.checkpoint_append_amg(path, _amg_host(m))
    good = open(path, "rb").read()
    capi.Checkpoint(path)  # the untouched file is accepted
    lev = _trailer_fields(good)
    assert len(lev) >= 2 and "P" in lev[0]
    # widen one column count: every column index stays inside it, so only the
    # shape check can catch it
    for which, lvl in (("A", 0), ("P", 0), ("R", 0), ("A", len(lev) - 1)):
        b = bytearray(good)
        at = lev[lvl][which] + 4
        nc = int(np.frombuffer(bytes(b[at:at + 4]), np.int32)[0])
        b[at:at + 4] = np.int32(nc + 1).tobytes()
        bad = tmp_path / f"bad_{which}{lvl}.simcp"
        bad.write_bytes(bytes(b))
        with pytest.raises(capi.MiError):
            capi.Checkpoint(bad)
