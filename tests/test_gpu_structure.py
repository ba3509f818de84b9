"""Device matrix-structure builder (ocsl_build_structure) against the oracle's
restatement of CFVM::define_matrix_structure, the reference's own Sod
structure and the vectorised hex-mesh generator."""
import ctypes
import os

import numpy as np
import pytest

import oracle
from opencslsolver_amd import gen
from opencslsolver_amd._native import hip_lib
from opencslsolver_amd.solver import OcslError, build_structure

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _check_slots(rp, col, fl, fr, fs, ds):
    fl = fl.astype(np.int64)
    fr = fr.astype(np.int64)
    assert np.all(col[fs[:, 0]] == fr) and np.all(col[fs[:, 1]] == fl)
    assert np.all((fs[:, 0] >= rp[fl].astype(np.int64)) & (fs[:, 0] < rp[fl + 1].astype(np.int64)))
    assert np.all((fs[:, 1] >= rp[fr].astype(np.int64)) & (fs[:, 1] < rp[fr + 1].astype(np.int64)))
    n = rp.shape[0] - 1
    assert np.all(col[ds] == np.arange(n))
    assert np.all((ds >= rp[:-1].astype(np.int64)) & (ds < rp[1:].astype(np.int64)))


def test_sod_structure_equals_reference():
    m = np.load(os.path.join(GOLDEN, "flow_sod1000_mesh.npz"))
    d = np.load(os.path.join(GOLDEN, "sod_cg_2steps.npz"))
    fl, fr, bl = m["mesh_face_left"], m["mesh_face_right"], m["mesh_bc_left"]
    rp, col, fs, ds = build_structure(1000, fl, fr, bl, slots=True)
    assert np.array_equal(rp, d["row_ptr"]) and np.array_equal(col, d["col"])
    _check_slots(rp, col, fl, fr, fs, ds)


@pytest.mark.parametrize("n,seed", [(4, None), (9, 1), (17, 2)])
def test_hex_mesh_structure(n, seed):
    perm = None if seed is None else np.random.default_rng(seed).permutation(n ** 3)
    mesh, rp_g, col_g = gen.hex_mesh(n, perm=perm)
    rp, col, fs, ds = build_structure(n ** 3, mesh["face_left"], mesh["face_right"], mesh["bc_left"], slots=True)
    assert np.array_equal(rp, rp_g) and np.array_equal(col, col_g)
    _check_slots(rp, col, mesh["face_left"], mesh["face_right"], fs, ds)


def test_headline_mesh_structure():
    """The 128^3 mesh of the headline system (6.2 M faces)."""
    mesh, rp_g, col_g = gen.hex_mesh(128)
    rp, col = build_structure(128 ** 3, mesh["face_left"], mesh["face_right"], mesh["bc_left"])
    assert np.array_equal(rp, rp_g) and np.array_equal(col, col_g)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_meshes_with_repeats_and_isolated_cells(seed):
    """Random face lists: faces listed twice (either orientation), cells with
    several boundary faces, cells with none or without any face, rows longer
    than a hex cell's."""
    rng = np.random.default_rng(seed)
    n = 700
    F = 1600
    fl = rng.integers(0, n - 40, F).astype(np.int32)  # the last 40 cells get no face
    fr = rng.integers(0, n - 40, F).astype(np.int32)
    keep = fl != fr
    fl, fr = fl[keep], fr[keep]
    fl = np.concatenate([fl, fr[:50]])
    fr = np.concatenate([fr, fl[:50]])  # 50 repeated faces, reversed
    bl = rng.integers(0, n, 900).astype(np.int32)
    rp, col, fs, ds = build_structure(n, fl, fr, bl, slots=True)
    rp_o, col_o = oracle.define_matrix_structure(n, fl, fr, bl)
    assert np.array_equal(rp, rp_o) and np.array_equal(col, col_o)
    _check_slots(rp, col, fl, fr, fs, ds)


def test_device_pointers_and_size_query():
    import torch

    mesh, rp_g, col_g = gen.hex_mesh(12, perm=np.random.default_rng(5).permutation(12 ** 3))
    dev = torch.device("cuda:0")
    fl = torch.from_numpy(mesh["face_left"]).to(dev)
    fr = torch.from_numpy(mesh["face_right"]).to(dev)
    bl = torch.from_numpy(mesh["bc_left"]).to(dev)
    n = 12 ** 3
    lib = hip_lib()
    nnz = ctypes.c_int64()
    rp = torch.empty(n + 1, dtype=torch.int64, device=dev)
    vp = ctypes.c_void_p
    st = lib.ocsl_build_structure(n, fl.numel(), vp(fl.data_ptr()), vp(fr.data_ptr()), bl.numel(),
                                  vp(bl.data_ptr()), 1, vp(rp.data_ptr()), None, 0, ctypes.byref(nnz), None, None)
    assert st == 0 and nnz.value == col_g.shape[0]
    torch.cuda.synchronize()
    assert np.array_equal(rp.cpu().numpy().view(np.uint64), rp_g)
    col = torch.empty(nnz.value, dtype=torch.int64, device=dev)
    fs = torch.empty(2 * fl.numel(), dtype=torch.int64, device=dev)
    st = lib.ocsl_build_structure(n, fl.numel(), vp(fl.data_ptr()), vp(fr.data_ptr()), bl.numel(),
                                  vp(bl.data_ptr()), 1, vp(rp.data_ptr()), vp(col.data_ptr()), nnz.value,
                                  ctypes.byref(nnz), vp(fs.data_ptr()), None)
    assert st == 0
    torch.cuda.synchronize()
    assert np.array_equal(col.cpu().numpy().view(np.uint64), col_g)
    _, _, fs_h, ds_h = build_structure(n, mesh["face_left"], mesh["face_right"], mesh["bc_left"], slots=True)
    assert np.array_equal(fs.cpu().numpy().reshape(-1, 2), fs_h)


def test_errors():
    with pytest.raises(OcslError):
        build_structure(4, [0, 1], [1, 4], [])  # cell 4 of 4
    with pytest.raises(OcslError):
        build_structure(4, [0, 2], [1, 2], [])  # one cell on both sides
    with pytest.raises(OcslError):
        build_structure(4, [0], [1], [7])  # boundary face of cell 7
    lib = hip_lib()
    fl = np.array([0, 1], np.int32)
    fr = np.array([1, 2], np.int32)
    col = np.empty(2, np.uint64)
    nnz = ctypes.c_int64()
    vp = ctypes.c_void_p
    st = lib.ocsl_build_structure(3, 2, fl.ctypes.data_as(vp), fr.ctypes.data_as(vp), 0, None, 0, None,
                                  col.ctypes.data_as(vp), 2, ctypes.byref(nnz), None, None)
    assert st != 0 and nnz.value == 7  # 7 entries do not fit in 2


def test_meshes_without_interior_faces():
    """Cells with no interior face: each row holds its diagonal, and the
    sentinel when the cell owns boundary faces."""
    rp, col = build_structure(3, [], [], [])
    assert rp.tolist() == [0, 1, 2, 3] and col.tolist() == [0, 1, 2]
    rp, col, fs, ds = build_structure(1, [], [], [0, 0, 0, 0, 0, 0], slots=True)
    assert rp.tolist() == [0, 2] and col.tolist() == [0, oracle.SENTINEL_COL]
    assert fs.shape == (0, 2) and ds.tolist() == [0]
    rp_o, col_o = oracle.define_matrix_structure(1, [], [], [0, 0, 0, 0, 0, 0])
    assert np.array_equal(rp, rp_o) and np.array_equal(col, col_o)
