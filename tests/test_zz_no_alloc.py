"""hf_flux_stage_noalloc (the flux stage that allocates nothing, its tables in
a static device array) against hf_flux_stage_fluxes through the C ABI: two
calls in a row, the second on a smaller set of boxes, each bit-identical in
the updated state, the face fluxes and the statistics; the size limits are
refused with HF_EINVAL."""
from ctypes import byref, c_int64

import numpy as np
import pytest

import hyperflow_amd as hf
from hyperflow_amd import _lib
from hyperflow_amd.amr import face_fluxes
from tests.conftest import H2
from tests.test_gpu_flux import smooth_box

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _layout(oracle, sizes, seed):
    """Boxes of the given (nx, ny) packed into one input and one output
    array, with per-box face-flux arrays."""
    nv = oracle.ns + 3
    ins, outs, boxes, fl = [], [], [], []
    in_off = out_off = 0
    for k, (nx, ny) in enumerate(sizes):
        G = smooth_box(oracle, nx, ny, seed=seed + k)
        gx, gy = nx + 4, ny + 4
        boxes.append(_lib.hf_box(nx, ny, [0, 1, 4, 15][k % 4], 0, in_off, gx, gx * gy, out_off, nx, nx * ny,
                                 2e-4, 1.5e-4))
        ins.append(G.reshape(nv, -1))
        in_off += gx * gy
        out_off += nx * ny
        fl.append((torch.zeros((nv, ny, nx + 1), dtype=torch.float64, device="cuda"),
                   torch.zeros((nv, ny + 1, nx), dtype=torch.float64, device="cuda")))
    # one plane per component: element n of box b at n*plane + offset_b
    Uin = np.concatenate(ins, axis=1)
    for b in boxes:
        b.in_plane = Uin.shape[1]
        b.out_plane = out_off
    return torch.from_numpy(np.ascontiguousarray(Uin)).cuda(), boxes, fl, out_off


def _run(fn, mech, boxes, Uin, nout, fl):
    nb = len(boxes)
    arr = (_lib.hf_box * nb)(*boxes)
    desc = (_lib.hf_face_fluxes * nb)(*[face_fluxes(fx, fy) for fx, fy in fl])
    out = torch.zeros((mech.nv, nout), dtype=torch.float64, device="cuda")
    st, ff = _lib.hf_flux_stats(), c_int64(-1)
    r = fn(mech.handle, arr, nb, Uin.data_ptr(), out.data_ptr(), 0, desc, byref(st), byref(ff))
    assert r == 0, _lib.load().hf_last_error(mech.handle)
    return out.cpu().numpy(), [(fx.cpu().numpy(), fy.cpu().numpy()) for fx, fy in fl], st


@pytest.mark.parametrize("sizes", [[(40, 33), (15, 15), (64, 17), (31, 48), (7, 5)], [(20, 9), (3, 30)]])
def test_noalloc_matches_allocating_stage_twice(oracle_h2, sizes):
    print(f"[gpu] no-allocation flux stage, boxes {sizes} then {sizes[:1]}", flush=True)
    L = _lib.load()
    mech = hf.Mechanism(H2)
    for call, sz in enumerate((sizes, sizes[:1])):    # the second call is smaller
        Uin, boxes, fl_a, nout = _layout(oracle_h2, sz, seed=17 + call)
        fl_b = [(torch.full_like(fx, 7.0), torch.full_like(fy, 7.0)) for fx, fy in fl_a]
        ref, ref_fl, st_a = _run(L.hf_flux_stage_fluxes, mech, boxes, Uin, nout, fl_a)
        got, got_fl, st_b = _run(L.hf_flux_stage_noalloc, mech, boxes, Uin, nout, fl_b)
        assert np.array_equal(got, ref)
        for (a, b), (c, d) in zip(got_fl, ref_fl):
            assert np.array_equal(a, c) and np.array_equal(b, d)
        assert st_b.cells == st_a.cells
        # faces computed count a face two tiles share twice: equal for the same tile shape
        # (the allocating stage may be trying another shape for its tuner)
        if (st_a.tile_x, st_a.tile_y) == (st_b.tile_x, st_b.tile_y):
            assert (st_b.faces, st_b.first_order_sides) == (st_a.faces, st_a.first_order_sides)
        print(f"[gpu]   call {call}: {len(sz)} boxes, {st_b.faces} faces, identical", flush=True)


def test_noalloc_without_fluxes_matches_flux_stage(oracle_h2):
    L = _lib.load()
    mech = hf.Mechanism(H2)
    Uin, boxes, _, nout = _layout(oracle_h2, [(33, 21), (16, 16)], seed=5)
    nb = len(boxes)
    arr = (_lib.hf_box * nb)(*boxes)
    outs = []
    for fn in (L.hf_flux_stage, L.hf_flux_stage_noalloc):
        out = torch.zeros((mech.nv, nout), dtype=torch.float64, device="cuda")
        st, ff = _lib.hf_flux_stats(), c_int64(-1)
        args = (mech.handle, arr, nb, Uin.data_ptr(), out.data_ptr(), 0)
        r = fn(*args, byref(st), byref(ff)) if fn is L.hf_flux_stage else fn(*args, None, byref(st), byref(ff))
        assert r == 0
        outs.append(out.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])


def test_noalloc_limits_and_host_arrays(oracle_h2):
    L = _lib.load()
    mech = hf.Mechanism(H2)
    st, ff = _lib.hf_flux_stats(), c_int64(-1)
    # 4097 boxes
    many = (_lib.hf_box * 4097)(*[_lib.hf_box(1, 1, 0, 0, 0, 5, 25, 0, 1, 1, 1e-4, 1e-4)] * 4097)
    dU = torch.zeros((mech.nv, 25), dtype=torch.float64, device="cuda")
    dO = torch.zeros((mech.nv, 1), dtype=torch.float64, device="cuda")
    assert L.hf_flux_stage_noalloc(mech.handle, many, 4097, dU.data_ptr(), dO.data_ptr(), 0, None, byref(st),
                                   byref(ff)) == _lib.HF_EINVAL
    # one box of 65537 tiles: 15 x 15 cells per tile
    big = _lib.hf_box(15 * 257, 15 * 256, 0, 0, 0, 15 * 257 + 4, (15 * 257 + 4) * (15 * 256 + 4), 0, 15 * 257,
                      15 * 257 * 15 * 256, 1e-4, 1e-4)
    one = (_lib.hf_box * 1)(big)
    assert L.hf_flux_stage_noalloc(mech.handle, one, 1, dU.data_ptr(), dO.data_ptr(), 0, None, byref(st),
                                   byref(ff)) == _lib.HF_EINVAL
    assert "tiles" in L.hf_last_error(mech.handle).decode()
    # host arrays are refused (the allocating stage would stage them)
    hU = np.zeros((mech.nv, 25))
    hO = np.zeros((mech.nv, 1))
    small = (_lib.hf_box * 1)(_lib.hf_box(1, 1, 0, 0, 0, 5, 25, 0, 1, 1, 1e-4, 1e-4))
    assert L.hf_flux_stage_noalloc(mech.handle, small, 1, hU.ctypes.data, hO.ctypes.data, 0, None, byref(st),
                                   byref(ff)) == _lib.HF_EINVAL
