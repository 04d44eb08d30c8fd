"""Allocation-free nonbonded execute (GPU): tm_potential_execute_device_noalloc
against tm_potential_execute_device on the same device inputs, all four
outputs compared bit for bit (and both against tm_potential_execute_host).
Two calls in a row, the second on a smaller system (the static scratch then
holds the first call's larger neighborlist), then the first system again;
plus the size limit and the rejected forms. Device buffers come from the HIP
runtime the library itself links (libamdhip64.so.7), through ctypes."""

import ctypes

import numpy as np
import pytest

from timemachine_amd import _lib
from timemachine_amd import testsystems as ts

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip(gpu):
    rt = ctypes.CDLL("libamdhip64.so.7")  # the copy the library already loaded
    rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    rt.hipFree.argtypes = [ctypes.c_void_p]
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rt.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    return rt


class _Buf:
    def __init__(self, rt, nbytes):
        self.rt, self.nbytes = rt, nbytes
        self.ptr = ctypes.c_void_p()
        assert rt.hipMalloc(ctypes.byref(self.ptr), nbytes) == 0

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        assert self.rt.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, dtype):
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype)
        assert self.rt.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        return out

    def free(self):
        self.rt.hipFree(self.ptr)


def _create(sys_, interpolated=False):
    plane = _lib.i32(sys_.lambda_plane_idxs)
    offset = _lib.i32(sys_.lambda_offset_idxs)
    h = ctypes.c_void_p()
    name = "tm_nonbonded_all_pairs_interpolated_create" if interpolated else "tm_nonbonded_all_pairs_create"
    # n_atom_idxs = -1: every atom (no subset)
    _lib.call(name, sys_.N, _lib.pi(plane), _lib.pi(offset), sys_.beta, sys_.cutoff, None, -1, ctypes.byref(h))
    return h


def _execute(rt, entry, h, sys_, lam):
    """One call of entry on fresh caller-zeroed device outputs; returns them as uint64."""
    N = sys_.N
    ins = [_Buf(rt, 8 * a.size) for a in (sys_.x, sys_.nb_params, sys_.box)]
    for b, a in zip(ins, (sys_.x, sys_.nb_params, sys_.box)):
        b.put(np.asarray(a, np.float64))
    outs = [_Buf(rt, 8 * n) for n in (N * 3, N * 3, N, N)]
    for b in outs:
        assert rt.hipMemset(b.ptr, 0, b.nbytes) == 0
    assert rt.hipDeviceSynchronize() == 0
    _lib.call(entry, h, N, N * 3, ins[0].ptr, ins[1].ptr, ins[2].ptr, lam, *[b.ptr for b in outs], None)
    assert rt.hipDeviceSynchronize() == 0
    res = [b.get(np.uint64) for b in outs]
    for b in ins + outs:
        b.free()
    return res


def _host(h, sys_, lam):
    N = sys_.N
    out = [np.zeros(n, np.uint64) for n in (N * 3, N * 3, N, N)]
    _lib.call("tm_potential_execute_host", h, N, N * 3, _lib.pd(_lib.f64(sys_.x)), _lib.pd(_lib.f64(sys_.nb_params)),
              _lib.pd(_lib.f64(sys_.box)), lam, *[_lib.pu64(o) for o in out])
    return out


def _same(a, b):
    for name, u, v in zip(("du_dx", "du_dp", "du_dl", "u"), a, b):
        np.testing.assert_array_equal(u, v, err_msg=name)


@pytest.fixture(scope="module")
def systems(gpu):
    big = ts.ligand_in_water()  # 6372 atoms, ligand offsets: alchemical tiles at lambda 0.5
    small = ts.water_box()  # 2661 atoms
    return big, small


def test_noalloc_equals_execute_device_bitwise(systems, hip):
    big, small = systems
    hb, hs = _create(big), _create(small)
    try:
        for sys_, h, lam in ((big, hb, 0.5), (small, hs, 0.0), (big, hb, 0.5)):
            print(f"noalloc vs execute_device: N={sys_.N} lambda={lam}", flush=True)
            host = _host(h, sys_, lam)
            assert np.any(host[0] != 0) and host[3][0] != 0
            ref = _execute(hip, "tm_potential_execute_device", h, sys_, lam)
            _same(host, ref)
            got = _execute(hip, "tm_potential_execute_device_noalloc", h, sys_, lam)
            _same(ref, got)
            if lam == 0.5:
                assert ref[2][0] != 0  # du/dlambda of the alchemical tiles
    finally:
        _lib.load().tm_potential_release(hb)
        _lib.load().tm_potential_release(hs)


def test_noalloc_rejects_oversize_and_other_forms(systems):
    lib = _lib.load()
    big, _ = systems
    n = 30000  # over the 16 MiB scratch
    h = ctypes.c_void_p()
    z = np.zeros(n, np.int32)
    _lib.call("tm_nonbonded_all_pairs_create", n, _lib.pi(z), _lib.pi(z), 2.0, 1.2, None, -1, ctypes.byref(h))
    try:
        rc = lib.tm_potential_execute_device_noalloc(h, n, n * 3, None, None, None, 0.0, None, None, None, None, None)
        assert rc == 1 and "bytes of scratch, more than the 16777216-byte maximum" in lib.tm_last_error().decode()
    finally:
        lib.tm_potential_release(h)
    hi = _create(big, interpolated=True)
    try:
        rc = lib.tm_potential_execute_device_noalloc(hi, big.N, big.N * 6, None, None, None, 0.0, None, None, None,
                                                     None, None)
        assert rc == 1 and "no atom subset, row subset or interpolation" in lib.tm_last_error().decode()
    finally:
        lib.tm_potential_release(hi)
    hb = ctypes.c_void_p()
    _lib.call("tm_harmonic_bond_create", 1, _lib.pi(np.array([0, 1], np.int32)), None, None, ctypes.byref(hb))
    try:
        rc = lib.tm_potential_execute_device_noalloc(hb, 2, 2, None, None, None, 0.0, None, None, None, None, None)
        assert rc == 1 and "not a NonbondedAllPairs" in lib.tm_last_error().decode()
    finally:
        lib.tm_potential_release(hb)
