"""The C ABI library (CPU, no GPU needed): it loads, exports every symbol the
header declares, the Python binding declares exactly that set, and validation
errors raised before any device allocation carry the reference's messages."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from timemachine_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def _header_symbols():
    text = (ROOT / "include" / "tm_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tm_[A-Za-z0-9_]+)\s*\(", text)))


def test_library_loads_and_exports_header_symbols():
    lib = _lib.load()
    syms = _header_symbols()
    assert len(syms) > 40
    for s in syms:
        assert hasattr(lib, s), s


def test_binding_declares_the_header():
    assert sorted(_lib.declared_symbols()) == _header_symbols()


def test_tile_size():
    assert _lib.load().tm_tile_size() == 64


def _err():
    return _lib.load().tm_last_error().decode()


def test_validation_errors_before_allocation():
    lib = _lib.load()
    h = ctypes.c_void_p()
    idx = np.array([0, 1, 2], dtype=np.int32)
    rc = lib.tm_harmonic_bond_create(1, _lib.pi(np.array([1, 1], np.int32)), None, None, ctypes.byref(h))
    assert rc == 1 and _err() == "src == dst"
    rc = lib.tm_harmonic_angle_create(1, _lib.pi(np.array([0, 1, 0], np.int32)), None, None, ctypes.byref(h))
    assert rc == 1 and _err() == "angle triplets must be unique"
    rc = lib.tm_periodic_torsion_create(1, _lib.pi(np.array([0, 1, 2, 2], np.int32)), None, None, ctypes.byref(h))
    assert rc == 1 and _err() == "torsion quads must be unique"
    pl = np.zeros(3, np.int32)
    of = np.zeros(2, np.int32)
    rc = lib.tm_nonbonded_pair_list_create(3, 1, _lib.pi(np.array([2, 2], np.int32)), _lib.pd(np.ones(2)),
                                           _lib.pi(pl), _lib.pi(np.zeros(3, np.int32)), 2.0, 1.2, 1, ctypes.byref(h))
    assert rc == 1 and _err() == "illegal pair with src == dst: 2, 2"
    rc = lib.tm_barostat_create(3, 1.0, 300.0, 1, _lib.pi(np.array([3], np.int32)), _lib.pi(np.array([0, 1, 2], np.int32)),
                                0, 0, None, 1, ctypes.byref(h))
    assert rc == 1 and _err() == "Barostat interval must be greater than 0"
    rc = lib.tm_barostat_create(3, 1.0, 300.0, 1, _lib.pi(np.array([2], np.int32)), _lib.pi(np.array([0, 5], np.int32)),
                                5, 0, None, 1, ctypes.byref(h))
    assert rc == 1 and _err() == "Grouped indices must be between 0 and N"
    del idx, of


def test_python_level_messages():
    from timemachine_amd.lib import custom_ops

    with pytest.raises(RuntimeError, match="lambda offset idxs and plane idxs need to be equivalent"):
        custom_ops.NonbondedAllPairs_f32("", np.zeros(3, np.int32), np.zeros(2, np.int32), 2.0, 1.2)
    with pytest.raises(RuntimeError, match="atom indices must be unique"):
        custom_ops.NonbondedAllPairs_f32("", np.zeros(3, np.int32), np.zeros(3, np.int32), 2.0, 1.2, np.array([0, 0]))
    with pytest.raises(NotImplementedError):
        custom_ops.HarmonicBond_f64(np.array([[0, 1]]))
    with pytest.raises(RuntimeError, match=r"bond_idxs.size\(\) must be exactly 2\*k!"):
        custom_ops.HarmonicBond_f32(np.array([0, 1, 2]))


def test_fixed_point_decode():
    v = np.array([2**36, (-(2**36)) % 2**64, 3 * 2**35], dtype=np.uint64)
    np.testing.assert_array_equal(_lib.fixed_to_float(v), [1.0, -1.0, 1.5])
    assert _lib.wrapping_sum(np.array([2**63, 2**63], dtype=np.uint64)) == 0
