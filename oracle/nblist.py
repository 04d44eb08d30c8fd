"""Neighborlist references.

build_reference_ixn_list restates tests/test_nblist.py:70-94 (and image_coords,
:66-67) with the block size as a parameter: for row block r, every column atom
j >= first atom of the block whose f64 periodic distance to at least one row
atom is < cutoff (the block's own atoms included, distance 0).

build_reference_ixn_list_with_subset restates tests/test_nblist.py:95-131 (row
subset: row blocks taken from row_idxs in the order given, columns = the other
atoms); fp32_ixn_list_rows is its FP32 counterpart.

fp32_ixn_list is the exact FP32 predicate of the HIP neighborlist (the
min-image form of kernels/k_neighborlist.cuh:374-384 with rint ties-to-even):
  d = x_i - x_j; d -= b * rint(d * (1/b)) per axis; d2 = dx*dx + dy*dy + dz*dz;
  interacting iff d2 < c*c, with c = float(cutoff + padding), b = float(box).
It is implemented in C (nblist_ref.c) so FP32 rounding is exactly IEEE; the
result must equal the GPU set bit for bit.
"""

from __future__ import annotations

import ctypes
from pathlib import Path
from typing import List

import numpy as np

_HERE = Path(__file__).resolve().parent


def image_coords(x, box_diag):
    return x - box_diag * np.floor(x / box_diag + 0.5)


def build_reference_ixn_list(coords, box, cutoff, block_size=64) -> List[List[int]]:
    coords = np.asarray(coords, dtype=np.float64)
    N = coords.shape[0]
    box_diag = np.diagonal(np.asarray(box))
    out = []
    col = coords[None, :, :]
    for rb in range((N + block_size - 1) // block_size):
        r0 = rb * block_size
        rows = coords[r0 : min(r0 + block_size, N)][:, None, :]
        d = np.linalg.norm(image_coords(rows - col, box_diag), axis=-1)
        d[:, :r0] = cutoff
        out.append(np.argwhere(np.any(d < cutoff, axis=0)).reshape(-1).tolist())
    return out


def build_reference_ixn_list_with_subset(coords, box, cutoff, row_idxs, block_size=64) -> List[List[int]]:
    coords = np.asarray(coords, dtype=np.float64)
    N = coords.shape[0]
    row_idxs = np.asarray(row_idxs).reshape(-1)
    col_idxs = np.delete(np.arange(N), row_idxs)
    box_diag = np.diagonal(np.asarray(box))
    col = coords[col_idxs][None, :, :]
    out = []
    for rb in range((len(row_idxs) + block_size - 1) // block_size):
        rows = coords[row_idxs[rb * block_size : (rb + 1) * block_size]][:, None, :]
        d = np.linalg.norm(image_coords(rows - col, box_diag), axis=-1)
        out.append(col_idxs[np.any(d < cutoff, axis=0)].tolist())
    return out


def brute_force_ixn_list(coords, box, cutoff, block_size=64) -> List[List[int]]:
    """Pure-Python double loop (tiny N): the same set, no vectorisation."""
    coords = np.asarray(coords, dtype=np.float64)
    N = coords.shape[0]
    L = np.diagonal(np.asarray(box))
    out = []
    for rb in range((N + block_size - 1) // block_size):
        r0 = rb * block_size
        s = []
        for j in range(r0, N):
            hit = False
            for i in range(r0, min(r0 + block_size, N)):
                d2 = 0.0
                for k in range(3):
                    dk = coords[i, k] - coords[j, k]
                    dk -= L[k] * np.floor(dk / L[k] + 0.5)
                    d2 += dk * dk
                if np.sqrt(d2) < cutoff:
                    hit = True
                    break
            if hit:
                s.append(j)
        out.append(s)
    return out


_clib = None


def _lib():
    global _clib
    if _clib is None:
        path = _HERE / "_build" / "liboracle.so"
        if not path.exists():
            from oracle import build as _b

            _b.build()
        lib = ctypes.CDLL(str(path))
        lib.oracle_fp32_ixn_mask.restype = ctypes.c_int
        lib.oracle_fp32_ixn_mask.argtypes = [
            ctypes.c_int,
            ctypes.POINTER(ctypes.c_double),
            ctypes.POINTER(ctypes.c_double),
            ctypes.c_double,
            ctypes.c_int,
            ctypes.POINTER(ctypes.c_uint8),
        ]
        lib.oracle_fp32_ixn_mask_rows.restype = ctypes.c_int
        lib.oracle_fp32_ixn_mask_rows.argtypes = [
            ctypes.c_int,
            ctypes.POINTER(ctypes.c_double),
            ctypes.POINTER(ctypes.c_double),
            ctypes.c_double,
            ctypes.c_int,
            ctypes.c_int,
            ctypes.POINTER(ctypes.c_int),
            ctypes.POINTER(ctypes.c_uint8),
        ]
        lib.oracle_round_to_fixed.restype = ctypes.c_uint64
        lib.oracle_round_to_fixed.argtypes = [ctypes.c_double]
        _clib = lib
    return _clib


def fp32_ixn_list(coords, box, cutoff, block_size=64) -> List[List[int]]:
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    box = np.ascontiguousarray(box, dtype=np.float64)
    N = coords.shape[0]
    nb = (N + block_size - 1) // block_size
    mask = np.zeros((nb, N), dtype=np.uint8)
    _lib().oracle_fp32_ixn_mask(
        N,
        coords.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
        box.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
        float(cutoff),
        block_size,
        mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
    )
    return [np.nonzero(mask[r])[0].tolist() for r in range(nb)]


def fp32_ixn_list_rows(coords, box, cutoff, row_idxs, block_size=64) -> List[List[int]]:
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    box = np.ascontiguousarray(box, dtype=np.float64)
    rows = np.ascontiguousarray(np.asarray(row_idxs).reshape(-1), dtype=np.int32)
    N = coords.shape[0]
    nb = (rows.size + block_size - 1) // block_size
    mask = np.zeros((nb, N), dtype=np.uint8)
    rc = _lib().oracle_fp32_ixn_mask_rows(
        N,
        coords.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
        box.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
        float(cutoff),
        block_size,
        rows.size,
        rows.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
        mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
    )
    if rc != 0:
        raise MemoryError("oracle_fp32_ixn_mask_rows")
    return [np.nonzero(mask[r])[0].tolist() for r in range(nb)]


def round_to_fixed(v: float) -> int:
    """llrint(v) as uint64 (0x8000000000000000 out of range), the fixed-point rounding."""
    return int(_lib().oracle_round_to_fixed(float(v)))
