"""ctypes binding of ``libtm_amd.so`` (declared in ``include/tm_amd.h``).

The library is built in-tree by ``timemachine_amd.build``. There is no CPU
fallback: if the library is missing this module raises on first use.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_uint32, c_uint64, c_void_p
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libtm_amd.so"

_P = c_void_p
_PD = POINTER(c_double)
_PI = POINTER(c_int32)
_PU64 = POINTER(c_uint64)
_PPV = POINTER(c_void_p)

# name: (restype, argtypes)
_SIGS = {
    "tm_last_error": (c_char_p, []),
    "tm_tile_size": (c_int, []),
    "tm_device_count": (c_int, [POINTER(c_int)]),
    "tm_nonbonded_all_pairs_create": (c_int, [c_int, _PI, _PI, c_double, c_double, _PI, c_int, _PPV]),
    "tm_nonbonded_interaction_group_create": (
        c_int,
        [c_int, c_int, _PI, _PI, _PI, c_double, c_double, c_int, _PPV],
    ),
    "tm_nonbonded_all_pairs_interpolated_create": (c_int, [c_int, _PI, _PI, c_double, c_double, _PI, c_int, _PPV]),
    "tm_nonbonded_all_pairs_set_nblist_padding": (c_int, [_P, c_double]),
    "tm_nonbonded_all_pairs_disable_hilbert_sort": (c_int, [_P]),
    "tm_nonbonded_all_pairs_num_tiles": (c_int, [_P, POINTER(c_int)]),
    "tm_nonbonded_pair_list_create": (
        c_int,
        [c_int, c_int, _PI, _PD, _PI, _PI, c_double, c_double, c_int, _PPV],
    ),
    "tm_nonbonded_pair_list_interpolated_create": (
        c_int,
        [c_int, c_int, _PI, _PD, _PI, _PI, c_double, c_double, c_int, _PPV],
    ),
    "tm_harmonic_bond_create": (c_int, [c_int, _PI, _PI, _PI, _PPV]),
    "tm_harmonic_angle_create": (c_int, [c_int, _PI, _PI, _PI, _PPV]),
    "tm_periodic_torsion_create": (c_int, [c_int, _PI, _PI, _PI, _PPV]),
    "tm_summed_potential_create": (c_int, [c_int, _PPV, _PI, _PPV]),
    "tm_fanout_summed_potential_create": (c_int, [c_int, _PPV, _PPV]),
    "tm_potential_release": (None, [_P]),
    "tm_potential_execute_device": (
        c_int,
        [_P, c_int, c_int, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "tm_potential_execute_device_noalloc": (
        c_int,
        [_P, c_int, c_int, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "tm_potential_execute_host": (
        c_int,
        [_P, c_int, c_int, _PD, _PD, _PD, c_double, _PU64, _PU64, _PU64, _PU64],
    ),
    "tm_potential_execute_batch_host": (
        c_int,
        [_P, c_int, c_int, c_int, c_int, c_int, _PD, _PD, _PD, _PD, _PU64, _PU64, _PU64, _PU64],
    ),
    "tm_potential_du_dp_fixed_to_float": (c_int, [_P, c_int, c_int, _PU64, _PD]),
    "tm_potential_execute_dl_dp_device": (
        c_int,
        [_P, c_int, c_int, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p],
    ),
    "tm_potential_execute_dl_dp_host": (c_int, [_P, c_int, c_int, _PD, _PD, _PD, c_double, _PU64]),
    "tm_potential_execute_hvp_device": (
        c_int,
        [_P, c_int, c_int, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p, c_double, c_void_p, c_void_p,
         c_void_p, c_void_p],
    ),
    "tm_potential_execute_hvp_host": (
        c_int,
        [_P, c_int, c_int, _PD, _PD, _PD, c_double, _PD, _PD, c_double, _PD, _PD, _PD],
    ),
    "tm_bound_potential_create": (c_int, [_P, c_int, _PI, _PD, _PPV]),
    "tm_bound_potential_release": (None, [_P]),
    "tm_bound_potential_size": (c_int, [_P, POINTER(c_int)]),
    "tm_bound_potential_execute_host": (c_int, [_P, c_int, _PD, _PD, c_double, _PU64, _PU64, _PU64]),
    "tm_neighborlist_create": (c_int, [c_int, _PPV]),
    "tm_neighborlist_release": (None, [_P]),
    "tm_neighborlist_resize": (c_int, [_P, c_int]),
    "tm_neighborlist_build_host": (c_int, [_P, c_int, _PD, _PD, c_double, POINTER(c_int)]),
    "tm_neighborlist_get_tiles": (c_int, [_P, c_int, _PI, POINTER(c_uint32)]),
    "tm_neighborlist_compute_block_bounds_host": (c_int, [_P, c_int, _PD, _PD, _PD, _PD]),
    "tm_neighborlist_set_row_idxs": (c_int, [_P, c_int, POINTER(c_uint32)]),
    "tm_neighborlist_reset_row_idxs": (c_int, [_P]),
    "tm_neighborlist_set_pieces": (c_int, [_P, c_int]),
    "tm_langevin_integrator_create": (c_int, [c_int, c_double, c_double, _PD, _PD, c_uint64, _PPV]),
    "tm_integrator_release": (None, [_P]),
    "tm_barostat_create": (
        c_int,
        [c_int, c_double, c_double, c_int, _PI, _PI, c_int, c_int, _PPV, c_uint64, _PPV],
    ),
    "tm_barostat_set_interval": (c_int, [_P, c_int]),
    "tm_barostat_get_interval": (c_int, [_P, POINTER(c_int)]),
    "tm_barostat_set_pressure": (c_int, [_P, c_double]),
    "tm_barostat_release": (None, [_P]),
    "tm_context_create": (c_int, [c_int, _PD, _PD, _PD, _P, c_int, _PPV, _P, _PPV]),
    "tm_context_release": (None, [_P]),
    "tm_context_num_atoms": (c_int, [_P, POINTER(c_int)]),
    "tm_context_step": (c_int, [_P, c_double]),
    "tm_context_multiple_steps": (c_int, [_P, c_int, _PD, c_int, c_int, _PD, _PD, _PD]),
    "tm_context_multiple_steps_U": (c_int, [_P, c_double, c_int, c_int, _PD, c_int, c_int, _PD, _PD, _PD]),
    "tm_batch_multiple_steps_U": (c_int, [_P, c_int, c_int, _PD, c_int, c_int, _PD, _PD, _PD, c_int]),
    "tm_context_get_x_t": (c_int, [_P, _PD]),
    "tm_context_get_v_t": (c_int, [_P, _PD]),
    "tm_context_get_box": (c_int, [_P, _PD]),
    "tm_context_set_x_t": (c_int, [_P, _PD]),
    "tm_context_get_du_dx_t_minus_1": (c_int, [_P, _PU64]),
    "tm_batch_create": (c_int, [c_int, _PPV, _PD, _PPV]),
    "tm_batch_release": (None, [_P]),
    "tm_batch_num_systems": (c_int, [_P, POINTER(c_int)]),
    "tm_batch_total_atoms": (c_int, [_P, POINTER(c_int)]),
    "tm_batch_multiple_steps": (c_int, [_P, c_int, _PD, c_int, c_int, _PD, _PD, _PD, c_int]),
    "tm_batch_energies": (c_int, [_P, _PD, _PD]),
    "tm_batch_fire_minimize": (c_int, [_P, c_int, _PD, _PI]),
    "tm_batch_max_force": (c_int, [_P, _PD]),
    "tm_fire_minimize": (c_int, [c_int, _PD, _PD, c_int, _PPV, c_int, _PD, _PD]),
    "tm_batch_get_x": (c_int, [_P, c_int, _PD]),
    "tm_batch_get_v": (c_int, [_P, c_int, _PD]),
    "tm_batch_get_box": (c_int, [_P, c_int, _PD]),
    "tm_batch_write_back": (c_int, [_P]),
    "tm_batch_repack": (c_int, [_P, c_int, _PI, c_int, _PPV, _PD, _PPV]),
    "tm_batch_create_windows": (
        c_int,
        [c_int, c_int, _PD, _PD, _PD, _P, _PU64, c_int, _PPV, _P, _PU64, _PD, _PPV],
    ),
    "tm_batch_gather": (c_int, [c_int, _PPV, _PI, _PD, _PPV]),
    "tm_batch_device_bytes": (c_int, [_P, _PU64]),
    "tm_batch_set_tile_grid": (c_int, [_P, c_int]),
    "tm_batch_set_tile_waves": (c_int, [_P, c_int]),
    "tm_batch_set_build_pieces": (c_int, [_P, c_int]),
    "tm_batch_set_nblist_padding": (c_int, [_P, c_double]),
    "tm_batch_set_grad_observer": (c_int, [_P, c_int]),
    "tm_batch_observe": (c_int, [_P]),
    "tm_batch_grad_size": (c_int, [_P, c_int, POINTER(c_int)]),
    "tm_batch_grad_stats": (c_int, [_P, c_int, POINTER(c_int), _PD, _PD, _PD, _PD, _PD]),
    "tm_batch_adjoint_forward": (c_int, [_P, c_int, c_int, c_int, _PD]),
    "tm_batch_adjoint_backward": (c_int, [_P, _PD, _PD, _PD, _PD, c_int]),
    "tm_batch_profile": (c_int, [_P, c_int, _PD]),
    "tm_batch_get_barostat_move": (c_int, [_P, c_int, _PD]),
    "tm_batch_get_status": (c_int, [_P, _PI]),
    "tm_batch_set_raise_faults": (c_int, [_P, c_int]),
    "tm_batch_total_tiles": (c_int, [_P, POINTER(c_int)]),
}

_lib = None


def load() -> ctypes.CDLL:
    """Loads the library once; raises (no fallback) when it was not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m timemachine_amd.build` "
                "(there is no CPU fallback for the HIP path)"
            )
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def declared_symbols() -> list[str]:
    return list(_SIGS)


def check(rc: int) -> None:
    if rc != 0:
        msg = load().tm_last_error()
        raise RuntimeError(msg.decode() if msg else f"tm_amd error {rc}")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))


def device_count() -> int:
    n = c_int(0)
    call("tm_device_count", ctypes.byref(n))
    return n.value


# ---- numpy helpers -----------------------------------------------------------


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def pd(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(_PD)


def pi(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(_PI)


def pu64(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(_PU64)


def handle_array(handles) -> ctypes.Array:
    arr = (c_void_p * max(len(handles), 1))()
    for i, h in enumerate(handles):
        arr[i] = h.value if isinstance(h, c_void_p) else h
    return arr


FIXED_EXPONENT = 2.0**36


def fixed_to_float(v) -> np.ndarray | float:
    """uint64 fixed point (2^36) -> float64, as FIXED_TO_FLOAT (fixed_point.hpp:15-17)."""
    a = np.asarray(v, dtype=np.uint64).view(np.int64).astype(np.float64) / FIXED_EXPONENT
    return a if a.ndim else float(a)


def wrapping_sum(v: np.ndarray) -> np.uint64:
    """Sum of uint64 values modulo 2^64."""
    with np.errstate(over="ignore"):
        return np.sum(np.asarray(v, dtype=np.uint64), dtype=np.uint64)
