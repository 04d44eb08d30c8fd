"""Drop-in replacement for the reference's ``timemachine.lib.custom_ops``.

Class names, constructor signatures, methods, return shapes and error
messages follow the reference's pybind11 module
(timemachine/cpp/src/wrap_kernels.cpp, stubs in timemachine/lib/custom_ops.pyi)
so that ``lib.potentials``-style wrappers resolving ``type(name) + "_f32"``
keep working. Every call goes through the C ABI of ``libtm_amd.so``; there
is no CPU path.

Differences a caller can observe:
  * arithmetic is FP32 with 64-bit fixed-point accumulation; ``*_f64`` names
    exist but raise (the MI355X path is the FP32 one);
  * the neighborlist works on 64-atom blocks (one wave64 per tile);
    ``Neighborlist_f32.get_nblist`` returns the reference's 32-atom row
    blocks by default (``block_size=64`` gives the tiles' own blocks);
  * Langevin and barostat noise come from a counter-based generator
    (Philox4x32-10), so trajectories differ from curand's but are bitwise
    reproducible, checkpointable, and identical alone or in a batch.
"""

from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import numpy as np

from timemachine_amd import _lib
from timemachine_amd._lib import call, f64, fixed_to_float, i32, pd, pi, pu64, wrapping_sum

FIXED_EXPONENT_DU_DSIG = 2.0**37
FIXED_EXPONENT_DU_DEPS = 2.0**38

# per-system fault bits of BatchedContexts.status()
STATUS_BOX_TOO_SMALL = 1
STATUS_BOX_NOT_ORTHO = 2
STATUS_TILE_OVERFLOW = 4
STATUS_NONFINITE = 8
STATUS_FIXED_INVALID = 16
STATUS_MESSAGES = {
    STATUS_BOX_TOO_SMALL: "Cutoff with padding is more than half of the box width, neighborlist is no longer reliable",
    STATUS_BOX_NOT_ORTHO: "Provided non-ortholinear box, unable to compute nonbonded energy",
    STATUS_TILE_OVERFLOW: "neighborlist tile buffer overflow",
    STATUS_NONFINITE: "coordinates or velocities are not finite",
    STATUS_FIXED_INVALID: "a force left the fixed-point range (non-finite or huge pair terms)",
}


def status_message(bits: int) -> str:
    return "; ".join(msg for bit, msg in STATUS_MESSAGES.items() if bits & bit)


class _Handle:
    _release = None

    def __init__(self):
        self._h = ctypes.c_void_p()
        self._keep = []

    @property
    def handle(self):
        return self._h

    def __del__(self):
        # At interpreter shutdown module globals may already be None.
        try:
            h = getattr(self, "_h", None)
            rel = type(self)._release
            lib = _lib._lib
        except (AttributeError, TypeError):
            return
        if h is not None and h.value and rel is not None and lib is not None:
            getattr(lib, rel)(h)
            self._h = ctypes.c_void_p()


# ---------------------------------------------------------------- potentials


class Potential(_Handle):
    """Base class; mirrors the reference's Potential binding (wrap_kernels.cpp:334-652)."""

    _release = "tm_potential_release"

    def _du_dp_to_float(self, N: int, P: int, du_dp_u64: np.ndarray, shape) -> np.ndarray:
        out = np.empty(P, dtype=np.float64)
        if P:
            call("tm_potential_du_dp_fixed_to_float", self._h, N, P, pu64(du_dp_u64), pd(out))
        return out.reshape(shape)

    def execute_selective(self, coords, params, box, lam, compute_du_dx, compute_du_dp, compute_du_dl, compute_u):
        coords = f64(coords)
        params = f64(params)
        box = f64(box)
        N = coords.shape[0]
        P = params.size
        du_dx = np.zeros(N * 3, dtype=np.uint64) if compute_du_dx else None
        du_dp = np.zeros(max(P, 1), dtype=np.uint64) if compute_du_dp else None
        du_dl = np.zeros(max(N, 1), dtype=np.uint64) if compute_du_dl else None
        u = np.zeros(max(N, 1), dtype=np.uint64) if compute_u else None
        call(
            "tm_potential_execute_host",
            self._h,
            N,
            P,
            pd(coords),
            pd(params.reshape(-1)) if P else None,
            pd(box),
            float(lam),
            pu64(du_dx),
            pu64(du_dp),
            pu64(du_dl),
            pu64(u),
        )
        r_dx = fixed_to_float(du_dx).reshape(N, 3) if compute_du_dx else None
        r_dp = self._du_dp_to_float(N, P, du_dp, params.shape) if compute_du_dp else None
        r_dl = fixed_to_float(wrapping_sum(du_dl)) if compute_du_dl else None
        r_u = fixed_to_float(wrapping_sum(u)) if compute_u else None
        return r_dx, r_dp, r_dl, r_u

    def execute(self, coords, params, box, lam):
        return self.execute_selective(coords, params, box, lam, True, True, True, True)

    def execute_fixed_raw(self, coords, params, box, lam):
        """Raw 64-bit fixed-point outputs (du_dx [N*3], du_dp [P], du_dl [N], u [N]) before decoding."""
        coords = f64(coords)
        params = f64(params)
        box = f64(box)
        N = coords.shape[0]
        P = params.size
        out = [np.zeros(N * 3, np.uint64), np.zeros(max(P, 1), np.uint64), np.zeros(max(N, 1), np.uint64),
               np.zeros(max(N, 1), np.uint64)]
        call("tm_potential_execute_host", self._h, N, P, pd(coords), pd(params.reshape(-1)) if P else None, pd(box),
             float(lam), *[pu64(a) for a in out])
        return out

    def execute_du_dx(self, coords, params, box, lam):
        return self.execute_selective(coords, params, box, lam, True, False, False, False)[0]

    def execute_dl_dp(self, coords, params, box, lam):
        """d(du/dlambda)/dp, shaped like params. Extension of this library (the
        reference obtains it from its JAX potentials); used for gradients of
        lambda-window averages of du/dlambda."""
        coords = f64(coords)
        params = f64(params)
        box = f64(box)
        N = coords.shape[0]
        P = params.size
        out = np.zeros(max(P, 1), dtype=np.uint64)
        call("tm_potential_execute_dl_dp_host", self._h, N, P, pd(coords), pd(params.reshape(-1)) if P else None,
             pd(box), float(lam), pu64(out))
        return self._du_dp_to_float(N, P, out, params.shape)

    def execute_hvp(self, coords, params, box, lam, v_coords=None, v_params=None, v_lam=0.0,
                    compute_hx=True, compute_hp=True, compute_hl=True):
        """Hessian-vector product: the directional derivatives of (du_dx, du_dp,
        du_dl) along the tangent (v_coords [N, 3], v_params shaped like params,
        v_lam), i.e. H @ (v_coords, v_params, v_lam) with H the Hessian of U in
        (coords, params, lambda). By the symmetry of H this is also the
        vector-Jacobian product of the gradient with the tangent, which is what
        reverse-mode differentiation through forces needs. None tangents are
        zero. Returns (hx [N, 3], hp shaped like params, hl float), None for
        outputs not computed. Extension of this library (the reference takes
        second derivatives from its JAX potentials)."""
        coords = f64(coords)
        params = f64(params)
        box = f64(box)
        N = coords.shape[0]
        P = params.size
        vx = None if v_coords is None else f64(v_coords).reshape(N, 3)
        vp = None if v_params is None or P == 0 else f64(v_params).reshape(P)
        hx = np.zeros(N * 3, np.float64) if compute_hx else None
        hp = np.zeros(max(P, 1), np.float64) if compute_hp else None
        hl = np.zeros(1, np.float64) if compute_hl else None
        call("tm_potential_execute_hvp_host", self._h, N, P, pd(coords), pd(params.reshape(-1)) if P else None,
             pd(box), float(lam), None if vx is None else pd(vx), None if vp is None else pd(vp), float(v_lam),
             pd(hx), pd(hp), pd(hl))
        r_x = hx.reshape(N, 3) if compute_hx else None
        r_p = hp[:P].reshape(params.shape) if compute_hp else None
        r_l = float(hl[0]) if compute_hl else None
        return r_x, r_p, r_l

    def execute_selective_batch(
        self, coords, params, boxes, lambs, compute_du_dx, compute_du_dp, compute_du_dl, compute_u
    ):
        coords = f64(coords)
        params = f64(params)
        boxes = f64(boxes)
        lambs = f64(lambs).reshape(-1)
        if coords.ndim != 3 and boxes.ndim != 3:
            raise RuntimeError("coords and boxes must have 3 dimensions")
        if coords.shape[0] != boxes.shape[0]:
            raise RuntimeError("number of batches of coords and boxes don't match")
        if params.ndim < 2:
            raise RuntimeError("parameters must have at least 2 dimensions")
        if lambs.size < 1:
            raise RuntimeError("lambdas must have at least 1 value")
        Bc, N = coords.shape[0], coords.shape[1]
        Bp = params.shape[0]
        P = params.size // Bp
        Bl = lambs.size
        T = Bc * Bp * Bl
        du_dx = np.zeros(T * N * 3, dtype=np.uint64) if compute_du_dx else None
        du_dp = np.zeros(max(T * P, 1), dtype=np.uint64) if compute_du_dp else None
        du_dl = np.zeros(T * N, dtype=np.uint64) if compute_du_dl else None
        u = np.zeros(T * N, dtype=np.uint64) if compute_u else None
        call(
            "tm_potential_execute_batch_host",
            self._h,
            Bc,
            N,
            Bp,
            P,
            Bl,
            pd(coords),
            pd(params.reshape(-1)) if P else None,
            pd(boxes),
            pd(lambs),
            pu64(du_dx),
            pu64(du_dp),
            pu64(du_dl),
            pu64(u),
        )
        out = [None, None, None, None]
        if compute_du_dx:
            out[0] = fixed_to_float(du_dx).reshape(Bc, Bp, Bl, N, 3)
        if compute_du_dp:
            flat = np.stack([self._du_dp_to_float(N, P, du_dp[i * P : (i + 1) * P], (P,)) for i in range(T)])
            out[1] = flat.reshape((Bc, Bp, Bl) + params.shape[1:])
        if compute_du_dl:
            out[2] = np.array([fixed_to_float(wrapping_sum(du_dl[i * N : (i + 1) * N])) for i in range(T)]).reshape(
                Bc, Bp, Bl
            )
        if compute_u:
            out[3] = np.array([fixed_to_float(wrapping_sum(u[i * N : (i + 1) * N])) for i in range(T)]).reshape(
                Bc, Bp, Bl
            )
        return tuple(out)


_BONDED_SIZE_MSG = {
    2: "bond_idxs.size() must be exactly 2*k!",
    3: "angle_idxs.size() must be exactly 3*A",
    4: "torsion_idxs.size() must be exactly 4*k",
}


def _bonded(create, idxs, arity, lamb_mult, lamb_offset):
    """Validates in the reference's order (harmonic_bond.cu:15-33) and creates the potential."""
    idxs = i32(idxs).reshape(-1)
    if idxs.size % arity != 0:
        raise RuntimeError(_BONDED_SIZE_MSG[arity])
    n = idxs.size // arity
    mult = None if lamb_mult is None else i32(lamb_mult).reshape(-1)
    off = None if lamb_offset is None else i32(lamb_offset).reshape(-1)
    if mult is not None and mult.size > 0 and mult.size != n:
        raise RuntimeError("bad lambda_mult size()")
    if off is not None and off.size > 0 and off.size != n:
        raise RuntimeError("bad lambda_offset size()")
    h = ctypes.c_void_p()
    call(
        create,
        n,
        pi(idxs),
        pi(mult) if mult is not None and mult.size else None,
        pi(off) if off is not None and off.size else None,
        ctypes.byref(h),
    )
    return h, [idxs, mult, off]


class HarmonicBond_f32(Potential):
    def __init__(self, bond_idxs, lamb_mult=None, lamb_offset=None):
        super().__init__()
        self._h, self._keep = _bonded("tm_harmonic_bond_create", bond_idxs, 2, lamb_mult, lamb_offset)


class HarmonicAngle_f32(Potential):
    def __init__(self, angle_idxs, lamb_mult=None, lamb_offset=None):
        super().__init__()
        self._h, self._keep = _bonded("tm_harmonic_angle_create", angle_idxs, 3, lamb_mult, lamb_offset)


class PeriodicTorsion_f32(Potential):
    def __init__(self, angle_idxs, lamb_mult=None, lamb_offset=None):
        super().__init__()
        self._h, self._keep = _bonded("tm_periodic_torsion_create", angle_idxs, 4, lamb_mult, lamb_offset)


_IDENTITY_TRANSFORMS = ("lambda",)


def _check_transforms(*exprs):
    for e in exprs:
        if e not in _IDENTITY_TRANSFORMS:
            raise RuntimeError(
                f"lambda transform {e!r} is not supported: only the identity transform 'lambda' is compiled in"
            )


class NonbondedAllPairs_f32(Potential):
    """NonbondedAllPairs (wrap_kernels.cpp:835-895); kernel_dir is accepted and ignored."""

    _create = "tm_nonbonded_all_pairs_create"

    def __init__(
        self,
        kernel_dir,
        lambda_plane_idxs_i,
        lambda_offset_idxs_i,
        beta,
        cutoff,
        atom_idxs_i=None,
        transform_lambda_charge="lambda",
        transform_lambda_sigma="lambda",
        transform_lambda_epsilon="lambda",
        transform_lambda_w="lambda",
    ):
        super().__init__()
        _check_transforms(transform_lambda_charge, transform_lambda_sigma, transform_lambda_epsilon, transform_lambda_w)
        plane = i32(lambda_plane_idxs_i)
        offset = i32(lambda_offset_idxs_i)
        N = offset.size
        if plane.size != offset.size:
            raise RuntimeError("lambda offset idxs and plane idxs need to be equivalent")
        atoms = None
        n_atoms = -1
        if atom_idxs_i is not None:
            atoms = i32(atom_idxs_i)
            if len(set(atoms.tolist())) != atoms.size:
                raise RuntimeError("atom indices must be unique")
            n_atoms = atoms.size
        call(
            self._create,
            N,
            pi(plane),
            pi(offset),
            float(beta),
            float(cutoff),
            pi(atoms),
            n_atoms,
            ctypes.byref(self._h),
        )
        self._keep = [plane, offset, atoms]

    def set_nblist_padding(self, padding):
        call("tm_nonbonded_all_pairs_set_nblist_padding", self._h, float(padding))

    def disable_hilbert_sort(self):
        call("tm_nonbonded_all_pairs_disable_hilbert_sort", self._h)

    def num_tiles(self) -> int:
        n = ctypes.c_int(0)
        call("tm_nonbonded_all_pairs_num_tiles", self._h, ctypes.byref(n))
        return n.value


class NonbondedPairList_f32(Potential):
    """NonbondedPairList (wrap_kernels.cpp:957-1014)."""

    _negated = False
    _create = "tm_nonbonded_pair_list_create"

    def __init__(
        self,
        kernel_dir,
        pair_idxs_i,
        scales_i,
        lambda_plane_idxs_i,
        lambda_offset_idxs_i,
        beta,
        cutoff,
        transform_lambda_charge="lambda",
        transform_lambda_sigma="lambda",
        transform_lambda_epsilon="lambda",
        transform_lambda_w="lambda",
    ):
        super().__init__()
        _check_transforms(transform_lambda_charge, transform_lambda_sigma, transform_lambda_epsilon, transform_lambda_w)
        pairs = i32(pair_idxs_i).reshape(-1)
        scales = f64(scales_i).reshape(-1)
        plane = i32(lambda_plane_idxs_i)
        offset = i32(lambda_offset_idxs_i)
        if pairs.size % 2 != 0:
            raise RuntimeError(f"pair_idxs.size() must be even, but got {pairs.size}")
        M = pairs.size // 2
        if scales.size // 2 != M:
            raise RuntimeError(f"expected same number of pairs and scale tuples, but got {M} != {scales.size // 2}")
        call(
            self._create,
            offset.size,
            M,
            pi(pairs),
            pd(scales),
            pi(plane),
            pi(offset),
            float(beta),
            float(cutoff),
            1 if self._negated else 0,
            ctypes.byref(self._h),
        )
        self._keep = [pairs, scales, plane, offset]


class NonbondedPairList_f32_negated(NonbondedPairList_f32):
    _negated = True


class NonbondedAllPairs_f32_interpolated(NonbondedAllPairs_f32):
    """NonbondedAllPairs with parameters [2N, 3] (initial, final) interpolated as
    (1 - lambda) p0 + lambda p1 (identity transforms only)."""

    _create = "tm_nonbonded_all_pairs_interpolated_create"


class NonbondedPairList_f32_interpolated(NonbondedPairList_f32):
    _create = "tm_nonbonded_pair_list_interpolated_create"


class NonbondedPairList_f32_negated_interpolated(NonbondedPairList_f32):
    _negated = True
    _create = "tm_nonbonded_pair_list_interpolated_create"


class NonbondedInteractionGroup_f32(NonbondedAllPairs_f32):
    """NonbondedInteractionGroup (wrap_kernels.cpp:896-954): pairs (i, j) with i among
    the row atoms and j among the others."""

    _interpolated = 0

    def __init__(
        self,
        kernel_dir,
        row_atom_idxs_i,
        lambda_plane_idxs_i,
        lambda_offset_idxs_i,
        beta,
        cutoff,
        transform_lambda_charge="lambda",
        transform_lambda_sigma="lambda",
        transform_lambda_epsilon="lambda",
        transform_lambda_w="lambda",
    ):
        Potential.__init__(self)
        _check_transforms(transform_lambda_charge, transform_lambda_sigma, transform_lambda_epsilon, transform_lambda_w)
        rows = i32(row_atom_idxs_i)
        plane = i32(lambda_plane_idxs_i)
        offset = i32(lambda_offset_idxs_i)
        if plane.size != offset.size:
            raise RuntimeError("lambda offset idxs and plane idxs need to be equivalent")
        call(
            "tm_nonbonded_interaction_group_create",
            offset.size,
            rows.size,
            pi(rows),
            pi(plane),
            pi(offset),
            float(beta),
            float(cutoff),
            self._interpolated,
            ctypes.byref(self._h),
        )
        self._keep = [rows, plane, offset]


class NonbondedInteractionGroup_f32_interpolated(NonbondedInteractionGroup_f32):
    _interpolated = 1


class _CompositePotential(Potential):
    def get_potentials(self) -> List[Potential]:
        return list(self._pots)


class FanoutSummedPotential(_CompositePotential):
    """FanoutSummedPotential (wrap_kernels.cpp:1052-1063)."""

    def __init__(self, potentials: Sequence[Potential]):
        super().__init__()
        self._pots = list(potentials)
        arr = _lib.handle_array([p.handle for p in self._pots])
        call("tm_fanout_summed_potential_create", len(self._pots), arr, ctypes.byref(self._h))


class SummedPotential(_CompositePotential):
    """SummedPotential (wrap_kernels.cpp:1035-1050)."""

    def __init__(self, potentials: Sequence[Potential], params_sizes: Sequence[int]):
        super().__init__()
        self._pots = list(potentials)
        if len(self._pots) != len(params_sizes):
            raise RuntimeError("number of potentials != number of parameter sizes")
        sizes = i32(params_sizes)
        arr = _lib.handle_array([p.handle for p in self._pots])
        call("tm_summed_potential_create", len(self._pots), arr, pi(sizes), ctypes.byref(self._h))
        self._keep = [sizes]


def _f64_unavailable(name):
    class _Unavailable:
        def __init__(self, *args, **kwargs):
            raise NotImplementedError(
                f"{name}: the MI355X path has one precision (FP64 pair arithmetic with FP32 parameters and "
                "64-bit fixed-point accumulation, more accurate than the reference's f32 kernels); use the _f32 class"
            )

    _Unavailable.__name__ = name
    return _Unavailable


HarmonicBond_f64 = _f64_unavailable("HarmonicBond_f64")
HarmonicAngle_f64 = _f64_unavailable("HarmonicAngle_f64")
PeriodicTorsion_f64 = _f64_unavailable("PeriodicTorsion_f64")
NonbondedAllPairs_f64 = _f64_unavailable("NonbondedAllPairs_f64")
NonbondedPairList_f64 = _f64_unavailable("NonbondedPairList_f64")
NonbondedPairList_f64_negated = _f64_unavailable("NonbondedPairList_f64_negated")
NonbondedAllPairs_f64_interpolated = _f64_unavailable("NonbondedAllPairs_f64_interpolated")
NonbondedInteractionGroup_f64 = _f64_unavailable("NonbondedInteractionGroup_f64")
NonbondedInteractionGroup_f64_interpolated = _f64_unavailable("NonbondedInteractionGroup_f64_interpolated")
NonbondedPairList_f64_interpolated = _f64_unavailable("NonbondedPairList_f64_interpolated")
NonbondedPairList_f64_negated_interpolated = _f64_unavailable("NonbondedPairList_f64_negated_interpolated")
Neighborlist_f64 = _f64_unavailable("Neighborlist_f64")


class BoundPotential(_Handle):
    """BoundPotential (wrap_kernels.cpp:654-723)."""

    _release = "tm_bound_potential_release"

    def __init__(self, potential: Potential, params):
        super().__init__()
        params = f64(params)
        shape = i32(params.shape)
        self._potential = potential
        self._params = params
        call(
            "tm_bound_potential_create",
            potential.handle,
            shape.size,
            pi(shape),
            pd(params.reshape(-1)) if params.size else None,
            ctypes.byref(self._h),
        )

    def get_potential(self) -> Potential:
        return self._potential

    def size(self) -> int:
        n = ctypes.c_int(0)
        call("tm_bound_potential_size", self._h, ctypes.byref(n))
        return n.value

    @property
    def params(self) -> np.ndarray:
        return self._params

    def _execute_raw(self, coords, box, lam):
        coords = f64(coords)
        box = f64(box)
        N = coords.shape[0]
        du_dx = np.zeros(N * 3, dtype=np.uint64)
        du_dl = np.zeros(max(N, 1), dtype=np.uint64)
        u = np.zeros(max(N, 1), dtype=np.uint64)
        call(
            "tm_bound_potential_execute_host",
            self._h,
            N,
            pd(coords),
            pd(box),
            float(lam),
            pu64(du_dx),
            pu64(du_dl),
            pu64(u),
        )
        return du_dx, du_dl, u

    def execute(self, coords, box, lam):
        du_dx, du_dl, u = self._execute_raw(coords, box, lam)
        N = np.asarray(coords).shape[0]
        return fixed_to_float(du_dx).reshape(N, 3), fixed_to_float(wrapping_sum(du_dl)), fixed_to_float(wrapping_sum(u))

    def execute_fixed(self, coords, box, lam) -> np.ndarray:
        _, _, u = self._execute_raw(coords, box, lam)
        return np.array([wrapping_sum(u)], dtype=np.uint64)


# ---------------------------------------------------------------- neighborlist


class Neighborlist_f32(_Handle):
    """Neighborlist (wrap_kernels.cpp:29-87) on 64-atom blocks."""

    _release = "tm_neighborlist_release"
    BLOCK = 64

    def __init__(self, N: int):
        super().__init__()
        self._N = int(N)
        self._rows = 0  # row atoms of a row subset, 0 = all pairs
        self._row_idxs = np.zeros(0, dtype=np.int64)
        call("tm_neighborlist_create", self._N, ctypes.byref(self._h))

    def resize(self, size: int):
        call("tm_neighborlist_resize", self._h, int(size))
        self._N = int(size)
        self._rows = 0

    def reset_row_idxs(self):
        call("tm_neighborlist_reset_row_idxs", self._h)
        self._rows = 0

    def set_pieces(self, pieces: int):
        """Waves per row block of later builds (0 = automatic); the tile set does not change."""
        call("tm_neighborlist_set_pieces", self._h, int(pieces))

    def set_row_idxs(self, idxs):
        """Later builds pair only these atoms (rows, blocks of 64 in the order given)
        with every other atom (neighborlist.cu:215-243)."""
        idxs = np.ascontiguousarray(np.asarray(idxs).reshape(-1), dtype=np.uint32)
        call("tm_neighborlist_set_row_idxs", self._h, idxs.size,
             idxs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
        self._rows = idxs.size
        self._row_idxs = idxs.astype(np.int64)

    def _build(self, coords, box, cutoff):
        coords = f64(coords)
        box = f64(box)
        n = ctypes.c_int(0)
        call("tm_neighborlist_build_host", self._h, coords.shape[0], pd(coords), pd(box), float(cutoff), ctypes.byref(n))
        rows = np.zeros(max(n.value, 1), dtype=np.int32)
        atoms = np.zeros(max(n.value, 1) * self.BLOCK, dtype=np.uint32)
        call(
            "tm_neighborlist_get_tiles",
            self._h,
            max(n.value, 1),
            pi(rows),
            atoms.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
        )
        return rows[: n.value], atoms[: n.value * self.BLOCK].reshape(n.value, self.BLOCK)

    def get_tiles(self, coords, box, cutoff):
        """Raw tiles: (row block per tile [T], column atoms [T, 64], 0xFFFFFFFF = padding)."""
        return self._build(coords, box, cutoff)

    def get_nblist(self, coords, box, cutoff, block_size: int = 32) -> List[List[int]]:
        """Interacting column atoms of every row block (wrap_kernels.cpp:55-73):
        row blocks of ``block_size`` atoms (32, the reference's, or 64, the
        tiles' own) in atom order, or in the order of set_row_idxs. The 32-atom
        lists are the 64-atom block's columns kept by the build's own FP32
        predicate against each half's atoms (every such column is in the
        64-atom list: it is within the cutoff of one of the block's atoms)."""
        if block_size not in (32, self.BLOCK):
            raise RuntimeError(f"Block size must be 32 or {self.BLOCK}.")
        coords64 = f64(coords)
        N = coords64.shape[0]
        rows, atoms = self._build(coords64, box, cutoff)
        n_rows = self._rows or N
        n_blocks = (n_rows + self.BLOCK - 1) // self.BLOCK
        out: List[List[int]] = [[] for _ in range(n_blocks)]
        for r, a in zip(rows, atoms):
            out[int(r)].extend(int(j) for j in a if j < N)
        if block_size == self.BLOCK:
            return out
        # the build's predicate: FP32 coordinates, d - b rint(d / b) per axis
        # with b = float(box), then dx^2 + dy^2 + dz^2 < float(cutoff)^2
        x32 = coords64.astype(np.float32)
        b = np.diagonal(f64(box)).astype(np.float32)
        ib = (np.float32(1.0) / b).astype(np.float32)
        cut = np.float32(cutoff)
        cut2 = np.float32(cut * cut)
        row_atoms = self._row_idxs if self._rows else np.arange(N)
        halves: List[List[int]] = []
        for r in range((n_rows + 31) // 32):
            cols = np.asarray(out[r // 2], dtype=np.int64)
            ri = row_atoms[32 * r: min(32 * r + 32, n_rows)]
            if not self._rows:
                cols = cols[cols >= 32 * r]  # upper triangle from the half's first atom
            if cols.size == 0:
                halves.append([])
                continue
            d = x32[ri][:, None, :] - x32[cols][None, :, :]
            d = (d - b * np.rint(d * ib)).astype(np.float32)
            d2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
            keep = np.any(d2 < cut2, axis=0)
            halves.append(sorted(int(j) for j in cols[keep]))
        return halves

    def compute_block_bounds(self, coords, box, block_size):
        if block_size != self.BLOCK:
            raise RuntimeError(f"Block size must be {self.BLOCK}.")
        coords = f64(coords)
        box = f64(box)
        if coords.shape[1] != 3:
            raise RuntimeError("Dimensions must be 3")
        N = coords.shape[0]
        nb = (N - self._rows + self.BLOCK - 1) // self.BLOCK  # the column blocks of a row subset
        ctr = np.zeros((nb, 3))
        ext = np.zeros((nb, 3))
        call("tm_neighborlist_compute_block_bounds_host", self._h, N, pd(coords), pd(box), pd(ctr), pd(ext))
        return ctr, ext


# ---------------------------------------------------------------- integrator, barostat, context


class Integrator(_Handle):
    _release = "tm_integrator_release"


class LangevinIntegrator(Integrator):
    """LangevinIntegrator(dt, ca, cbs, ccs, seed) (wrap_kernels.cpp:314-332)."""

    def __init__(self, dt, ca, cbs, ccs, seed):
        super().__init__()
        cbs = f64(cbs)
        ccs = f64(ccs)
        self.N = cbs.size
        call(
            "tm_langevin_integrator_create",
            cbs.size,
            float(dt),
            float(ca),
            pd(cbs),
            pd(ccs),
            int(seed) & 0xFFFFFFFFFFFFFFFF,
            ctypes.byref(self._h),
        )


class MonteCarloBarostat(_Handle):
    """MonteCarloBarostat(N, pressure, temperature, group_idxs, interval, bps, seed) (wrap_kernels.cpp:1016-1033)."""

    _release = "tm_barostat_release"

    def __init__(self, N, pressure, temperature, group_idxs, interval, bps, seed):
        super().__init__()
        sizes = i32([len(g) for g in group_idxs])
        atoms = i32(np.concatenate([np.asarray(g, dtype=np.int32) for g in group_idxs]) if len(group_idxs) else [])
        self._bps = list(bps)
        arr = _lib.handle_array([bp.handle for bp in self._bps])
        call(
            "tm_barostat_create",
            int(N),
            float(pressure),
            float(temperature),
            sizes.size,
            pi(sizes),
            pi(atoms),
            int(interval),
            len(self._bps),
            arr,
            int(seed) & 0xFFFFFFFFFFFFFFFF,
            ctypes.byref(self._h),
        )

    def set_interval(self, interval):
        call("tm_barostat_set_interval", self._h, int(interval))

    def get_interval(self) -> int:
        n = ctypes.c_int(0)
        call("tm_barostat_get_interval", self._h, ctypes.byref(n))
        return n.value

    def set_pressure(self, pressure):
        call("tm_barostat_set_pressure", self._h, float(pressure))


class Context(_Handle):
    """Context(x0, v0, box, integrator, bps, barostat=None) (wrap_kernels.cpp:89-305)."""

    _release = "tm_context_release"

    def __init__(self, x0, v0, box, integrator, bps, barostat: Optional[MonteCarloBarostat] = None):
        super().__init__()
        x0 = f64(x0)
        v0 = f64(v0)
        box = f64(box)
        N = x0.shape[0]
        if v0.shape[0] != N:
            raise RuntimeError("v0 N != x0 N")
        if v0.shape[1] != x0.shape[1]:
            raise RuntimeError("v0 D != x0 D")
        self.N = N
        self._integrator = integrator
        self._bps = list(bps)
        self._barostat = barostat
        arr = _lib.handle_array([bp.handle for bp in self._bps])
        call(
            "tm_context_create",
            N,
            pd(x0),
            pd(v0),
            pd(box),
            integrator.handle,
            len(self._bps),
            arr,
            barostat.handle if barostat is not None else None,
            ctypes.byref(self._h),
        )

    def step(self, lam):
        call("tm_context_step", self._h, float(lam))

    def multiple_steps(self, lambda_schedule, store_du_dl_interval=0, store_x_interval=0):
        sched = f64(lambda_schedule).reshape(-1)
        n = sched.size
        dl_int = n if store_du_dl_interval <= 0 else int(store_du_dl_interval)
        x_int = n if store_x_interval <= 0 else int(store_x_interval)
        # an empty schedule resolves both intervals to 0 (context.cu:60-65)
        if dl_int <= 0:
            raise RuntimeError("store_du_dl_interval <= 0")
        if x_int <= 0:
            raise RuntimeError("store_x_interval <= 0")
        if n % x_int != 0:
            print("warning:: length of lambda_schedule modulo store_x_interval does not equal zero")
        if n % dl_int != 0:
            print("warning:: length of lambda_schedule modulo store_du_dl_interval does not equal zero")
        K = n // dl_int
        F = n // x_int
        du_dl = np.zeros(K)
        xs = np.zeros((F, self.N, 3))
        boxes = np.zeros((F, 3, 3))
        call(
            "tm_context_multiple_steps",
            self._h,
            n,
            pd(sched),
            dl_int,
            x_int,
            pd(du_dl),
            pd(xs),
            pd(boxes),
        )
        return du_dl, xs, boxes

    def multiple_steps_U(self, lamb, n_steps, lambda_windows, store_u_interval, store_x_interval):
        """Steps at lamb; energies of the state after every store_u_interval-th step at
        each lambda of lambda_windows (context.cu:139-233). Returns (u [F, K], xs, boxes)."""
        windows = f64(lambda_windows).reshape(-1)
        n_steps = int(n_steps)
        u_int = n_steps if store_u_interval <= 0 else int(store_u_interval)
        x_int = n_steps if store_x_interval <= 0 else int(store_x_interval)
        if u_int <= 0:
            raise RuntimeError("store_u_interval <= 0")
        if x_int <= 0:
            raise RuntimeError("store_x_interval <= 0")
        if n_steps % x_int != 0:
            print("warning:: n_steps modulo store_x_interval does not equal zero")
        if n_steps % u_int != 0:
            print("warning:: n_steps modulo store_u_interval does not equal zero")
        UF = n_steps // u_int
        F = n_steps // x_int
        u = np.zeros((UF, windows.size))
        xs = np.zeros((F, self.N, 3))
        boxes = np.zeros((F, 3, 3))
        call("tm_context_multiple_steps_U", self._h, float(lamb), n_steps, int(windows.size), pd(windows), u_int,
             x_int, pd(u), pd(xs), pd(boxes))
        return u, xs, boxes

    def set_x_t(self, x):
        x = f64(x)
        call("tm_context_set_x_t", self._h, pd(x))

    def get_x_t(self) -> np.ndarray:
        out = np.zeros((self.N, 3))
        call("tm_context_get_x_t", self._h, pd(out))
        return out

    def get_v_t(self) -> np.ndarray:
        out = np.zeros((self.N, 3))
        call("tm_context_get_v_t", self._h, pd(out))
        return out

    def get_box(self) -> np.ndarray:
        out = np.zeros((3, 3))
        call("tm_context_get_box", self._h, pd(out))
        return out

    def _get_du_dx_t_minus_1(self) -> np.ndarray:
        out = np.zeros(self.N * 3, dtype=np.uint64)
        call("tm_context_get_du_dx_t_minus_1", self._h, pu64(out))
        return fixed_to_float(out).reshape(self.N, 3)


class BatchedContexts(_Handle):
    """Steps many Contexts together in one engine (no reference counterpart).

    Each context keeps its own potentials, integrator, barostat and seed; the
    batch packs their current states into one HBM arena and advances all of
    them with the same kernel launches (captured into hipGraphs). A system's
    trajectory is bitwise identical to running its Context alone.
    """

    _release = "tm_batch_release"

    def __init__(self, contexts: Sequence[Context], lambdas: Sequence[float]):
        super().__init__()
        self._ctxs = list(contexts)
        lam = f64(lambdas).reshape(-1)
        if lam.size != len(self._ctxs):
            raise RuntimeError("one lambda per context is required")
        self._lam = lam
        arr = _lib.handle_array([c.handle for c in self._ctxs])
        call("tm_batch_create", len(self._ctxs), arr, pd(lam), ctypes.byref(self._h))
        self.sizes = [c.N for c in self._ctxs]

    def __len__(self):
        return len(self._ctxs)

    @classmethod
    def _wrap(cls, handle, contexts, lambdas, sizes=None):
        obj = cls.__new__(cls)
        _Handle.__init__(obj)
        obj._h = handle
        obj._ctxs = list(contexts)
        obj._lam = f64(lambdas).reshape(-1)
        obj.sizes = list(sizes) if sizes is not None else [c.N for c in obj._ctxs]
        return obj

    @classmethod
    def from_windows(cls, x0, v0s, box, integrator, seeds, bps, lambdas, barostat=None, barostat_seeds=None):
        """W lambda windows of one system in one engine, no Context per window.

        Window w starts from x0 and box with velocities v0s[w], the bound
        potentials ``bps``, ``integrator``'s coefficients with seed seeds[w] and,
        with a barostat, its settings with seed barostat_seeds[w]. Bitwise the
        batch of the W Contexts built from the same arguments."""
        x0 = f64(x0)
        N = x0.shape[0]
        lam = f64(lambdas).reshape(-1)
        W = lam.size
        v0s = f64(v0s).reshape(W, N, 3)
        seeds = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint64).reshape(-1))
        if seeds.size != W:
            raise RuntimeError("one integrator seed per window is required")
        bseeds = None
        if barostat is not None:
            bseeds = np.ascontiguousarray(np.asarray(barostat_seeds, dtype=np.uint64).reshape(-1))
            if bseeds.size != W:
                raise RuntimeError("one barostat seed per window is required")
        bps = list(bps)
        arr = _lib.handle_array([bp.handle for bp in bps])
        h = ctypes.c_void_p()
        call("tm_batch_create_windows", W, N, pd(x0), pd(v0s), pd(f64(box)), integrator.handle, pu64(seeds),
             len(bps), arr, barostat.handle if barostat is not None else None, pu64(bseeds), pd(lam),
             ctypes.byref(h))
        return cls._wrap(h, [None] * W, lam, sizes=[N] * W)

    @staticmethod
    def gather(parts, lambdas):
        """A new batch whose system i is system parts[i][1] of batch parts[i][0]
        (state copied on the device); the sources are left unchanged."""
        parts = list(parts)
        lam = f64(lambdas).reshape(-1)
        if lam.size != len(parts):
            raise RuntimeError("one lambda per system of the new batch is required")
        srcs = _lib.handle_array([b.handle for b, _ in parts])
        idx = i32([int(s) for _, s in parts])
        h = ctypes.c_void_p()
        call("tm_batch_gather", len(parts), srcs, pi(idx), pd(lam), ctypes.byref(h))
        return BatchedContexts._wrap(h, [b._ctxs[int(s)] for b, s in parts], lam,
                                     sizes=[b.sizes[int(s)] for b, s in parts])

    def repack(self, keep: Sequence[int], new_contexts: Sequence[Context], lambdas: Sequence[float]):
        """A new batch with systems ``keep`` of this one (state handed over on the
        device) followed by ``new_contexts``; lambdas for all of them in that order.
        This batch is left unchanged (release it when done)."""
        keep = i32(list(keep)).reshape(-1)
        new_contexts = list(new_contexts)
        lam = f64(lambdas).reshape(-1)
        if lam.size != keep.size + len(new_contexts):
            raise RuntimeError("one lambda per system of the new batch is required")
        arr = _lib.handle_array([c.handle for c in new_contexts]) if new_contexts else None
        h = ctypes.c_void_p()
        call("tm_batch_repack", self._h, int(keep.size), pi(keep) if keep.size else None, len(new_contexts), arr,
             pd(lam), ctypes.byref(h))
        return BatchedContexts._wrap(h, [self._ctxs[int(k)] for k in keep] + new_contexts, lam,
                                     sizes=[self.sizes[int(k)] for k in keep] + [c.N for c in new_contexts])

    def device_bytes(self) -> int:
        n = ctypes.c_uint64(0)
        call("tm_batch_device_bytes", self._h, ctypes.byref(n))
        return int(n.value)

    def set_tile_grid(self, grid: int):
        call("tm_batch_set_tile_grid", self._h, int(grid))

    def set_nblist_padding(self, padding: float):
        call("tm_batch_set_nblist_padding", self._h, float(padding))

    def set_build_pieces(self, pieces: int):
        """Waves per row block in the neighborlist build (1, 2, 4 or 8; 0 =
        automatic). Results do not depend on it."""
        call("tm_batch_set_build_pieces", self._h, int(pieces))

    def set_tile_waves(self, waves: int):
        """Waves per SIMD the pair kernel is compiled for: 4 (default), 3 or 2 (extension)."""
        call("tm_batch_set_tile_waves", self._h, int(waves))

    def autotune(self, steps: int = 45, waves=(4, 3, 2), grids=(32768, 65536), paddings=(0.075, 0.1, 0.15),
                 store_du_dl_interval: int = 1, use_graph: bool = True) -> dict:
        """Chooses the pair kernel's occupancy build and grid, then the
        neighborlist padding (extension).

        Every candidate is timed on a scratch copy of this batch made on the
        device (repack), so this batch's trajectories are untouched; results do
        not depend on any of the settings (fixed-point sums are
        order-independent and the tile kernel tests the exact cutoff), only
        the speed does. The occupancy build (4, 3 or 2 waves per SIMD) trades
        latency hiding against registers, the grid (one-wave workgroups)
        trades load balance against scheduling, padding trades neighborlist
        rebuilds against tiles. Stage 1 times waves x grids at the first
        padding in ``paddings`` that is 0.1 nm (else the first); stage 2 times
        the paddings at the stage-1 winner. Applies and returns the fastest
        setting."""
        import time

        scratch = self.repack(range(len(self)), [], self._lam)
        kw = dict(store_du_dl_interval=store_du_dl_interval, use_graph=use_graph)
        scratch.multiple_steps(steps, **kw)
        table = []

        def stage(name, cands):
            # The step time drifts while a freshly built system relaxes, so the
            # candidates are timed twice, in one order and then in the reverse
            # order, and ranked by their mean.
            times = {c: [] for c in cands}
            for sweep in (cands, cands[::-1]):
                for w, g, pad in sweep:
                    scratch.set_tile_waves(w)
                    scratch.set_tile_grid(g)
                    scratch.set_nblist_padding(pad)
                    scratch.multiple_steps(max(steps // 3, 1), **kw)
                    t0 = time.perf_counter()
                    scratch.multiple_steps(steps, **kw)
                    times[(w, g, pad)].append(1e3 * (time.perf_counter() - t0) / steps)
            rows = [{"stage": name, "waves": w, "grid": g, "padding": pad, "ms_per_step": sum(v) / len(v)}
                    for (w, g, pad), v in times.items()]
            table.extend(rows)
            return min(rows, key=lambda r: r["ms_per_step"])

        pad0 = 0.1 if any(abs(p - 0.1) < 1e-12 for p in paddings) else float(paddings[0])
        best = stage("occupancy x grid", [(int(w), int(g), pad0) for w in waves for g in grids])
        best = stage("padding", [(best["waves"], best["grid"], float(p)) for p in paddings])
        del scratch
        self.set_tile_waves(best["waves"])
        self.set_tile_grid(best["grid"])
        self.set_nblist_padding(best["padding"])
        return {"waves": best["waves"], "grid": best["grid"], "padding": best["padding"],
                "ms_per_step": best["ms_per_step"], "candidates": table}

    def multiple_steps(self, n_steps, lambda_schedules=None, store_du_dl_interval=0, store_x_interval=0, use_graph=True):
        S = len(self._ctxs)
        sched = None
        if lambda_schedules is not None:
            sched = f64(lambda_schedules).reshape(S, n_steps)
        K = n_steps // store_du_dl_interval if store_du_dl_interval > 0 else 0
        F = n_steps // store_x_interval if store_x_interval > 0 else 0
        du_dl = np.zeros((S, max(K, 0)))
        total = sum(self.sizes)
        xs = np.zeros(max(F * total * 3, 1))
        boxes = np.zeros((S, max(F, 0), 3, 3))
        call(
            "tm_batch_multiple_steps",
            self._h,
            int(n_steps),
            pd(sched),
            int(store_du_dl_interval),
            int(store_x_interval),
            pd(du_dl) if K else None,
            pd(xs) if F else None,
            pd(boxes) if F else None,
            1 if use_graph else 0,
        )
        frames = []
        off = 0
        for n in self.sizes:
            frames.append(xs[off : off + F * n * 3].reshape(F, n, 3))
            off += F * n * 3
        return du_dl, frames, boxes

    def multiple_steps_U(self, n_steps, lambda_windows, store_u_interval, store_x_interval=0, use_graph=True):
        """Every system steps at its own lambda; the state after every store_u_interval-th
        step is evaluated at each of lambda_windows. Returns (u [S, F, K], frames, boxes)."""
        S = len(self._ctxs)
        windows = f64(lambda_windows).reshape(-1)
        if store_u_interval <= 0:
            raise RuntimeError("store_u_interval <= 0")
        UF = n_steps // store_u_interval
        F = n_steps // store_x_interval if store_x_interval > 0 else 0
        u = np.zeros((S, UF, windows.size))
        total = sum(self.sizes)
        xs = np.zeros(max(F * total * 3, 1))
        boxes = np.zeros((S, max(F, 0), 3, 3))
        call("tm_batch_multiple_steps_U", self._h, int(n_steps), int(windows.size), pd(windows),
             int(store_u_interval), int(store_x_interval), pd(u), pd(xs) if F else None, pd(boxes) if F else None,
             1 if use_graph else 0)
        frames = []
        off = 0
        for n in self.sizes:
            frames.append(xs[off : off + F * n * 3].reshape(F, n, 3))
            off += F * n * 3
        return u, frames, boxes

    def energies(self):
        S = len(self._ctxs)
        u = np.zeros(S)
        dl = np.zeros(S)
        call("tm_batch_energies", self._h, pd(u), pd(dl))
        return u, dl

    def fire_minimize(self, lambda_schedules=None, n_steps=None, active=None):
        """FIRE minimisation of every system (or those with ``active`` set) on the
        device, in place (md/minimizer.py:35-73 for each system at once):
        ``lambda_schedules`` [S, n_sched] (the state starts at entry 0, one step per
        further entry), or ``n_steps`` steps at the current lambdas."""
        S = len(self._ctxs)
        if lambda_schedules is not None:
            sched = f64(lambda_schedules).reshape(S, -1)
            n_sched = sched.shape[1]
        else:
            if n_steps is None:
                raise RuntimeError("give lambda_schedules or n_steps")
            sched = None
            n_sched = int(n_steps) + 1
        act = None
        if active is not None:
            act = i32(np.asarray(active, dtype=bool).astype(np.int32)).reshape(-1)
            if act.size != S:
                raise RuntimeError("one active flag per system is required")
        call("tm_batch_fire_minimize", self._h, n_sched, pd(sched), pi(act))

    def max_force(self):
        """Largest per-atom norm of the total force of every system at its current
        state and lambda."""
        out = np.zeros(len(self._ctxs))
        call("tm_batch_max_force", self._h, pd(out))
        return out

    def get_x_t(self, s):
        out = np.zeros((self.sizes[s], 3))
        call("tm_batch_get_x", self._h, int(s), pd(out))
        return out

    def get_v_t(self, s):
        out = np.zeros((self.sizes[s], 3))
        call("tm_batch_get_v", self._h, int(s), pd(out))
        return out

    def get_box(self, s):
        out = np.zeros((3, 3))
        call("tm_batch_get_box", self._h, int(s), pd(out))
        return out

    def write_back(self):
        call("tm_batch_write_back", self._h)

    def status(self) -> np.ndarray:
        """Fault bits of every system so far (STATUS_* flags; 0 = healthy)."""
        out = np.zeros(len(self._ctxs), dtype=np.int32)
        call("tm_batch_get_status", self._h, pi(out))
        return out

    def set_raise_faults(self, raise_faults: bool):
        """False: multiple_steps returns with faulted systems flagged in status()
        instead of raising for the whole batch."""
        call("tm_batch_set_raise_faults", self._h, 1 if raise_faults else 0)

    def barostat_move(self, s: int) -> dict:
        """The barostat move system s made in the last step (all zero if none)."""
        out = np.zeros(9)
        call("tm_batch_get_barostat_move", self._h, int(s), pd(out))
        keys = ["triggered", "accepted", "u_before", "u_after", "delta_u", "delta_volume", "r2", "volume_scale",
                "length_scale"]
        return dict(zip(keys, out.tolist()))

    def profile(self, reps: int = 20) -> dict:
        """Per-kernel-group durations (ms) timed with HIP events on the batch's stream."""
        ms = np.zeros(4)
        call("tm_batch_profile", self._h, int(reps), pd(ms))
        return dict(zip(["prepare", "nonbonded", "pairlist_bonded", "integrator"], ms.tolist()))

    def total_tiles(self) -> int:
        n = ctypes.c_int(0)
        call("tm_batch_total_tiles", self._h, ctypes.byref(n))
        return n.value

    # -- gradient observer (extension): statistics for d<du/dlambda>/dp

    def set_grad_observer(self, interval: int):
        """Observe every `interval` steps of multiple_steps (at the du_dl store
        states); resets the sums. 0 turns it off."""
        call("tm_batch_set_grad_observer", self._h, int(interval))

    def observe(self):
        """One observation of every system at its current state."""
        call("tm_batch_observe", self._h)

    def grad_size(self, s: int) -> int:
        n = ctypes.c_int(0)
        call("tm_batch_grad_size", self._h, int(s), ctypes.byref(n))
        return n.value

    def grad_stats(self, s: int) -> dict:
        """Sums over observed frames of system s: n, A, A^2 (A = du/dlambda) and,
        per flat parameter (nonbonded [N,3], bond [B,2], angle [A,2], torsion
        [T,3]), g = du/dp, A g and h = d(du/dlambda)/dp."""
        P = self.grad_size(s)
        n = ctypes.c_int(0)
        a = ctypes.c_double(0.0)
        aa = ctypes.c_double(0.0)
        g, ag, h = np.zeros(P), np.zeros(P), np.zeros(P)
        call("tm_batch_grad_stats", self._h, int(s), ctypes.byref(n), ctypes.byref(a), ctypes.byref(aa), pd(g),
             pd(ag), pd(h))
        return {"n": n.value, "sum_a": a.value, "sum_aa": aa.value, "sum_g": g, "sum_ag": ag, "sum_h": h}

    # -- trajectory adjoint (extension): reverse mode through the Langevin steps

    def adjoint_forward(self, n_steps: int, du_dl_interval: int, checkpoint_interval: int) -> np.ndarray:
        """Run n_steps of every system (as multiple_steps without a lambda
        schedule) keeping checkpoints every checkpoint_interval steps, for
        adjoint_backward. Returns the stored du_dl [S, n_steps // du_dl_interval]."""
        F = n_steps // du_dl_interval if du_dl_interval > 0 else 0
        out = np.zeros((len(self), max(F, 1)), np.float64)
        call("tm_batch_adjoint_forward", self._h, int(n_steps), int(du_dl_interval), int(checkpoint_interval),
             pd(out))
        self._adj_frames = F
        return out[:, :F].copy()

    def adjoint_backward(self, dl_bar, want_initial_state: bool = False, horizon: int = 0):
        """Gradient of a loss on the du_dl values of the last adjoint_forward,
        given its cotangent dl_bar [S, F]: a list with each system's gradient
        in its flat parameters (grad_size(s) each, grad_stats order), and with
        want_initial_state also the cotangents of the initial coordinates and
        velocities ([total atoms, 3] each). horizon > 0 truncates the reverse
        sweep to blocks of that many steps."""
        S = len(self)
        F = getattr(self, "_adj_frames", None)
        if F is None:
            raise RuntimeError("Engine::adjoint_backward: call adjoint_forward first")
        dl_bar = f64(dl_bar)
        if dl_bar.size != S * F:
            raise RuntimeError(f"adjoint_backward: dl_bar needs {S} x {F} values (systems x stored du_dl frames)")
        dl_bar = dl_bar.reshape(S, F)
        sizes = [self.grad_size(s) for s in range(S)]
        grad = np.zeros(max(sum(sizes), 1))
        A = int(sum(self.sizes))
        xb = np.zeros((A, 3)) if want_initial_state else None
        vb = np.zeros((A, 3)) if want_initial_state else None
        call("tm_batch_adjoint_backward", self._h, pd(np.ascontiguousarray(dl_bar)), pd(grad), pd(xb), pd(vb),
             int(horizon))
        offs = np.cumsum([0] + sizes)
        grads = [grad[offs[s]:offs[s + 1]].copy() for s in range(S)]
        return (grads, xb, vb) if want_initial_state else grads
