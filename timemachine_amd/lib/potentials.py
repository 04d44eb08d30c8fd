"""Picklable potential wrappers with the interface of the reference's
``timemachine.lib.potentials`` (timemachine/lib/potentials.py:13-460):
``bind(params)``, ``unbound_impl(precision)`` resolving ``<Name>_f32`` in
:mod:`timemachine_amd.lib.custom_ops`, and ``bound_impl(precision)``.
"""

from __future__ import annotations

from typing import List

import numpy as np

from timemachine_amd.lib import custom_ops

_KERNEL_DIR = "<compiled-in>"  # the reference passes its kernel source directory here


def _suffix(precision) -> str:
    return "_f64" if precision == np.float64 else "_f32"


class CustomOpWrapper:
    def __init__(self, *args):
        self.args = list(args)
        self.params = None

    def bind(self, params):
        self.params = params
        return self

    def unbound_impl(self, precision):
        ctor = getattr(custom_ops, type(self).__name__ + _suffix(precision))
        return ctor(*self.args)

    def bound_impl(self, precision):
        if self.params is None:
            raise ValueError("This op has not been bound to parameters.")
        return custom_ops.BoundPotential(self.unbound_impl(precision), self.params)


class SummedPotential(CustomOpWrapper):
    def __init__(self, potentials: List[CustomOpWrapper], params_init: List[np.ndarray]):
        if len(potentials) != len(params_init):
            raise ValueError("number of potentials != number of parameter arrays")
        self._potentials = potentials
        self._sizes = [np.asarray(p).size for p in params_init]
        self.params = None

    def unbound_impl(self, precision):
        return custom_ops.SummedPotential([p.unbound_impl(precision) for p in self._potentials], self._sizes)


class FanoutSummedPotential(CustomOpWrapper):
    def __init__(self, potentials: List[CustomOpWrapper]):
        self._potentials = potentials
        self.params = None

    def unbound_impl(self, precision):
        return custom_ops.FanoutSummedPotential([p.unbound_impl(precision) for p in self._potentials])


class BondedWrapper(CustomOpWrapper):
    def get_idxs(self):
        return self.args[0]

    def set_idxs(self, new_idxs):
        self.args[0] = new_idxs

    def get_lambda_mult(self):
        return self.args[1] if len(self.args) > 1 else None

    def get_lambda_offset(self):
        return self.args[2] if len(self.args) > 2 else None

    def set_lambda_mult_and_offset(self, mult, offset):
        if len(self.args) > 1:
            self.args[1] = mult
            self.args[2] = offset
        else:
            self.args += [mult, offset]


class HarmonicBond(BondedWrapper):
    pass


class HarmonicAngle(BondedWrapper):
    pass


class PeriodicTorsion(BondedWrapper):
    pass


class NonbondedImplWrapper(custom_ops.FanoutSummedPotential):
    """FanoutSummedPotential that forwards the neighborlist knobs to its parts."""

    def disable_hilbert_sort(self):
        for impl in self.get_potentials():
            if hasattr(impl, "disable_hilbert_sort"):
                impl.disable_hilbert_sort()

    def set_nblist_padding(self, padding):
        for impl in self.get_potentials():
            if hasattr(impl, "set_nblist_padding"):
                impl.set_nblist_padding(padding)


class NonbondedCustomOpWrapper(CustomOpWrapper):
    def unbound_impl(self, precision):
        ctor = getattr(custom_ops, type(self).__name__ + _suffix(precision))
        return ctor(_KERNEL_DIR, *self.args)


class Nonbonded(NonbondedCustomOpWrapper):
    """Nonbonded(exclusion_idxs, scale_factors, lambda_plane_idxs, lambda_offset_idxs, beta, cutoff)
    = NonbondedAllPairs + NonbondedPairListNegated over the exclusions (lib/potentials.py:269-301)."""

    def __init__(self, *args):
        exclusion_idxs = np.asarray(args[0])
        uniq = {tuple(sorted((int(a), int(b)))) for a, b in exclusion_idxs.reshape(-1, 2)}
        assert len(uniq) == exclusion_idxs.reshape(-1, 2).shape[0]
        super().__init__(*args)

    def unbound_impl(self, precision):
        all_pairs = NonbondedAllPairs(
            self.get_lambda_plane_idxs(), self.get_lambda_offset_idxs(), self.get_beta(), self.get_cutoff()
        ).unbound_impl(precision)
        exclusions = NonbondedPairListNegated(
            self.get_exclusion_idxs(),
            self.get_scale_factors(),
            self.get_lambda_plane_idxs(),
            self.get_lambda_offset_idxs(),
            self.get_beta(),
            self.get_cutoff(),
        ).unbound_impl(precision)
        return NonbondedImplWrapper([all_pairs, exclusions])

    def get_exclusion_idxs(self):
        return self.args[0]

    def set_exclusion_idxs(self, x):
        self.args[0] = x

    def get_scale_factors(self):
        return self.args[1]

    def set_scale_factors(self, x):
        self.args[1] = x

    def get_lambda_plane_idxs(self):
        return self.args[2]

    def get_lambda_offset_idxs(self):
        return self.args[3]

    def set_lambda_plane_idxs(self, val):
        self.args[2] = val

    def set_lambda_offset_idxs(self, val):
        self.args[3] = val

    def get_beta(self):
        return self.args[4]

    def get_cutoff(self):
        return self.args[5]

    def interpolate(self):
        """The interpolated variant: parameters [2N, 3] = (initial, final), used at
        lambda as (1 - lambda) initial + lambda final (lib/potentials.py:333-345)."""
        return NonbondedInterpolated(*self.args)


class NonbondedInterpolated(Nonbonded):
    """Nonbonded over [2N, 3] interpolated parameters (lib/potentials.py:347-368)."""

    def unbound_impl(self, precision):
        all_pairs = NonbondedAllPairsInterpolated(
            self.get_lambda_plane_idxs(), self.get_lambda_offset_idxs(), self.get_beta(), self.get_cutoff(), None,
            *self.args[6:]
        ).unbound_impl(precision)
        exclusions = NonbondedPairListNegatedInterpolated(
            self.get_exclusion_idxs(),
            self.get_scale_factors(),
            self.get_lambda_plane_idxs(),
            self.get_lambda_offset_idxs(),
            self.get_beta(),
            self.get_cutoff(),
            *self.args[6:],
        ).unbound_impl(precision)
        return NonbondedImplWrapper([all_pairs, exclusions])


class NonbondedAllPairs(NonbondedCustomOpWrapper):
    pass


class NonbondedAllPairsInterpolated(NonbondedAllPairs):
    def unbound_impl(self, precision):
        ctor = getattr(custom_ops, "NonbondedAllPairs" + _suffix(precision) + "_interpolated")
        return ctor(_KERNEL_DIR, *self.args)


class NonbondedInteractionGroup(NonbondedCustomOpWrapper):
    """NonbondedInteractionGroup(row_atom_idxs, lambda_plane_idxs, lambda_offset_idxs, beta, cutoff)
    (lib/potentials.py:389-390)."""


class NonbondedInteractionGroupInterpolated(NonbondedInteractionGroup):
    def unbound_impl(self, precision):
        ctor = getattr(custom_ops, "NonbondedInteractionGroup" + _suffix(precision) + "_interpolated")
        return ctor(_KERNEL_DIR, *self.args)


class NonbondedPairList(NonbondedCustomOpWrapper):
    def get_idxs(self):
        return self.args[0]

    def get_rescale_mask(self):
        return self.args[1]

    def get_beta(self):
        return self.args[4]

    def get_cutoff(self):
        return self.args[5]


class NonbondedPairListNegated(NonbondedCustomOpWrapper):
    def unbound_impl(self, precision):
        ctor = getattr(custom_ops, "NonbondedPairList" + _suffix(precision) + "_negated")
        return ctor(_KERNEL_DIR, *self.args)


class NonbondedPairListInterpolated(NonbondedPairList):
    def unbound_impl(self, precision):
        ctor = getattr(custom_ops, "NonbondedPairList" + _suffix(precision) + "_interpolated")
        return ctor(_KERNEL_DIR, *self.args)


class NonbondedPairListNegatedInterpolated(NonbondedCustomOpWrapper):
    def unbound_impl(self, precision):
        ctor = getattr(custom_ops, "NonbondedPairList" + _suffix(precision) + "_negated_interpolated")
        return ctor(_KERNEL_DIR, *self.args)
