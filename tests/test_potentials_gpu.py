"""Potential composition and ownership on the GPU (restates
tests/test_potentials.py:16-93 and :208-250 of the reference).

* BoundPotential / SummedPotential keep their component potentials alive on
  the C side (refcounted handles), also after every Python reference to the
  components is gone;
* SummedPotential validates its lengths and its parameter size with the
  reference's messages (summed_potential.cu:19-68);
* FanoutSummedPotential equals SummedPotential on shared parameters;
* construct_differentiable_interface_fast (one SummedPotential custom op)
  equals the sum of the per-potential executions, bit for bit in fixed point.
"""

import gc

import numpy as np
import pytest
import torch

from timemachine_amd import _lib
from timemachine_amd import testsystems as ts
from timemachine_amd.fe import functional
from timemachine_amd.lib import custom_ops
from timemachine_amd.lib.potentials import FanoutSummedPotential, HarmonicBond, SummedPotential

pytestmark = pytest.mark.gpu
F32 = np.float32


@pytest.fixture
def harmonic_bond(gpu):
    bond_idxs = np.array([[0, 1], [0, 2]], dtype=np.int32)
    params = np.ones(shape=(2, 2), dtype=np.float32)
    return HarmonicBond(bond_idxs).bind(params)


def execute_bound_impl(bp):
    coords = np.zeros(shape=(3, 3), dtype=np.float32)
    box = np.diag(np.ones(3))
    return bp.execute(coords, box, 1)


def test_bound_potential_keeps_referenced_potential_alive(harmonic_bond):
    bp = harmonic_bond.bound_impl(F32)
    execute_bound_impl(bp)


def test_bound_potential_get_potential(harmonic_bond):
    unbound_impl = harmonic_bond.unbound_impl(F32)
    bound_impl = custom_ops.BoundPotential(unbound_impl, harmonic_bond.params)
    assert unbound_impl is bound_impl.get_potential()


def test_bound_potential_outlives_python_potential(harmonic_bond):
    """The C handle of the bound potential holds its potential: with the Python
    potential object gone (its handle released), execution is unchanged."""
    unbound = harmonic_bond.unbound_impl(F32)
    bp = custom_ops.BoundPotential(unbound, harmonic_bond.params)
    ref = execute_bound_impl(bp)
    bp._potential = None
    del unbound
    gc.collect()
    out = execute_bound_impl(bp)
    for a, b in zip(out, ref):
        np.testing.assert_array_equal(a, b)


def test_summed_potential_raises_on_inconsistent_lengths(harmonic_bond):
    with pytest.raises(ValueError) as excinfo:
        SummedPotential([harmonic_bond], [])
    assert str(excinfo.value) == "number of potentials != number of parameter arrays"


def test_summed_potential_keeps_referenced_potentials_alive(harmonic_bond):
    sp = SummedPotential([harmonic_bond], [harmonic_bond.params]).bind(harmonic_bond.params)
    execute_bound_impl(sp.bound_impl(F32))


def test_summed_potential_outlives_python_components(harmonic_bond):
    """Releasing every Python handle of the components leaves the summed
    potential working: the C side holds them (shared ownership)."""
    impls = [harmonic_bond.unbound_impl(F32) for _ in range(2)]
    summed = custom_ops.SummedPotential(impls, [harmonic_bond.params.size] * 2)
    params = np.concatenate([harmonic_bond.params.reshape(-1)] * 2)
    x = np.random.default_rng(0).random((3, 3))
    ref = summed.execute(x, params, np.eye(3) * 3.0, 0.5)
    summed._pots = []
    del impls
    gc.collect()
    out = summed.execute(x, params, np.eye(3) * 3.0, 0.5)
    for a, b in zip(out, ref):
        np.testing.assert_array_equal(a, b)


def test_summed_potential_get_potentials(harmonic_bond):
    impls = [harmonic_bond.unbound_impl(F32) for _ in range(2)]
    params_sizes = [harmonic_bond.params.size] * 2
    summed_impl = custom_ops.SummedPotential(impls, params_sizes)
    assert set(id(p) for p in summed_impl.get_potentials()) == set(id(p) for p in impls)


def test_summed_potential_invalid_parameters_size(harmonic_bond):
    sp = SummedPotential([harmonic_bond], [harmonic_bond.params])
    n = harmonic_bond.params.size
    with pytest.raises(RuntimeError) as e:
        execute_bound_impl(sp.bind(np.empty(0)).bound_impl(F32))
    assert f"SummedPotential::execute_device(): expected {n} parameters, got 0" in str(e.value)
    with pytest.raises(RuntimeError) as e:
        execute_bound_impl(sp.bind(np.ones(n + 1)).bound_impl(F32))
    assert f"SummedPotential::execute_device(): expected {n} parameters, got {n + 1}" in str(e.value)


def test_fanout_summed_potential_consistency(gpu):
    rng = np.random.default_rng(2022)
    num_atoms, num_bonds = 10, 10
    box = 3.0 * np.eye(3)
    coords = rng.uniform(0, 1, size=(num_atoms, 3)).astype(np.float32).astype(np.float64)

    def random_bond_idxs():
        return np.array([rng.choice(num_atoms, size=(2,), replace=False) for _ in range(num_bonds)], dtype=np.int32)

    hb1, hb2 = HarmonicBond(random_bond_idxs()), HarmonicBond(random_bond_idxs())
    params = rng.uniform(0, 1, size=(num_bonds, 2))
    summed = SummedPotential([hb1, hb2], [params, params])
    fanout = FanoutSummedPotential([hb1, hb2])
    du_dx_ref, du_dps_ref, du_dl_ref, u_ref = summed.unbound_impl(F32).execute(
        coords, np.concatenate([params.reshape(-1)] * 2), box, 0.1)
    du_dx, du_dp, du_dl, u = fanout.unbound_impl(F32).execute(coords, params, box, 0.1)
    np.testing.assert_array_equal(du_dx_ref, du_dx)
    np.testing.assert_allclose(np.asarray(du_dps_ref).reshape(2, -1).sum(0), np.asarray(du_dp).reshape(-1),
                               rtol=1e-8, atol=1e-8)
    assert du_dl_ref == du_dl and u_ref == u


def test_differentiable_interface_fast_equals_sum_of_potentials(gpu):
    """construct_differentiable_interface_fast(ubps, params): energy, coordinate and
    parameter gradients equal the sum over the per-potential executions; the
    energy and forces bit for bit in fixed point (the summed op and the separate
    ops add the same integers)."""
    sys_ = ts.ligand_in_water(side=3.0, n_waters=800, n_heavy=12, n_ligand=22, seed=3)
    bps = sys_.bound_potentials()
    params = [np.asarray(bp.params, dtype=np.float64) for bp in bps]
    U = functional.construct_differentiable_interface_fast(bps, params)
    lam = 0.4
    x = torch.tensor(sys_.x, requires_grad=True)
    ps = [torch.tensor(p, requires_grad=True) for p in params]
    u = U(x, ps, sys_.box, lam)
    u.backward()
    total_u = np.uint64(0)
    total_dx = np.zeros(sys_.N * 3, dtype=np.uint64)
    for bp, p, tp in zip(bps, params, ps):
        du_dx, du_dp, du_dl, u_raw = bp.unbound_impl(F32).execute_fixed_raw(sys_.x, p, sys_.box, lam)
        total_dx += du_dx  # modulo 2^64
        with np.errstate(over="ignore"):
            total_u = total_u + _lib.wrapping_sum(u_raw)
        ref_dp = bp.unbound_impl(F32).execute(sys_.x, p, sys_.box, lam)[1]
        np.testing.assert_allclose(tp.grad.numpy(), np.asarray(ref_dp).reshape(p.shape), rtol=0, atol=0)
    assert float(u.detach()) == _lib.fixed_to_float(total_u)
    np.testing.assert_array_equal(x.grad.numpy().reshape(-1), _lib.fixed_to_float(total_dx))
