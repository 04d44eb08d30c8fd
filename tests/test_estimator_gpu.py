"""The reference's free-energy protocol on the device (GPU): every window of a
schedule FIRE-minimised, equilibrated and sampled in one batch, the
[K, F, K] energies from multiple_steps_U, then BAR and MBAR
(fe/estimator.py:200-423). The estimators themselves are pinned on the CPU
(tests/test_estimator.py); here the pipeline runs end to end, is bitwise
reproducible, and BAR and MBAR agree on the same data."""

import numpy as np
import pytest

from timemachine_amd import testsystems as ts
from timemachine_amd.fe import estimator as est

pytestmark = pytest.mark.gpu


def test_windows_to_delta_g(gpu):
    """Ligand charges scaled to 0.9 by parameter interpolation (no 4D offsets):
    neighbouring windows overlap well, so BAR and MBAR must agree."""
    import dataclasses

    s = ts.ligand_in_water(side=3.2, n_waters=1050, n_heavy=24, n_ligand=40, seed=5)
    p1 = s.nb_params.copy()
    p1[s.ligand_idxs, 0] *= 0.9
    s = dataclasses.replace(s, nb_params_final=p1, lambda_offset_idxs=np.zeros_like(s.lambda_offset_idxs))
    lams = np.linspace(0.0, 1.0, 6)
    runs = []
    for _ in range(2):
        U, _, _ = est.simulate_windows(s, lams, equil_steps=200, prod_steps=1000, u_interval=20, seed=3)
        runs.append(U)
    assert runs[0].shape == (6, 50, 6)
    np.testing.assert_array_equal(runs[0], runs[1])  # fixed point + counter-based noise
    kT = 0.0083144626 * 300.0
    r = est.deltaG_from_U_knk(runs[0], 1.0 / kT, lams)
    print(f"BAR {r['bar_dG']:.3f} +- {r['bar_dG_err']:.3f}  MBAR {r['mbar_dG']:.3f} +- {r['mbar_dG_err']:.3f} "
          "kJ/mol", flush=True)
    assert np.isfinite(r["bar_dG"]) and np.isfinite(r["mbar_dG"])
    assert r["bar_dG_err"] > 0 and r["mbar_dG_err"] > 0
    assert abs(r["bar_dG"] - r["mbar_dG"]) < 3 * (r["bar_dG_err"] + r["mbar_dG_err"])
