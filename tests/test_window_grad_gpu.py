"""Gradients of lambda-window averages of du/dlambda (GPU, through the C ABI).

* The in-run observer equals observing by hand at the same states, bit for bit.
* The observer's sums equal the FP64 oracle evaluated on the same frames
  (du/dlambda, du/dp and the double-backward d(du/dlambda)/dp), within the
  du_dp tolerance (1e-5 of the column scale; bonded blocks of sum g and
  sum h 1e-8 of the column scale plus one fixed-point quantum 2^-36 per
  frame; sums of A g take A's 1e-5).
* The thermodynamic gradient equals central finite differences (h = 1e-4) of
  the reweighted window average over the same frames.
* The autograd loss gives the same value and gradients with and without
  recomputation (bitwise), and equals the chain rule over the window
  gradients.
* The differentiable U of fe.functional returns the custom op's gradients.
"""

import numpy as np
import pytest
import torch

from oracle import bonded as obonded
from oracle import nonbonded as onb
from timemachine_amd import testsystems as ts
from timemachine_amd.fe import functional
from timemachine_amd.fe import window_gradients as wg

pytestmark = pytest.mark.gpu

LAMS = [0.25, 0.5, 0.75]
K = 5  # steps between frames
F = 4  # frames


@pytest.fixture(scope="module")
def small(gpu):
    return ts.ligand_in_water(side=2.9, n_waters=780, n_heavy=12, n_ligand=22, seed=7)


@pytest.fixture(scope="module")
def windows(small):
    return wg.LambdaWindows(small, LAMS, n_frames=F, frame_interval=K, equil_steps=0, seed=11)


@pytest.fixture(scope="module")
def by_hand(windows):
    """Observe by hand at the states the in-run observer uses (after K-1, 2K-1, ... steps)."""
    batch, keep = windows.batch()
    batch.set_grad_observer(0)
    frames = [[] for _ in LAMS]
    for _ in range(F):
        if K > 1:
            batch.multiple_steps(K - 1)
        for w in range(len(LAMS)):
            frames[w].append((batch.get_x_t(w), batch.get_box(w)))
        batch.observe()
        batch.multiple_steps(1)
    stats = [batch.grad_stats(w) for w in range(len(LAMS))]
    del batch, keep
    return frames, stats


def test_in_run_observer_equals_observing_by_hand(windows, by_hand):
    means, series, stats = windows.sample(observe=True)
    _, hand = by_hand
    for w in range(len(LAMS)):
        assert stats[w]["n"] == F == hand[w]["n"]
        for key in ("sum_a", "sum_aa"):
            assert stats[w][key] == hand[w][key], key
        for key in ("sum_g", "sum_ag", "sum_h"):
            assert np.array_equal(stats[w][key], hand[w][key]), key
        # the du_dl series stored with the same interval sees the same states
        assert means[w] == pytest.approx(stats[w]["sum_a"] / F, rel=1e-12, abs=1e-12)


def _oracle_frame(sys_, x, box, lam):
    nb = onb.nonbonded_pairs(x, sys_.nb_params, box, lam, sys_.exclusion_idxs, sys_.scales, sys_.beta,
                             sys_.cutoff, sys_.lambda_plane_idxs, sys_.lambda_offset_idxs)
    pairs = onb.candidate_pairs(x, box, sys_.cutoff)
    alch = sys_.lambda_offset_idxs != 0
    pairs = pairs[alch[pairs[:, 0]] | alch[pairs[:, 1]]]
    h_nb = onb.nonbonded_pairs_dl_dp(x, sys_.nb_params, box, lam, sys_.exclusion_idxs, sys_.scales, sys_.beta,
                                     sys_.cutoff, sys_.lambda_plane_idxs, sys_.lambda_offset_idxs, pairs=pairs)
    a = nb[2]
    gs = [nb[1].reshape(-1)]
    hs = [h_nb.reshape(-1)]
    for kind, idxs, params in (("bond", sys_.bond_idxs, sys_.bond_params), ("angle", sys_.angle_idxs, sys_.angle_params),
                               ("torsion", sys_.torsion_idxs, sys_.torsion_params)):
        _, gp, gl, _ = obonded.bonded(kind, x, params, lam, idxs)
        a += gl
        gs.append(gp.reshape(-1))
        hs.append(obonded.bonded_dl_dp(kind, x, params, lam, idxs).reshape(-1))
    return a, np.concatenate(gs), np.concatenate(hs)


def test_observer_sums_match_oracle(small, windows, by_hand):
    frames, stats = by_hand
    for w, lam in enumerate(LAMS):
        sa, sg, sag, sh = 0.0, 0.0, 0.0, 0.0
        for x, box in frames[w]:
            a, g, h = _oracle_frame(small, x, box, lam)
            sa, sg, sag, sh = sa + a, sg + g, sag + a * g, sh + h
        st = stats[w]
        print(f"[lam={lam}] sum A {st['sum_a']:.6f} oracle {sa:.6f}")
        assert abs(st["sum_a"] - sa) <= 1e-5 * max(abs(sa), F)
        for key, ref in (("sum_g", sg), ("sum_ag", sag), ("sum_h", sh)):
            got = windows.layout.split(st[key])
            exp = windows.layout.split(ref)
            for block in ("nonbonded", "bond", "angle", "torsion"):
                # A = du/dlambda carries the nonbonded tolerance into every A g sum
                tol = 1e-5 if (block == "nonbonded" or key == "sum_ag") else 1e-8
                # bonded parameters belong to one term: one fixed-point rounding per frame
                # (times |A| for the A g sums)
                a_scale = abs(st["sum_a"]) / F + 1.0
                quantum = F * 2.0**-36 * (a_scale if key == "sum_ag" else 1.0)
                for c in range(exp[block].shape[1] if exp[block].size else 0):
                    scale = np.abs(exp[block][:, c]).max()
                    if scale == 0.0:
                        assert np.all(got[block][:, c] == 0.0), (key, block, c)
                        continue
                    err = np.abs(got[block][:, c] - exp[block][:, c]).max()
                    assert err <= tol * scale + (0.0 if block == "nonbonded" else quantum), (
                        lam, key, block, c, err, scale)


def _reweighted_mean(small, frames_w, lam, nb_params, torsion_params, u_ref, kT):
    """sum_f A_f(p) w_f / sum_f w_f, w_f = exp(-(U_f(p) - U_f(p0)) / kT), on the GPU potentials."""
    s2 = ts.System(**{**small.__dict__, "nb_params": nb_params, "torsion_params": torsion_params})
    nb_impl = s2.nonbonded().unbound_impl(np.float32)
    from timemachine_amd.lib import potentials

    tor = potentials.PeriodicTorsion(s2.torsion_idxs).unbound_impl(np.float32)
    a_f, u_f = [], []
    for x, box in frames_w:
        _, _, dl, u = nb_impl.execute_selective(x, nb_params, box, lam, False, False, True, True)
        _, _, _, ut = tor.execute_selective(x, torsion_params, box, lam, False, False, False, True)
        a_f.append(dl)
        u_f.append(u + ut)
    u_f = np.asarray(u_f)
    if u_ref is None:
        return None, u_f
    logw = -(u_f - u_ref) / kT
    wts = np.exp(logw - logw.max())
    return float(np.sum(np.asarray(a_f) * wts) / np.sum(wts)), u_f


def test_thermodynamic_gradient_matches_finite_differences(small, windows, by_hand):
    frames, stats = by_hand
    h = 1e-4
    lig = small.ligand_idxs
    picks_nb = [(int(lig[0]), 0), (int(lig[3]), 1), (int(lig[5]), 2), (int(lig[-1]), 0)]
    picks_t = [0, len(small.torsion_idxs) // 2]
    for w, lam in enumerate(LAMS):
        G = windows.layout.split(wg.window_gradient(stats[w], windows.kT))
        _, u0 = _reweighted_mean(small, frames[w], lam, small.nb_params, small.torsion_params, None, windows.kT)
        for i, c in picks_nb:
            pp, pm = small.nb_params.copy(), small.nb_params.copy()
            pp[i, c] += h
            pm[i, c] -= h
            fp, _ = _reweighted_mean(small, frames[w], lam, pp, small.torsion_params, u0, windows.kT)
            fm, _ = _reweighted_mean(small, frames[w], lam, pm, small.torsion_params, u0, windows.kT)
            fd = (fp - fm) / (2 * h)
            print(f"[lam={lam}] nb ({i},{c}) analytic {G['nonbonded'][i, c]:.6e} fd {fd:.6e}")
            assert G["nonbonded"][i, c] == pytest.approx(fd, rel=2e-4, abs=2e-4 * max(1.0, abs(fd)))
        for t in picks_t:
            tp, tm = small.torsion_params.copy(), small.torsion_params.copy()
            tp[t, 0] += h
            tm[t, 0] -= h
            fp, _ = _reweighted_mean(small, frames[w], lam, small.nb_params, tp, u0, windows.kT)
            fm, _ = _reweighted_mean(small, frames[w], lam, small.nb_params, tm, u0, windows.kT)
            fd = (fp - fm) / (2 * h)
            print(f"[lam={lam}] torsion k {t} analytic {G['torsion'][t, 0]:.6e} fd {fd:.6e}")
            assert G["torsion"][t, 0] == pytest.approx(fd, rel=2e-4, abs=1e-6)


def test_autograd_loss_recompute_and_chain_rule(small, windows):
    target = 3.0
    out = []
    for recompute in (True, False):
        nb = torch.tensor(small.nb_params, dtype=torch.float64, requires_grad=True)
        tk = torch.tensor(small.torsion_params[:, 0], dtype=torch.float64, requires_grad=True)
        loss = wg.ti_loss(nb, tk, windows, target, recompute=recompute)
        loss.backward()
        out.append((float(loss.detach()), nb.grad.numpy().copy(), tk.grad.numpy().copy()))
    (l1, g1, t1), (l2, g2, t2) = out
    assert l1 == l2
    assert np.array_equal(g1, g2) and np.array_equal(t1, t2)
    # chain rule by hand
    means, _, stats = windows.sample(observe=True)
    c = wg.trapezoid_weights(LAMS)
    dl = 2 * (np.sum(c * means) - target)
    tot = sum(c[w] * wg.window_gradient(stats[w], windows.kT) for w in range(len(LAMS)))
    parts = windows.layout.split(dl * tot)
    np.testing.assert_allclose(g1, parts["nonbonded"], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(t1, parts["torsion"][:, 0], rtol=1e-12, atol=1e-14)
    # only ligand rows (and their near neighbours through the explicit term) carry gradient
    assert np.abs(g1[small.ligand_idxs]).sum() > 0


def test_differentiable_potential_energy(small):
    """torch gradients of U from fe.functional equal the custom op's outputs."""
    nbw = small.nonbonded()
    U = functional.construct_differentiable_interface([nbw])
    x = torch.tensor(small.x, dtype=torch.float64, requires_grad=True)
    p = torch.tensor(small.nb_params, dtype=torch.float64, requires_grad=True)
    lam = torch.tensor(0.4, dtype=torch.float64, requires_grad=True)
    u = U(x, [p], small.box, lam)
    u.backward()
    ref = nbw.unbound_impl(np.float32).execute(small.x, small.nb_params, small.box, 0.4)
    np.testing.assert_array_equal(x.grad.numpy(), ref[0])
    np.testing.assert_array_equal(p.grad.numpy(), ref[1])
    assert float(lam.grad) == ref[2]
    assert float(u.detach()) == ref[3]
