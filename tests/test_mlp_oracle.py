"""Pins the CPU restatement of the learned reconstruction (oracle/mlp.py): backward pass
against finite differences and against torch autograd fixtures, bf16 rounding against
torch's conversion, the bounded interface states, feature invariances and the van Leer
member of the family."""
import os

import numpy as np
import pytest

from oracle import mlp
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mlp_autograd.npz")


def _params(seed):
    rng = np.random.default_rng(seed)
    return {"W1": rng.normal(0, 0.3, (48, 64)), "b1": rng.normal(0, 0.1, 64),
            "W2": rng.normal(0, 0.2, (64, 64)), "b2": rng.normal(0, 0.1, 64),
            "W3": rng.normal(0, 0.3, (64, 16)), "b3": rng.normal(0, 0.1, 16)}


def test_backward_matches_finite_differences():
    P = _params(1)
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, (6, 48))
    tau = rng.uniform(-1, 1, (6, 8))
    wp = rng.uniform(0, 1, (6, 16))
    om = rng.uniform(0.2, 2.0, (6, 8))
    _, g = mlp.loss_and_grad(P, x, tau, wp, om)
    h = 1e-6
    for k in P:
        flat = P[k].reshape(-1)
        for idx in rng.choice(flat.size, size=min(12, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + h
            lp, _ = mlp.loss_and_grad(P, x, tau, wp, om)
            flat[idx] = old - h
            lm, _ = mlp.loss_and_grad(P, x, tau, wp, om)
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            assert g[k].reshape(-1)[idx] == pytest.approx(fd, rel=1e-6, abs=1e-9), (k, idx)


def test_backward_matches_torch_autograd_fixtures():
    F = np.load(GOLDEN)
    for seed in range(2):
        P = {k: F["s%d_%s" % (seed, k)] for k in ("W1", "b1", "W2", "b2", "W3", "b3")}
        loss, g = mlp.loss_and_grad(P, F["s%d_x" % seed], F["s%d_tau" % seed],
                                    F["s%d_wp" % seed], F["s%d_om" % seed])
        assert loss == pytest.approx(float(F["s%d_loss" % seed]), rel=1e-13)
        y, _ = mlp.forward(P, F["s%d_x" % seed])
        assert np.allclose(y, F["s%d_y" % seed], rtol=1e-13, atol=0)
        for k in P:
            ref = F["s%d_g%s" % (seed, k)]
            assert np.allclose(g[k], ref, rtol=1e-10, atol=1e-14 * np.max(np.abs(ref))), k


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(0, 1, 4000).astype(np.float32),
                        np.float32([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 0.0, 65504.0,
                                    1e-30, 3.0e38])])
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(mlp.bf16(x), ref)


def test_checked_weights_are_tvd():
    """Checked weights give slopes inside the TVD region: per cell, the deviation of each
    face state from the cell value has the sign of, and is at most, each one-sided
    difference."""
    rng = np.random.default_rng(7)
    S = rng.normal(0, 1, (400, 8, 6))
    w = mlp.weights(rng.uniform(-0.5, 0.5, (400, 16)), rng.uniform(0.2, 0.8, (400, 16)))
    w[0, 3] = np.nan
    wc = mlp.checked_weights(S, w, fallback="plm")
    vL, vR = mlp.face_states(S, wc)
    a, b, c, d = S[..., 1], S[..., 2], S[..., 3], S[..., 4]
    dl = vL - b   # deviation of cell i-1's right state
    dr = c - vR   # deviation of cell i's left state
    tol = 1e-12
    assert np.all((np.abs(dl) <= tol) | ((dl * (b - a) > 0) & (np.abs(dl) <= np.abs(b - a) + tol)
                                         & (dl * (c - b) > 0) & (np.abs(dl) <= np.abs(c - b) + tol)))
    assert np.all((np.abs(dr) <= tol) | ((dr * (d - c) > 0) & (np.abs(dr) <= np.abs(d - c) + tol)
                                         & (dr * (c - b) > 0) & (np.abs(dr) <= np.abs(c - b) + tol)))
    # weights kept where they pass, van Leer where they fail
    ok = np.isclose(wc, w)
    assert 0.0 < np.mean(ok) < 1.0
    vl = mlp.van_leer_weights(S)
    assert np.allclose(wc[~ok], vl[~ok])
    # the van Leer member passes its own check
    assert np.allclose(mlp.checked_weights(S, vl, fallback="plm"), vl)


def test_ppm_member_reproduces_ppm_and_is_the_default_fallback():
    from oracle import oracle as O
    rng = np.random.default_rng(12)
    S = rng.normal(0, 1, (60, 8, 6))
    S[:, 0] = np.abs(S[:, 0]) + 0.1   # density positive
    S[:, 4] = np.abs(S[:, 4]) + 0.1   # pressure positive
    w = mlp.ppm_weights(S)
    vL, vR = mlp.face_states(S, w)
    for f in range(0, 60, 7):
        for q in range(8):
            _, left = O.ppm_states(*S[f, q, 0:5])
            right, _ = O.ppm_states(*S[f, q, 1:6])
            if q in (0, 4):
                if left <= 0:
                    left = S[f, q, 2] + 0.5 * O.vanleer(S[f, q, 3] - S[f, q, 2], S[f, q, 2] - S[f, q, 1])
                if right <= 0:
                    right = S[f, q, 3] - 0.5 * O.vanleer(S[f, q, 4] - S[f, q, 3], S[f, q, 3] - S[f, q, 2])
            assert vL[f, q] == pytest.approx(left, rel=1e-13, abs=1e-13)
            assert vR[f, q] == pytest.approx(right, rel=1e-13, abs=1e-13)
    # weights that fail the check are replaced by the PPM ones
    bad = np.full((60, 16), 2.0)
    assert np.allclose(mlp.checked_weights(S, bad), w)


def test_features_shift_and_scale_invariant():
    rng = np.random.default_rng(9)
    S = rng.normal(0, 1, (50, 8, 6))
    x0, _, _ = mlp.features(S)
    x1, _, _ = mlp.features(3.5 * S + 2.0)
    assert np.allclose(x0, x1, rtol=1e-12, atol=1e-12)
    assert np.max(np.abs(x0)) <= 1.0 + 1e-15
    # a uniform variable gives zero features
    S[:, 0, :] = 4.0
    x2, _, s = mlp.features(S)
    assert np.all(x2[:, :6] == 0.0) and np.all(s[:, 0] == 0.0)


def test_van_leer_member_reproduces_plm():
    rng = np.random.default_rng(11)
    S = rng.normal(0, 1, (300, 8, 6))
    w = mlp.van_leer_weights(S)
    vL, vR = mlp.face_states(S, w)
    # PLM van Leer states of the interface between cells 2 and 3 (oracle's limiter)
    vl = np.vectorize(O.vanleer)
    slope_l = vl(S[..., 3] - S[..., 2], S[..., 2] - S[..., 1])
    slope_r = vl(S[..., 4] - S[..., 3], S[..., 3] - S[..., 2])
    assert np.allclose(vL, S[..., 2] + 0.5 * slope_l, rtol=0, atol=1e-13)
    assert np.allclose(vR, S[..., 3] - 0.5 * slope_r, rtol=0, atol=1e-13)


def test_zero_output_layer_is_ppm():
    """With a zero output layer the corrected weights are exactly PPM's."""
    P = _params(3)
    P["W3"][:] = 0.0
    P["b3"][:] = 0.0
    rng = np.random.default_rng(13)
    S = rng.normal(0, 1, (80, 8, 6))
    S[:, 0] = np.abs(S[:, 0]) + 0.1
    S[:, 4] = np.abs(S[:, 4]) + 0.1
    x, _, _ = mlp.features(S)
    y, _ = mlp.forward(P, x)
    wp = mlp.ppm_weights(S)
    assert np.array_equal(mlp.checked_weights(S, mlp.weights(y, wp)), wp)


def test_bf16_forward_close_to_fp64():
    P = _params(2)
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, (64, 48))
    y64, _ = mlp.forward(P, x)
    yb, _ = mlp.forward(P, x, bf16_mode=True)
    assert np.max(np.abs(yb - y64)) < 5e-2 * max(1.0, np.max(np.abs(y64)))
