"""GPU checks of the learned limiter through the C ABI: FP32 training gradients and Adam
steps against oracle/mlp.py, BF16 MFMA inference against the oracle's BF16-rounded
forward, the PPM fallback, the MLP-mode solver against PPM (fallback weights), and
the flux-form invariants (conservation, div B) with a loaded network."""
import numpy as np
import pytest

from idefix_amd import learned, problems
from oracle import mlp

pytestmark = pytest.mark.gpu


def _batch(B, seed):
    """Training rows [B, 72] (features | PPM weights | loss weights) and targets [B, 8]."""
    rng = np.random.default_rng(seed)
    S = rng.normal(0, 1, (B, 8, 6))
    S[:, 0] = np.abs(S[:, 0]) + 0.5
    x, c, s = mlp.features(S)
    t = 0.5 * (S[:, :, 2] + S[:, :, 3]) + 0.3 * rng.normal(0, 1, (B, 8)) * s
    tau = np.where(s > 0, (t - c) / np.where(s > 0, s, 1.0), 0.0)
    wp = rng.uniform(0, 1, (B, 16))
    om = rng.uniform(0.3, 2.0, (B, 8))
    rows = np.concatenate([x, wp, om], axis=1).astype(np.float32).astype(np.float64)
    return rows, tau


def _split(rows):
    return rows[:, :48], rows[:, 48:64], rows[:, 64:72]


def _params(seed, scale_out=1.0):
    p = learned.init_params(seed)
    rng = np.random.default_rng(seed + 100)
    p = p + rng.normal(0, 0.05, p.shape).astype(np.float32)  # non-zero biases
    P = learned.unpack(p)
    P["W3"] = P["W3"] * scale_out
    P["b3"] = P["b3"] * scale_out
    return learned.pack(P)


def test_training_gradient_matches_oracle():
    import torch
    rows, tau = _batch(3000, 1)
    p = _params(3)
    tr = learned.Trainer(4096, p)
    X = torch.tensor(rows, dtype=torch.float32, device="cuda")
    T = torch.tensor(tau, dtype=torch.float32, device="cuda")
    loss, g = tr.loss_grad(X, T)
    P = learned.unpack(p)
    lo, go = mlp.loss_and_grad(P, *_split(rows)[:1], tau.astype(np.float32).astype(np.float64),
                               *_split(rows)[1:])
    assert loss == pytest.approx(lo, rel=1e-5)
    gref = learned.pack(go).astype(np.float64)
    err = np.abs(g - gref)
    assert np.max(err) <= 1e-4 * np.max(np.abs(gref)), np.max(err) / np.max(np.abs(gref))


def test_adam_steps_match_oracle():
    import torch
    rows, tau = _batch(1024, 2)
    p = _params(4)
    tr = learned.Trainer(1024, p)
    X = torch.tensor(rows, dtype=torch.float32, device="cuda")
    T = torch.tensor(tau, dtype=torch.float32, device="cuda")
    P = learned.unpack(p)
    m = {k: np.zeros_like(v) for k, v in P.items()}
    v = {k: np.zeros_like(v) for k, v in P.items()}
    xs, wp, om = _split(rows)
    ts = tau.astype(np.float32).astype(np.float64)
    for t in range(1, 6):
        tr.step(X, T, 1e-3)
        _, g = mlp.loss_and_grad(P, xs, ts, wp, om)
        mlp.adam_step(P, g, m, v, t, lr=1e-3)
    got = tr.params().astype(np.float64)
    ref = learned.pack(P).astype(np.float64)
    assert np.max(np.abs(got - ref)) < 2e-5
    # the loss decreases on a fixed batch
    l0 = mlp.loss_and_grad(learned.unpack(p), xs, ts, wp, om)[0]
    assert tr.loss() < l0


def _mlp_block(nx=(40, 36), precision="fp64", seed=5):
    from idefix_amd.datablock import DataBlock
    g = DataBlock(len(nx), nx, comps=3, mhd=True, recon="mlp", solver="hlld",
                  emf="uct_contact", precision=precision, cfl=0.4)
    Vc, Vs = problems.turbulence(g, seed=seed)
    g.upload(Vc, Vs)
    g.set_boundaries()
    return g


@pytest.mark.parametrize("nx,d", [((40, 36), 0), ((40, 36), 1), ((20, 12, 10), 2)])
def test_mfma_inference_matches_bf16_oracle(nx, d):
    """The network's corrections as the GPU applies them (weights with the network minus
    the same kernel's PPM weights, on every variable whose corrected weights pass the
    check) against the oracle's BF16-rounded forward on the same FP32 stencils.  Only the
    FP32 accumulation order and the feature reciprocal (v_rcp_f32, 1 ulp) differ, so
    99.5 % of the corrections agree to 2e-5 (a lane-layout error would put whole groups
    of outputs off at O(0.1)); the rest are inputs whose BF16 rounding flips."""
    g = _mlp_block(nx)
    p = _params(7, scale_out=0.05)
    g.set_mlp_params(None)
    Wp = g.mlp_faces(d).cpu().numpy()
    g.set_mlp_params(p)
    W = g.mlp_faces(d).cpu().numpy()
    Vc, _ = g.download()
    lo, hi = g.beg[d], g.end[d] + 1
    S = learned.stencils(Vc, d, lo, hi)               # [F, rest..., 8, 6]
    S32 = S.astype(np.float32).astype(np.float64)
    x, _, _ = mlp.features(S32)
    y, _ = mlp.forward(learned.unpack(p), x.astype(np.float32), bf16_mode=True)
    take = lambda A: np.moveaxis(np.moveaxis(np.take(A, np.arange(lo, hi), axis=3 - d), 3 - d, 0), 1, -1)
    Wk, Wpk = take(W), take(Wp)
    assert Wk.shape == y.shape
    applied = Wk != Wpk
    assert np.mean(applied) > 0.25
    # the correction each weight receives: +z[2v] on wL, -z[2v] on wR
    shift = mlp.weights(y.astype(np.float64), np.zeros(y.shape))
    zg = (Wk - Wpk)[applied]
    zo = shift[applied]
    err = np.abs(zg - zo)
    assert np.mean(err <= 2e-5 * (1.0 + np.abs(zo))) >= 0.995, np.mean(err <= 2e-5)
    assert np.mean(err) < 2e-4, np.mean(err)
    # and the weights the kernel keeps are the oracle's checked ones where they agree
    wo = mlp.checked_weights(S32, mlp.weights(y.astype(np.float64), Wpk.astype(np.float64)))
    assert np.mean(np.abs(Wk - wo) <= 1e-4) >= 0.99


@pytest.mark.parametrize("d", [0, 1])
def test_no_network_gives_ppm_weights(d):
    g = _mlp_block()
    g.set_mlp_params(None)
    W = g.mlp_faces(d).cpu().numpy()
    Vc, _ = g.download()
    lo, hi = g.beg[d], g.end[d] + 1
    S = learned.stencils(Vc, d, lo, hi).astype(np.float32).astype(np.float64)
    w = mlp.ppm_weights(S)
    Wk = np.moveaxis(np.moveaxis(np.take(W, np.arange(lo, hi), axis=3 - d), 3 - d, 0), 1, -1)
    # the interface states the weights produce agree with PPM in float precision
    vL, vR = mlp.face_states(S, Wk.astype(np.float64))
    rL, rR = mlp.face_states(S, w)
    scale = np.max(np.abs(S), axis=(-1,))
    assert np.max(np.abs(vL - rL) / (scale + 1e-30)) < 1e-5
    assert np.max(np.abs(vR - rR) / (scale + 1e-30)) < 1e-5


def test_mlp_mode_with_fallback_weights_tracks_ppm():
    """MLP mode without a network runs PPM through the learned-limiter path: the state
    after five RK2 cycles matches the PPM-mode state to the float precision of the stored
    weights (an index or variable-mapping error would show at O(1))."""
    from idefix_amd.datablock import DataBlock
    a = _mlp_block()
    a.set_mlp_params(None)
    b = DataBlock(2, (40, 36), comps=3, mhd=True, recon="ppm", solver="hlld",
                  emf="uct_contact", precision="fp64", cfl=0.4)
    Vb, Sb = problems.turbulence(b, seed=5)
    b.upload(Vb, Sb)
    for _ in range(5):
        a.cycle()
        b.cycle()
    Va, Sa = a.download()
    Vb, Sb = b.download()
    ia = Va[:, :, a.beg[1]:a.end[1], a.beg[0]:a.end[0]]
    ib = Vb[:, :, b.beg[1]:b.end[1], b.beg[0]:b.end[0]]
    scale = np.max(np.abs(ib), axis=(1, 2, 3), keepdims=True)
    assert np.max(np.abs(ia - ib) / scale) < 1e-5


def test_learned_mode_conserves_and_keeps_div_b():
    g = _mlp_block(precision="fp32")
    g.set_mlp_params(_params(9))
    Vc0, _ = g.download()

    def totals():
        import torch
        U = g.Uc[:, :, g.beg[1]:g.end[1], g.beg[0]:g.end[0]].double()
        return [float(torch.sum(U[n])) for n in range(5)]

    g.cycle()
    t1 = totals()
    for _ in range(4):
        g.cycle()
    t5 = totals()
    for a, b in zip(t1, t5):
        assert abs(a - b) <= 1e-10 * max(1.0, abs(a))
    assert g.check_divb() < 1e-10
