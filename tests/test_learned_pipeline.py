"""CPU checks of the learned limiter's data pipeline (torch on the CPU): block averaging
conserves the totals and keeps the discrete div B at zero, interface targets are exact
for linear data, and the samples' features and targets follow the specification."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from idefix_amd import learned  # noqa: E402
from oracle import mlp  # noqa: E402


def _fine(n=32, seed=0):
    rng = np.random.default_rng(seed)
    V = torch.tensor(rng.uniform(0.5, 1.5, (8, n, n)))
    # divergence-free faces from a corner potential A (periodic)
    A = rng.normal(0, 1, (n, n))
    bx = np.zeros((n, n + 1))
    by = np.zeros((n + 1, n))
    bx[:, :n] = (np.roll(A, -1, axis=0) - A) * n
    bx[:, n] = bx[:, 0]
    by[:n, :] = -(np.roll(A, -1, axis=1) - A) * n
    by[n, :] = by[0, :]
    V[5] = torch.tensor(0.5 * (bx[:, :-1] + bx[:, 1:]))
    V[6] = torch.tensor(0.5 * (by[:-1] + by[1:]))
    return V, torch.tensor(bx), torch.tensor(by)


def test_coarse_state_conserves_and_keeps_div_b():
    gamma = 5.0 / 3.0
    V, bx, by = _fine()
    Vc, bxc, byc = learned.coarse_state(V, bx, by, 4, gamma)
    rho, v, p, B = V[0], V[1:4], V[4], V[5:8]
    E = p / (gamma - 1) + 0.5 * rho * (v * v).sum(0) + 0.5 * (B * B).sum(0)
    rc, vc, pc, Bc = Vc[0], Vc[1:4], Vc[4], Vc[5:8]
    assert float(rc.sum()) * 16 == pytest.approx(float(rho.sum()), rel=1e-13)
    for c in range(3):
        assert float((rc * vc[c]).sum()) * 16 == pytest.approx(float((rho * v[c]).sum()), rel=1e-12, abs=1e-12)
    # energy with the coarse cell field from the coarse faces is the averaged energy
    Ec = pc / (gamma - 1) + 0.5 * rc * (vc * vc).sum(0) + 0.5 * (Bc * Bc).sum(0)
    assert float(Ec.sum()) * 16 == pytest.approx(float(E.sum()), rel=1e-12)
    div = (bxc[:, 1:] - bxc[:, :-1]) + (byc[1:] - byc[:-1])
    assert float(div.abs().max()) < 1e-11 * float(bxc.abs().max())


def test_face_targets_exact_for_linear_data():
    n, f = 32, 4
    x = (np.arange(n) + 0.5) / n
    V = torch.zeros((8, n, n), dtype=torch.float64)
    for q in range(8):
        V[q] = torch.tensor(np.tile(1.0 + 0.1 * q + 0.3 * x, (n, 1)))      # linear in x
    bx = torch.zeros((n, n + 1), dtype=torch.float64)
    by = torch.zeros((n + 1, n), dtype=torch.float64)
    tx, ty = learned.face_targets(V, bx, by, f)
    xf = np.arange(n // f) * f / n   # coarse x faces
    for q in range(8):
        if q == 5:
            continue  # the normal field comes from the faces
        want = 1.0 + 0.1 * q + 0.3 * xf
        got = tx[q].numpy()
        # interior faces (the periodic wrap makes face 0 a jump)
        assert np.allclose(got[:, 1:], want[None, 1:], rtol=0, atol=1e-14)


def test_samples_follow_the_feature_specification():
    V, bx, by = _fine(16, seed=3)
    Vc, _, _ = learned.coarse_state(V, bx, by, 2, 5.0 / 3.0)
    tgt = torch.rand((8, 8, 8), dtype=torch.float64)
    wp = torch.rand((16, 8, 8), dtype=torch.float64)
    X, T, Sc = learned.samples(Vc, tgt, 0, wp)
    assert X.shape == (64, 72) and T.shape == (64, 8) and Sc.shape == (64, 8)
    # face (J=2, I=5): stencil cells I-3 .. I+2 along x in face-frame order
    S = Vc[learned.FRAME[0]][:, 2, :].numpy()
    st = np.stack([S[:, (5 - 3 + s) % 8] for s in range(6)], axis=-1)
    x, c, s = mlp.features(st)
    k = 2 * 8 + 5
    assert np.allclose(X[k, :48].numpy(), x.astype(np.float32), atol=1e-6)
    # the PPM weights of the same face, and unit loss weights
    assert np.allclose(X[k, 48:64].numpy(), wp[:, 2, 5].numpy().astype(np.float32))
    assert np.all(X[k, 64:72].numpy() == 1.0)
    assert np.allclose(Sc[k].numpy(), s, rtol=1e-6)
    tau = np.clip((tgt[:, 2, 5].numpy()[learned.FRAME[0]] - c) / s, -1, 1)
    assert np.allclose(T[k].numpy(), tau, atol=1e-6)


def test_coarse_state3_conserves_and_keeps_div_b():
    from idefix_amd import learned3d
    gamma = 5.0 / 3.0
    n, f = 16, 4
    rng = np.random.default_rng(3)
    V = torch.tensor(rng.uniform(0.5, 1.5, (8, n, n, n)))
    # divergence-free faces from an edge-centred vector potential (periodic): B = curl A
    A = [rng.normal(0, 1, (n, n, n)) for _ in range(3)]   # Ax, Ay, Az on x, y, z edges
    d = lambda a, ax: (np.roll(a, -1, axis=ax) - a) * n    # forward difference
    bx = np.zeros((n, n, n + 1))
    by = np.zeros((n, n + 1, n))
    bz = np.zeros((n + 1, n, n))
    bx[:, :, :n] = d(A[2], 1) - d(A[1], 0)      # dAz/dy - dAy/dz  (axes: 0=z, 1=y, 2=x)
    by[:, :n, :] = d(A[0], 0) - d(A[2], 2)      # dAx/dz - dAz/dx
    bz[:n] = d(A[1], 2) - d(A[0], 1)            # dAy/dx - dAx/dy
    bx[:, :, n], by[:, n, :], bz[n] = bx[:, :, 0], by[:, 0, :], bz[0]
    div = (bx[:, :, 1:] - bx[:, :, :-1]) + (by[:, 1:] - by[:, :-1]) + (bz[1:] - bz[:-1])
    assert np.abs(div).max() < 1e-10 * np.abs(bx).max()
    V[5] = torch.tensor(0.5 * (bx[..., :-1] + bx[..., 1:]))
    V[6] = torch.tensor(0.5 * (by[:, :-1] + by[:, 1:]))
    V[7] = torch.tensor(0.5 * (bz[:-1] + bz[1:]))
    bx, by, bz = torch.tensor(bx), torch.tensor(by), torch.tensor(bz)
    Vc, bxc, byc, bzc = learned3d.coarse_state3(V, bx, by, bz, f, gamma)
    m = f ** 3
    rho, v, p, B = V[0], V[1:4], V[4], V[5:8]
    E = p / (gamma - 1) + 0.5 * rho * (v * v).sum(0) + 0.5 * (B * B).sum(0)
    rc, vc, pc, Bc = Vc[0], Vc[1:4], Vc[4], Vc[5:8]
    assert float(rc.sum()) * m == pytest.approx(float(rho.sum()), rel=1e-13)
    for c in range(3):
        assert float((rc * vc[c]).sum()) * m == pytest.approx(float((rho * v[c]).sum()), rel=1e-12, abs=1e-12)
    Ec = pc / (gamma - 1) + 0.5 * rc * (vc * vc).sum(0) + 0.5 * (Bc * Bc).sum(0)
    assert float(Ec.sum()) * m == pytest.approx(float(E.sum()), rel=1e-12)
    # the coarse faces are the mean of their f^2 fine faces, and div B stays zero
    assert float(bxc[1, 2, 3]) == pytest.approx(float(bx[f:2 * f, 2 * f:3 * f, 3 * f].mean()), rel=1e-14)
    assert float(bzc[2, 0, 1]) == pytest.approx(float(bz[2 * f, 0:f, f:2 * f].mean()), rel=1e-14)
    divc = (bxc[..., 1:] - bxc[..., :-1]) + (byc[:, 1:] - byc[:, :-1]) + (bzc[1:] - bzc[:-1])
    assert float(divc.abs().max()) < 1e-10 * float(bxc.abs().max())
    # the wrong pooling (faces of every fine plane, not every f-th) breaks div B
    bx_bad = bx[:, :, :n].reshape(n // f, f, n // f, f, n // f, f).mean((1, 3, 5))
    bx_bad = torch.cat([bx_bad, bx_bad[..., :1]], dim=-1)
    div_bad = (bx_bad[..., 1:] - bx_bad[..., :-1]) + (byc[:, 1:] - byc[:, :-1]) + (bzc[1:] - bzc[:-1])
    assert float(div_bad.abs().max()) > 1e-3 * float(bxc.abs().max())


def test_face_targets3_and_samples3_on_linear_data():
    """3-D interface targets are exact for data linear along the face normal (away from
    the periodic wrap), and the 3-D samples follow the 2-D specification: stencil
    features of the face-frame variables and normalised targets."""
    from idefix_amd import learned3d
    n, f = 16, 4
    N = n // f
    x = (np.arange(n) + 0.5) / n
    for d in range(3):
        V = torch.zeros((8, n, n, n), dtype=torch.float64)
        shape = [1, 1, 1]
        shape[2 - d] = n
        for q in range(8):
            V[q] = torch.tensor((1.0 + 0.1 * q + 0.3 * x).reshape(shape)).expand(n, n, n)
        bx = torch.zeros((n, n, n + 1), dtype=torch.float64)
        by = torch.zeros((n, n + 1, n), dtype=torch.float64)
        bz = torch.zeros((n + 1, n, n), dtype=torch.float64)
        tg = learned3d.face_targets3(V, bx, by, bz, f)[d]
        xf = np.arange(N) * f / n
        frame = learned.FRAME[d]
        normal_b = 5 + d
        for qf in range(8):
            q = frame[qf]
            if q == normal_b:
                continue
            want = 1.0 + 0.1 * q + 0.3 * xf
            got = np.moveaxis(tg[qf].numpy(), 2 - d, -1)[..., 1:]
            assert np.allclose(got, np.broadcast_to(want[1:], got.shape), rtol=0, atol=1e-14)
        # samples of a coarse state linear along d, with the exact face values as
        # targets: features (j - 2.5) / 2.5, normalised targets 0
        Nc = 8
        Vc = torch.zeros((8, Nc, Nc, Nc), dtype=torch.float64)
        xc = (np.arange(Nc) + 0.5) / Nc
        shp = [1, 1, 1]
        shp[2 - d] = Nc
        for q in range(8):
            Vc[q] = torch.tensor((1.0 + 0.1 * q + 0.3 * xc).reshape(shp)).expand(Nc, Nc, Nc)
        face = [0, 0, 0]
        face[2 - d] = 4  # stencil cells 1 .. 6, no periodic wrap
        tgt = torch.zeros((8, Nc, Nc, Nc), dtype=torch.float64)
        for qf, q in enumerate(learned.FRAME[d]):
            tgt[qf] = 1.0 + 0.1 * q + 0.3 * (4.0 / Nc)
        wp = torch.zeros((16, Nc, Nc, Nc), dtype=torch.float64)
        rows, tau = learned3d.samples3(Vc, tgt, d, wp)
        assert rows.shape == (Nc ** 3, 72) and tau.shape == (Nc ** 3, 8)
        r = rows.reshape(Nc, Nc, Nc, 72)[tuple(face)]
        feats = r[:48].reshape(8, 6).numpy()
        want = (np.arange(6) - 2.5) / 2.5
        assert np.allclose(feats, np.broadcast_to(want, (8, 6)), atol=1e-6)
        assert np.allclose(r[64:72].numpy(), 1.0)
        assert np.allclose(tau.reshape(Nc, Nc, Nc, 8)[tuple(face)].numpy(), 0.0, atol=1e-5)
