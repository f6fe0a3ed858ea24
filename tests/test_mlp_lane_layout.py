"""CPU emulation of the learned-limiter inference kernel's MFMA operand layout.

k_mlp_infer (idefix_amd/csrc/mlp.hpp) runs the 48 -> 64 -> 64 -> 16 network of 16
interfaces per tile as v_mfma_f32_16x16x16_bf16 products whose operands are built lane by
lane: lane l = 16 g + r holds face r of the tile and the face-frame variables 2g, 2g + 1
(features 12 g .. 12 g + 11), the weights enter as the A operand (rows = output units) and
the activations as the B operand (columns = faces).  The instruction's operand layout
(CDNA3/CDNA4 ISA, 16x16x16 with 4 BF16 per lane): lane l holds A[l % 16][4 (l / 16) + j]
and B[4 (l / 16) + j][l % 16] for j = 0..3, and receives D[4 (l / 16) + t][l % 16] for
t = 0..3.

The emulation builds every lane's fragments exactly as the kernel indexes them, applies
the instruction's layout, and must reproduce the CPU restatement's BF16 forward
(oracle/mlp.py, forward(..., bf16_mode=True)).  A mismatched layout (A and B fragments
disagreeing on which input a K slot holds) fails it.

The same emulation covers the earlier variant on v_mfma_f32_16x16x32_bf16 (8 BF16 per
lane: two 4-element halves concatenated), which missed the GPU inference test by a mean
1.1e-2 when built for 6 waves per SIMD: its fragments give the oracle's outputs under
either K layout the 8-element operand can have (8 consecutive K per lane group, or two
16-wide halves), because A and B are built from the same lane-group order; so its miss
was not an operand-layout error of the source.
"""
import numpy as np
import pytest

from oracle import mlp

F32 = np.float32


def _frag_layout(kind):
    """K index held by (lane group g = l / 16, element j) of an MFMA operand."""
    if kind == "x16":            # 16x16x16: 4 elements per lane
        return lambda g, j: 4 * g + j
    if kind == "x32_consecutive":  # 16x16x32: 8 consecutive K per lane group
        return lambda g, j: 8 * g + j
    if kind == "x32_halves":     # 16x16x32: two 16x16x16 halves
        return lambda g, j: 4 * g + j if j < 4 else 16 + 4 * g + (j - 4)
    raise ValueError(kind)


def mfma(a, b, c, layout_a, layout_b=None):
    """D = A B + C for per-lane fragments a, b [64, E] (E = 4 or 8) and accumulators
    c [64, 4] (float32); products exact, sum in float64, rounded to float32.  layout_b
    (default: layout_a) lets a test give the B operand a different K layout."""
    layout_b = layout_b or layout_a
    E = a.shape[1]
    K = 4 * E
    A = np.zeros((16, K))
    B = np.zeros((K, 16))
    for l in range(64):
        g, r = divmod(l, 16)
        for j in range(E):
            A[r, layout_a(g, j)] = a[l, j]
            B[layout_b(g, j), r] = b[l, j]
    D = A @ B
    out = np.empty((64, 4), dtype=F32)
    for l in range(64):
        g, r = divmod(l, 16)
        for t in range(4):
            out[l, t] = F32(D[4 * g + t, r] + np.float64(c[l, t]))
    return out


def _net(P):
    W1t = mlp.bf16(P["W1"].astype(F32)).T   # [64 out][48 in]
    W2t = mlp.bf16(P["W2"].astype(F32)).T   # [64][64]
    W3t = mlp.bf16(P["W3"].astype(F32)).T   # [16][64]
    return W1t, W2t, W3t, [P[k].astype(F32) for k in ("b1", "b2", "b3")]


def _lanes():
    return [divmod(l, 16) for l in range(64)]  # (g, r)


def infer_x16(P, X, layout_a="x16", layout_b="x16"):
    """The current kernel's tile: X [16 faces, 48] -> outputs [16, 16]."""
    W1t, W2t, W3t, (b1, b2, b3) = _net(P)
    la = layout_a if callable(layout_a) else _frag_layout(layout_a)
    lb = layout_b if callable(layout_b) else _frag_layout(layout_b)
    xb = mlp.bf16(X.astype(F32))
    bx = np.zeros((3, 64, 4), dtype=F32)        # bx[ks][lane][j] = feature 12 g + 4 ks + j
    for l, (g, r) in enumerate(_lanes()):
        for ks in range(3):
            bx[ks, l] = xb[r, 12 * g + 4 * ks: 12 * g + 4 * ks + 4]
    h1 = np.zeros((4, 64, 4), dtype=F32)
    for mt in range(4):
        acc = np.zeros((64, 4), dtype=F32)
        for ks in range(3):
            a = np.array([W1t[16 * mt + r, 12 * g + 4 * ks: 12 * g + 4 * ks + 4] for g, r in _lanes()])
            acc = mfma(a, bx[ks], acc, la, lb)
        for l, (g, r) in enumerate(_lanes()):
            h1[mt, l] = mlp.bf16(mlp.gelu(acc[l] + b1[16 * mt + 4 * g: 16 * mt + 4 * g + 4]).astype(F32))
    h2 = np.zeros((4, 64, 4), dtype=F32)
    for mt in range(4):
        acc = np.zeros((64, 4), dtype=F32)
        for ks in range(4):
            a = np.array([W2t[16 * mt + r, 16 * ks + 4 * g: 16 * ks + 4 * g + 4] for g, r in _lanes()])
            acc = mfma(a, h1[ks], acc, la, lb)
        for l, (g, r) in enumerate(_lanes()):
            h2[mt, l] = mlp.bf16(mlp.gelu(acc[l] + b2[16 * mt + 4 * g: 16 * mt + 4 * g + 4]).astype(F32))
    acc = np.zeros((64, 4), dtype=F32)
    for ks in range(4):
        a = np.array([W3t[r, 16 * ks + 4 * g: 16 * ks + 4 * g + 4] for g, r in _lanes()])
        acc = mfma(a, h2[ks], acc, la, lb)
    y = np.zeros((16, 16), dtype=F32)
    for l, (g, r) in enumerate(_lanes()):
        y[r, 4 * g: 4 * g + 4] = acc[l] + b3[4 * g: 4 * g + 4]
    return y


def infer_x32(P, X, layout):
    """The earlier 16x16x32 variant's tile (its fragment construction, linear outputs)."""
    W1t, W2t, W3t, (b1, b2, b3) = _net(P)
    l32, l16 = _frag_layout(layout), _frag_layout("x16")
    xb = mlp.bf16(X.astype(F32))
    x_lane = np.array([xb[r, 12 * g: 12 * g + 12] for g, r in _lanes()])  # [64, 12]
    h1 = np.zeros((4, 64, 4), dtype=F32)
    for mt in range(4):
        # one 32-step over features 12 g + 0..7, one 16-step over 12 g + 8..11
        a = np.array([W1t[16 * mt + r, 12 * g: 12 * g + 8] for g, r in _lanes()])
        acc = mfma(a, x_lane[:, 0:8], np.zeros((64, 4), F32), l32)
        a = np.array([W1t[16 * mt + r, 12 * g + 8: 12 * g + 12] for g, r in _lanes()])
        acc = mfma(a, x_lane[:, 8:12], acc, l16)
        for l, (g, r) in enumerate(_lanes()):
            h1[mt, l] = mlp.bf16(mlp.gelu(acc[l] + b1[16 * mt + 4 * g: 16 * mt + 4 * g + 4]).astype(F32))

    def layer(Wt, rows, h, b, nout_tiles):
        out = np.zeros((nout_tiles, 64, 4), dtype=F32)
        for mt in range(nout_tiles):
            acc = np.zeros((64, 4), dtype=F32)
            for st in range(2):
                a = np.array([np.concatenate([Wt[rows(mt, r), 32 * st + 4 * g: 32 * st + 4 * g + 4],
                                              Wt[rows(mt, r), 32 * st + 16 + 4 * g: 32 * st + 16 + 4 * g + 4]])
                              for g, r in _lanes()])
                bb = np.concatenate([h[2 * st], h[2 * st + 1]], axis=1)
                acc = mfma(a, bb, acc, l32)
            out[mt] = acc
        return out
    z2 = layer(W2t, lambda mt, r: 16 * mt + r, h1, b2, 4)
    h2 = np.zeros((4, 64, 4), dtype=F32)
    for mt in range(4):
        for l, (g, r) in enumerate(_lanes()):
            h2[mt, l] = mlp.bf16(mlp.gelu(z2[mt, l] + b2[16 * mt + 4 * g: 16 * mt + 4 * g + 4]).astype(F32))
    z3 = layer(W3t, lambda mt, r: r, h2, b3, 1)[0]
    y = np.zeros((16, 16), dtype=F32)
    for l, (g, r) in enumerate(_lanes()):
        y[r, 4 * g: 4 * g + 4] = z3[l] + b3[4 * g: 4 * g + 4]
    return y


def _case(seed):
    rng = np.random.default_rng(seed)
    P = mlp.init_params(seed)
    for k in ("b1", "b2", "b3"):
        P[k] = rng.normal(0, 0.1, P[k].shape)
    P["W3"] = rng.normal(0, 0.3, P["W3"].shape)
    X = rng.uniform(-1, 1, (16, 48))
    want, _ = mlp.forward(P, X, bf16_mode=True)
    return P, X, want


@pytest.mark.parametrize("seed", [0, 1])
def test_current_kernel_layout_reproduces_the_oracle(seed):
    P, X, want = _case(seed)
    got = infer_x16(P, X)
    scale = np.abs(want).max()
    assert np.abs(got - want).max() < 2e-6 * scale


@pytest.mark.parametrize("layout", ["x32_consecutive", "x32_halves"])
def test_earlier_x32_variant_layout_reproduces_the_oracle(layout):
    P, X, want = _case(2)
    got = infer_x32(P, X, layout)
    assert np.abs(got - want).max() < 2e-6 * np.abs(want).max()


def test_mismatched_fragment_layout_is_detected():
    # A and B fragments that disagree on which input a K slot holds change the result
    # well beyond the 2e-6 tolerance (the check above would fail on such a kernel)
    P, X, want = _case(3)
    perm = lambda g, j: 4 * g + (j + 1) % 4
    got = infer_x16(P, X, layout_a="x16", layout_b=perm)
    assert np.abs(got - want).max() > 1e-2 * np.abs(want).max()
