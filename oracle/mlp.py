"""CPU restatement of the learned reconstruction (test infrastructure: only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline may use it).

The reference has no learned reconstruction; this module is the specification the HIP
kernels are checked against.  One interface x_{i-1/2} along a sweep direction sees the
6-cell stencil i-3 .. i+2 of the 8 primitives in the face frame
(rho, v_n, v_t, v_b, p, B_n, B_t, B_b).

Features (48): per variable, c = (V[i-1] + V[i]) / 2 and s = max_j |V_j - c| over the
stencil; x_j = (V_j - c) / s (0 when s = 0); ordered [variable][stencil cell].

Network: z1 = x W1 + b1 (48 -> 64), h1 = gelu(z1); z2 = h1 W2 + b2 (64 -> 64),
h2 = gelu(z2); z3 = h2 W3 + b3 (64 -> 16), linear outputs, with
gelu(z) = z * sigmoid(1.702 z) (Hendrycks & Gimpel 2016, sigmoid form).

Output: z3 corrects the limiter weights of PPM by a common shift of the two states of
each variable v: wL_v = wL_PPM + z3[2v], wR_v = wR_PPM - z3[2v] (output 2v+1 unused;
the jump vR - vL stays PPM's), and the interface states are
    vL_v = V_v[i-1] + wL_v (V_v[i] - V_v[i-1]),   vR_v = V_v[i] - wR_v (V_v[i] - V_v[i-1]).
Van Leer PLM is the member wL = dvm / (dvm + dvp) (monotone data) or 0 (extremum); PPM is
the member w = (v_PPM - V[i-1]) / (V[i] - V[i-1]) (0 where V[i] = V[i-1]).  Each
variable's pair of weights is checked against the TVD region of slope limiters (Sweby
1984: limited slope between 0 and twice each one-sided difference): the left state's
deviation wL (V[i] - V[i-1]) must have the sign of V[i-1] - V[i-2] and not exceed it in
magnitude (or be 0), the right state's deviation wR (V[i] - V[i-1]) likewise against
V[i+1] - V[i], and 0 <= w <= 1.  A variable that fails, or has non-finite weights, takes
the PPM weights (the fallback; van Leer with `fallback="plm"`).

Training target: the interface value t_v of a finer solution; loss per interface
sum_v om_v^2 [((vL_v - t_v)/s_v)^2 + ((vR_v - t_v)/s_v)^2] (both terms 0 when s_v = 0),
mean over the batch, with the unchecked weights; om_v are per-sample loss weights (1 for
the scale-free loss, s_v / sigma_v to weigh absolute errors with a per-variable scale).

Inference on the GPU runs with BF16 inputs, weights and hidden activations and FP32
accumulation; `forward(..., bf16=True)` rounds at the same points.
"""
import numpy as np

NF, NH, NO = 48, 64, 16
GELU_K = 1.702


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def gelu(z):
    return z * sigmoid(GELU_K * z)


def gelu_grad(z):
    s = sigmoid(GELU_K * z)
    return s + GELU_K * z * s * (1.0 - s)


def bf16(x):
    """Round float32 values to bfloat16 (round to nearest, ties to even), returned as
    float32."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    rounding = ((u >> 16) & 1) + 0x7FFF
    r = ((u + rounding) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    nan = np.isnan(a)
    if np.any(nan):
        out = out.copy()
        out[nan] = np.nan
    return out


def init_params(seed=0):
    """Glorot-uniform weights, zero biases (float64)."""
    rng = np.random.default_rng(seed)

    def glorot(n_in, n_out):
        lim = np.sqrt(6.0 / (n_in + n_out))
        return rng.uniform(-lim, lim, size=(n_in, n_out))

    return {"W1": glorot(NF, NH), "b1": np.zeros(NH), "W2": glorot(NH, NH), "b2": np.zeros(NH),
            "W3": glorot(NH, NO), "b3": np.zeros(NO)}


def features(S):
    """S: [..., 8, 6] stencil (variables x cells i-3..i+2).  Returns (x [..., 48], c, s)
    with c, s of shape [..., 8]."""
    S = np.asarray(S, dtype=np.float64)
    c = 0.5 * (S[..., 2] + S[..., 3])
    s = np.max(np.abs(S - c[..., None]), axis=-1)
    safe = np.where(s > 0, s, 1.0)
    x = np.where(s[..., None] > 0, (S - c[..., None]) / safe[..., None], 0.0)
    return x.reshape(S.shape[:-2] + (NF,)), c, s


def forward(P, x, bf16_mode=False):
    """Returns (y, cache).  bf16_mode rounds x, the weights and h1, h2 to bfloat16 and
    accumulates in float32, as the MFMA inference does."""
    if bf16_mode:
        f = np.float32
        xb = bf16(x.astype(f))
        W1, W2, W3 = (bf16(P[k].astype(f)) for k in ("W1", "W2", "W3"))
        b1, b2, b3 = (P[k].astype(f) for k in ("b1", "b2", "b3"))
        z1 = (xb.astype(np.float64) @ W1.astype(np.float64)).astype(f) + b1
        h1 = bf16(gelu(z1).astype(f))
        z2 = (h1.astype(np.float64) @ W2.astype(np.float64)).astype(f) + b2
        h2 = bf16(gelu(z2).astype(f))
        z3 = (h2.astype(np.float64) @ W3.astype(np.float64)).astype(f) + b3
        return z3, None
    z1 = x @ P["W1"] + P["b1"]
    h1 = gelu(z1)
    z2 = h1 @ P["W2"] + P["b2"]
    h2 = gelu(z2)
    z3 = h2 @ P["W3"] + P["b3"]
    return z3, (x, z1, h1, z2, h2, z3)


def weights(y, wp):
    """Limiter weights from the network outputs y and the PPM weights wp: output 2v shifts
    both states of variable v, wL = wpL + y[2v], wR = wpR - y[2v] (output 2v+1 unused)."""
    w = np.array(wp, dtype=np.float64, copy=True)
    w[..., 0::2] = wp[..., 0::2] + y[..., 0::2]
    w[..., 1::2] = wp[..., 1::2] - y[..., 0::2]
    return w


def face_states(S, w):
    """Interface states (vL, vR) [..., 8] from the stencil S [..., 8, 6] and weights w."""
    a = S[..., 2]
    b = S[..., 3]
    wl = w[..., 0::2]
    wr = w[..., 1::2]
    return a + wl * (b - a), b - wr * (b - a)


def checked_weights(S, w, fallback="ppm"):
    """The weights the solver uses: per variable, the fallback's weights where a weight
    leaves the TVD region or is not finite."""
    a, b, c, d = S[..., 1], S[..., 2], S[..., 3], S[..., 4]
    jump = c - b
    wl = w[..., 0::2]
    wr = w[..., 1::2]
    dl = wl * jump
    dr = wr * jump
    okl = (wl >= 0) & (wl <= 1) & ((dl == 0) | ((dl * (b - a) > 0) & (np.abs(dl) <= np.abs(b - a))))
    okr = (wr >= 0) & (wr <= 1) & ((dr == 0) | ((dr * (d - c) > 0) & (np.abs(dr) <= np.abs(d - c))))
    ok = (okl & okr) | (jump == 0)
    ok &= np.isfinite(wl) & np.isfinite(wr)
    fb = ppm_weights(S) if fallback == "ppm" else van_leer_weights(S)
    out = np.array(w, dtype=np.float64, copy=True)
    out[..., 0::2] = np.where(ok, wl, fb[..., 0::2])
    out[..., 1::2] = np.where(ok, wr, fb[..., 1::2])
    return out


def ppm_weights(S):
    """The weights that reproduce PPM (getPPMStates, slopeLimiter.hpp:142-191, with the
    van Leer fallback of extrapolateToFaces.hpp:260-304 where rho or p would not be
    positive): the left state of the interface is cell i-1's right value (stencil cells
    0..4), the right state cell i's left value (cells 1..5)."""
    from . import oracle as O
    ppm = np.vectorize(O.ppm_states)
    vl_vec = np.vectorize(O.vanleer)
    _, left = ppm(S[..., 0], S[..., 1], S[..., 2], S[..., 3], S[..., 4])
    right, _ = ppm(S[..., 1], S[..., 2], S[..., 3], S[..., 4], S[..., 5])
    left = np.array(left, dtype=np.float64)
    right = np.array(right, dtype=np.float64)
    for q in (0, 4):  # rho, p in the face frame
        bad = left[..., q] <= 0
        plm = S[..., q, 2] + 0.5 * vl_vec(S[..., q, 3] - S[..., q, 2], S[..., q, 2] - S[..., q, 1])
        left[..., q] = np.where(bad, plm, left[..., q])
        bad = right[..., q] <= 0
        plm = S[..., q, 3] - 0.5 * vl_vec(S[..., q, 4] - S[..., q, 3], S[..., q, 3] - S[..., q, 2])
        right[..., q] = np.where(bad, plm, right[..., q])
    jump = S[..., 3] - S[..., 2]
    safe = np.where(jump != 0, jump, 1.0)
    w = np.empty(S.shape[:-2] + (16,))
    w[..., 0::2] = np.where(jump != 0, (left - S[..., 2]) / safe, 0.0)
    w[..., 1::2] = np.where(jump != 0, (S[..., 3] - right) / safe, 0.0)
    return w


def loss_and_grad(P, x, tau, wp, om=None):
    """Loss (mean over the batch) and parameter gradients.  x: [B, 48] features,
    tau: [B, 8] normalised targets (t - c) / s (0 where s = 0), wp: [B, 16] PPM weights,
    om: [B, 8] loss weights (1 when None)."""
    B = x.shape[0]
    om2 = np.ones((B, 8)) if om is None else np.asarray(om, dtype=np.float64) ** 2
    y, (x, z1, h1, z2, h2, z3) = forward(P, x)
    xv = x.reshape(B, 8, 6)
    d = xv[:, :, 3] - xv[:, :, 2]
    w = weights(y, wp)
    eL = xv[:, :, 2] + w[:, 0::2] * d - tau
    eR = xv[:, :, 3] - w[:, 1::2] * d - tau
    loss = np.sum(om2 * (eL ** 2 + eR ** 2)) / B
    dz3 = np.zeros_like(y)
    dz3[:, 0::2] = om2 * 2.0 * (eL + eR) * d / B
    g = {"W3": h2.T @ dz3, "b3": dz3.sum(0)}
    dh2 = dz3 @ P["W3"].T
    dz2 = dh2 * gelu_grad(z2)
    g["W2"] = h1.T @ dz2
    g["b2"] = dz2.sum(0)
    dh1 = dz2 @ P["W2"].T
    dz1 = dh1 * gelu_grad(z1)
    g["W1"] = x.T @ dz1
    g["b1"] = dz1.sum(0)
    return loss, g


def adam_step(P, g, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015) in place; t is the 1-based step count."""
    for k in P:
        m[k] = b1 * m[k] + (1 - b1) * g[k]
        v[k] = b2 * v[k] + (1 - b2) * g[k] * g[k]
        mh = m[k] / (1 - b1 ** t)
        vh = v[k] / (1 - b2 ** t)
        P[k] = P[k] - lr * mh / (np.sqrt(vh) + eps)
    return P


def van_leer_weights(S):
    """The weights w that reproduce van Leer PLM (slopeLimiter.hpp:54-57) for the two
    states of the interface: wL from cell i-1's slope, wR from cell i's slope."""
    a, b = S[..., 1], S[..., 2]
    c, d = S[..., 3], S[..., 4]
    jump = c - b

    def vl(dp, dm):
        return np.where(dp * dm > 0, 2.0 * dp * dm / np.where(dp + dm != 0, dp + dm, 1.0), 0.0)

    slope_l = vl(c - b, b - a)   # cell i-1
    slope_r = vl(d - c, c - b)   # cell i
    safe = np.where(jump != 0, jump, 1.0)
    wl = np.where(jump != 0, 0.5 * slope_l / safe, 0.0)
    wr = np.where(jump != 0, 0.5 * slope_r / safe, 0.0)
    y = np.empty(S.shape[:-2] + (16,))
    y[..., 0::2] = wl
    y[..., 1::2] = wr
    return y
