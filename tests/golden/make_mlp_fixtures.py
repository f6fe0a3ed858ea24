"""Writes tests/golden/mlp_autograd.npz: loss and parameter gradients of the learned
reconstruction's training loss (PPM weights shifted by the network, per-sample
loss weights) computed by torch autograd on the CPU (float64), an
independent check of oracle/mlp.py's hand-written backward pass.

Run from the repository root:  python tests/golden/make_mlp_fixtures.py
"""
import os

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    torch.set_default_dtype(torch.float64)
    rng = np.random.default_rng(2024)
    B = 8
    out = {}
    for seed in range(2):
        P = {"W1": rng.normal(0, 0.3, (48, 64)), "b1": rng.normal(0, 0.1, 64),
             "W2": rng.normal(0, 0.2, (64, 64)), "b2": rng.normal(0, 0.1, 64),
             "W3": rng.normal(0, 0.3, (64, 16)), "b3": rng.normal(0, 0.1, 16)}
        x = rng.uniform(-1, 1, (B, 48))
        tau = rng.uniform(-1, 1, (B, 8))
        wp = rng.uniform(0, 1, (B, 16))
        om = rng.uniform(0.2, 2.0, (B, 8))
        T = {k: torch.tensor(v, requires_grad=True) for k, v in P.items()}
        xt = torch.tensor(x)
        tt = torch.tensor(tau)

        def gelu(z):
            return z * torch.sigmoid(1.702 * z)

        h1 = gelu(xt @ T["W1"] + T["b1"])
        h2 = gelu(h1 @ T["W2"] + T["b2"])
        y = h2 @ T["W3"] + T["b3"]
        xv = xt.reshape(B, 8, 6)
        d = xv[:, :, 3] - xv[:, :, 2]
        shift = torch.repeat_interleave(y[:, 0::2], 2, dim=1) * torch.tensor([1.0, -1.0] * 8)
        w = torch.tensor(wp) + shift
        eL = xv[:, :, 2] + w[:, 0::2] * d - tt
        eR = xv[:, :, 3] - w[:, 1::2] * d - tt
        loss = (torch.tensor(om) ** 2 * (eL ** 2 + eR ** 2)).sum() / B
        loss.backward()
        for k, v in P.items():
            out["s%d_%s" % (seed, k)] = v
            out["s%d_g%s" % (seed, k)] = T[k].grad.numpy()
        out["s%d_x" % seed] = x
        out["s%d_tau" % seed] = tau
        out["s%d_wp" % seed] = wp
        out["s%d_om" % seed] = om
        out["s%d_loss" % seed] = np.array(loss.item())
        out["s%d_y" % seed] = y.detach().numpy()
    np.savez_compressed(os.path.join(HERE, "mlp_autograd.npz"), **out)


if __name__ == "__main__":
    main()
