"""The CPU reference of the learned-predictor MLP, pinned by central finite
differences of its own loss (float64, no bf16 rounding): every parameter
block's analytic gradient matches the difference quotient."""
import numpy as np

from mlp_ref import NPARAM, IN, IN_PAD, OUT, SHAPES, loss_grad, unpack, pack


def test_mlp_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    B = 8
    p = rng.standard_normal(NPARAM) * 0.2
    w = unpack(p)
    w["W1"][:, IN:] = 0.0
    p = pack(w)
    x = np.zeros((B, IN_PAD))
    x[:, :IN] = rng.standard_normal((B, IN))
    t = rng.standard_normal((B, OUT))
    mask = (rng.random((B, OUT)) < 0.7).astype(float)
    loss, g = loss_grad(p, x, t, mask, rounding=False)
    off = 0
    h = 1e-6
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        for idx in rng.choice(n, size=6, replace=False):
            if name == "W1" and (idx % IN_PAD) >= IN:
                continue
            e = np.zeros(NPARAM)
            e[off + idx] = h
            lp, _ = loss_grad(p + e, x, t, mask, rounding=False)
            lm, _ = loss_grad(p - e, x, t, mask, rounding=False)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[off + idx]) <= 1e-6 * max(1.0, abs(fd)), (name, idx, fd, g[off + idx])
        off += n
