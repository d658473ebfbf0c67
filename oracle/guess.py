"""CPU oracle of the learned initial guess. TEST INFRASTRUCTURE ONLY.

Plain numpy FP64 restatement of the per-cell MLP guess, used to check the HIP forward and
backward kernels and the feature / extrapolation kernels. Only tests/ and the smoke check use
it; the product never does.

The extrapolated guess it corrects is the one of the reference's dual-splitting scheme:
p0 = sum_i beta_i p^{n-i} (time_int_bdf_dual_splitting.cpp:489-501) with the constant-step
constants of ExtrapolationConstants::set_constant_time_step (extrapolation_constants.cpp:62-83).
The network itself has no reference counterpart.
"""
import numpy as np

# extrapolation_constants.cpp:62-83, orders 1..4 (constant time step)
BETA = {1: (1.0,), 2: (2.0, -1.0), 3: (3.0, -3.0, 1.0), 4: (4.0, -6.0, 4.0, -1.0)}


def unpack(params, n_features, width, n_out):
    """Split the flat parameter vector [W1|b1|W2|b2|W3|b3] (W row-major out x in)."""
    p = np.asarray(params, dtype=np.float64)
    o = 0
    shapes = [(width, n_features), (width,), (width, width), (width,), (n_out, width), (n_out,)]
    out = []
    for sh in shapes:
        k = int(np.prod(sh))
        out.append(p[o:o + k].reshape(sh))
        o += k
    assert o == p.size
    return out


def forward(params, X, width, n_out):
    W1, b1, W2, b2, W3, b3 = unpack(params, X.shape[1], width, n_out)
    Z1 = X @ W1.T + b1
    H1 = np.maximum(Z1, 0.0)
    Z2 = H1 @ W2.T + b2
    H2 = np.maximum(Z2, 0.0)
    Y = H2 @ W3.T + b3
    return Y, (Z1, H1, Z2, H2)


def loss_and_grad(params, X, T, width, n_out):
    """Mean squared error over all entries and its gradient (flat, same layout)."""
    W1, b1, W2, b2, W3, b3 = unpack(params, X.shape[1], width, n_out)
    Y, (Z1, H1, Z2, H2) = forward(params, X, width, n_out)
    n = Y.size
    dY = 2.0 * (Y - T) / n
    gW3 = dY.T @ H2
    gb3 = dY.sum(0)
    dH2 = (dY @ W3) * (Z2 > 0)
    gW2 = dH2.T @ H1
    gb2 = dH2.sum(0)
    dH1 = (dH2 @ W2) * (Z1 > 0)
    gW1 = dH1.T @ X
    gb1 = dH1.sum(0)
    loss = float(((Y - T) ** 2).mean())
    return loss, np.concatenate([g.ravel() for g in (gW1, gb1, gW2, gb2, gW3, gb3)])


def features(history, rhs, n_cell_dofs, scales, target=None):
    """Features of every cell: backward differences of orders 0..H-1 of the history (newest
    first), then the rhs, each divided by its scale; targets (x - extrapolation) / s_t.
    scales = [s_0 .. s_{H-1}, s_b, s_t]."""
    H = len(history)
    h = [np.asarray(v, dtype=np.float64).reshape(-1, n_cell_dofs) for v in history]
    diffs = [h[0]]
    if H > 1:
        diffs.append(h[0] - h[1])
    if H > 2:
        diffs.append(h[0] - 2 * h[1] + h[2])
    if H > 3:
        diffs.append(h[0] - 3 * h[1] + 3 * h[2] - h[3])
    cols = [d / scales[k] for k, d in enumerate(diffs)]
    cols.append(np.asarray(rhs, dtype=np.float64).reshape(-1, n_cell_dofs) / scales[H])
    X = np.concatenate(cols, axis=1)
    T = None
    if target is not None:
        ex = sum(b * v for b, v in zip(BETA[H], h))
        T = (np.asarray(target, dtype=np.float64).reshape(-1, n_cell_dofs) - ex) / scales[H + 1]
    return X, T


def extrapolate(history):
    return sum(b * np.asarray(v, dtype=np.float64) for b, v in zip(BETA[len(history)], history))
