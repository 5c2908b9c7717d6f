"""TEST INFRASTRUCTURE ONLY: CPU restatement (NumPy, float64) of the per-front precision
classifier's forward pass and loss gradient, the checker for the GPU kernels (mlp.hip).

The reference has no such network; this follows the layout and loss the library documents for
rv_mlp_*: parameters W1[16][64], b1[64], W2[64][64], b2[64], W3[64][3], b3[3] in one array,
ReLU hidden layers, softmax output, class-weighted mean cross-entropy
L = sum_n w[y_n] (-log p[n, y_n]) / sum_n w[y_n].  Its gradient is pinned by central
differences in tests/test_mlp_oracle.py.
"""
import numpy as np

NI, NH, NO = 16, 64, 3
SHAPES = [("W1", (NI, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH, NO)), ("b3", (NO,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def forward(params, x):
    P = unpack(params)
    x = np.asarray(x, dtype=np.float64)
    a1 = x @ P["W1"] + P["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ P["W2"] + P["b2"]
    h2 = np.maximum(a2, 0.0)
    z = h2 @ P["W3"] + P["b3"]
    z = z - z.max(axis=1, keepdims=True)
    e = np.exp(z)
    prob = e / e.sum(axis=1, keepdims=True)
    return prob, (x, h1, h2)


def loss_grad(params, x, y, class_weight=None):
    P = unpack(params)
    y = np.asarray(y)
    cw = np.ones(NO) if class_weight is None else np.asarray(class_weight, dtype=np.float64)
    prob, (x, h1, h2) = forward(params, x)
    w = cw[y] / cw[y].sum()
    n = len(y)
    loss = float(np.sum(-w * np.log(np.maximum(prob[np.arange(n), y], 1e-300))))
    d3 = prob.copy()
    d3[np.arange(n), y] -= 1.0
    d3 *= w[:, None]
    d2 = (d3 @ P["W3"].T) * (h2 > 0)
    d1 = (d2 @ P["W2"].T) * (h1 > 0)
    g = {"W1": x.T @ d1, "b1": d1.sum(0), "W2": h1.T @ d2, "b2": d2.sum(0), "W3": h2.T @ d3, "b3": d3.sum(0)}
    return loss, np.concatenate([g[name].ravel() for name, _ in SHAPES])
