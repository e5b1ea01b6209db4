"""ORACLE - test infrastructure only.

FP64 numpy restatement of the warm-start network: in -> hidden (GELU) ->
hidden (GELU) -> out, mean-squared-error loss, exact (erf) GELU, parameters
packed as W1[hidden x in], b1, W2[hidden x hidden], b2, W3[out x hidden], b3.
It is the checker of the HIP forward / backward kernels.  The network has no
counterpart in the reference (which has no learned component); this oracle is
pinned by central finite differences in tests/test_mlp_oracle.py.
"""
import numpy as np
from scipy.special import erf


def unpack(p, n_in, hid, n_out):
    o = 0
    W1 = p[o:o + hid * n_in].reshape(hid, n_in); o += hid * n_in
    b1 = p[o:o + hid]; o += hid
    W2 = p[o:o + hid * hid].reshape(hid, hid); o += hid * hid
    b2 = p[o:o + hid]; o += hid
    W3 = p[o:o + n_out * hid].reshape(n_out, hid); o += n_out * hid
    b3 = p[o:o + n_out]; o += n_out
    assert o == p.size
    return W1, b1, W2, b2, W3, b3


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(p, X, n_in, hid, n_out):
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, hid, n_out)
    Z1 = X @ W1.T + b1
    H1 = gelu(Z1)
    Z2 = H1 @ W2.T + b2
    H2 = gelu(Z2)
    return H2 @ W3.T + b3, (Z1, H1, Z2, H2)


def loss_grad(p, X, T, n_in, hid, n_out):
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, hid, n_out)
    Y, (Z1, H1, Z2, H2) = forward(p, X, n_in, hid, n_out)
    D = Y - T
    loss = float(np.mean(D * D))
    dY = 2.0 * D / D.size
    gW3 = dY.T @ H2
    gb3 = dY.sum(0)
    dZ2 = (dY @ W3) * gelu_grad(Z2)
    gW2 = dZ2.T @ H1
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2) * gelu_grad(Z1)
    gW1 = dZ1.T @ X
    gb1 = dZ1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g
