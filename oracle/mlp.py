"""Oracle (test infrastructure only): float64 numpy restatement of the
learned peak proposal's network, loss and gradients.  The reference has no
such network; this is the checker for the HIP forward/backward kernels.

  z_i = h_{i-1} W_i + b_i, h_i = GELU(z_i) = z_i Phi(z_i) (hidden layers),
  y = h_depth W_out + b_out; candidate k = y[3k:3k+3] / |y[3k:3k+3]|;
  loss = mean over rows of  min_k d_k + (lambda / K) sum_k d_k,
  d_k = 1 - c_k . t.

Parameters are one flat float vector in the order W_1, b_1, ..., W_out,
b_out with W_i stored [in][out] row-major.  Pinned by central finite
differences of the loss (tests/test_mlp.py).
"""
import math

import numpy as np
from scipy.special import erf


def shapes(n_in, width, depth, K):
    dims = [n_in] + [width] * depth + [3 * K]
    return [(dims[i], dims[i + 1]) for i in range(depth + 1)]


def unpack(p, n_in, width, depth, K):
    out, off = [], 0
    for din, dout in shapes(n_in, width, depth, K):
        W = p[off:off + din * dout].reshape(din, dout)
        off += din * dout
        b = p[off:off + dout]
        off += dout
        out.append((W, b))
    return out


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2 * math.pi)


def forward(p, x, n_in, width, depth, K):
    layers = unpack(np.asarray(p, dtype=np.float64), n_in, width, depth, K)
    h = np.asarray(x, dtype=np.float64)
    zs, hs = [], [h]
    for i, (W, b) in enumerate(layers):
        z = h @ W + b
        zs.append(z)
        h = gelu(z) if i < depth else z
        hs.append(h)
    return h, zs, hs


def loss_grad(p, x, t, n_in, width, depth, K, lam):
    """Mean loss and its gradient with respect to the flat parameters."""
    p = np.asarray(p, dtype=np.float64)
    y, zs, hs = forward(p, x, n_in, width, depth, K)
    rows = y.shape[0]
    Y = y.reshape(rows, K, 3)
    nrm = np.linalg.norm(Y, axis=2) + 1e-12
    c = Y / nrm[:, :, None]
    dot = np.einsum("rkj,rj->rk", c, t)
    d = 1.0 - dot
    best = np.argmin(d, axis=1)
    loss = d[np.arange(rows), best] + lam / K * d.sum(axis=1)
    w = np.full((rows, K), lam / K)
    w[np.arange(rows), best] += 1.0
    g = -(w[:, :, None] * (t[:, None, :] - dot[:, :, None] * c) / nrm[:, :, None]) / rows
    dz = g.reshape(rows, 3 * K)
    layers = unpack(p, n_in, width, depth, K)
    grads = [None] * len(layers)
    for i in range(len(layers) - 1, -1, -1):
        W, b = layers[i]
        grads[i] = (hs[i].T @ dz, dz.sum(axis=0))
        if i == 0:
            break
        dz = (dz @ W.T) * gelu_d(zs[i - 1])
    flat = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
    return float(loss.mean()), flat, loss


def features(alm, L):
    """Real-form coefficients at unit RMS: Re a_l0 (l = 0..L), then for
    m = 1..L, l = m..L the pairs (Re a_lm, Im a_lm) in storage order."""
    a = np.asarray(alm)
    n0 = L + 1
    # the m > 0 coefficients occupy storage offsets L+1 .. L+1+L(L+1)/2-1
    pos = a[..., n0:n0 + L * (L + 1) // 2]
    f = np.concatenate([a[..., :n0].real, np.stack([pos.real, pos.imag], axis=-1).reshape(*pos.shape[:-1], -1)], axis=-1)
    rms = np.sqrt(np.mean(f * f, axis=-1, keepdims=True))
    return f / np.where(rms > 0, rms, 1.0)
