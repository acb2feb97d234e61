"""NumPy restatement of the surrogate potential: descriptor, MLP forward/backward, Adam.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  The GPU trains and evaluates the
same network with BF16 MFMA operands and FP32 accumulation; this oracle rounds the same
operands to bfloat16 (round to nearest even) and accumulates in float64, so the two agree
up to accumulation order and the rare bfloat16 rounding flip that order causes.

Descriptor (reference ``simulation_utilities/tensorflow_descriptors/distance_descriptors.py``,
``atm_atm_dists`` :102-113 and the like-atom sorting of :115-152): inverse interatomic
distances over a list of atom pairs, sorted in descending order within each group of pairs
of the same element pair.  Pinned by permutation invariance (swapping like atoms leaves the
descriptor unchanged, the property the reference tests at tests/test_cupy_dist.py:107,142)
and by closed-form distances.

Network: x_n = (desc - mean) * inv_std; h0 = silu(W0 x_n + b0); h1 = silu(W1 h0 + b1);
h2 = silu(W2 h1 + b2); y = w3 . h2 + b3; V = y * y_std + y_mean + q(x_n), q a quadratic
baseline (``quad_terms``; all-zero coefficients for a plain network).  MFMA operands are
bf16(W_l) and bf16(input of layer l); the output layer runs in FP32 on unrounded h2.
Pinned by central differences of the loss (``tests/test_oracle_mlp.py``); the baseline's
term order and least-squares fit by an exactly representable quadratic
(``tests/test_oracle_mlp.py``).
"""
import numpy as np


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), a, out).reshape(np.shape(x))


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


# ---- descriptors ----------------------------------------------------------------------
def pair_groups(elements):
    """All atom pairs (i < j) grouped by unordered element pair, group order of first
    appearance; returns (pairs, group_sizes) with pairs listed group by group."""
    pairs = [(i, j) for i in range(len(elements)) for j in range(i + 1, len(elements))]
    keys = []
    for i, j in pairs:
        k = tuple(sorted((elements[i], elements[j])))
        if k not in keys:
            keys.append(k)
    out, sizes = [], []
    for k in keys:
        g = [(i, j) for (i, j) in pairs if tuple(sorted((elements[i], elements[j]))) == k]
        out += g
        sizes.append(len(g))
    return out, sizes


def descriptor(cds, pairs, sizes):
    """Inverse distances of `pairs`, sorted descending within each group."""
    cds = np.asarray(cds, dtype=np.float64)
    inv = np.stack([1.0 / np.linalg.norm(cds[:, i] - cds[:, j], axis=1) for i, j in pairs], axis=1)
    out = np.empty_like(inv)
    o = 0
    for s in sizes:
        out[:, o:o + s] = -np.sort(-inv[:, o:o + s], axis=1)
        o += s
    return out


# ---- quadratic baseline ---------------------------------------------------------------
def quad_terms(xn):
    """Terms of the quadratic baseline on normalised features xn (B, D): 1, x_i, then
    x_i x_j for i <= j in row order (the order of pvd_mlp.quad)."""
    xn = np.asarray(xn, np.float64)
    B, D = xn.shape
    cols = [np.ones(B)] + [xn[:, i] for i in range(D)]
    cols += [xn[:, i] * xn[:, j] for i in range(D) for j in range(i, D)]
    return np.stack(cols, axis=1)


def fit_quad(xn, v):
    """Least-squares coefficients of the quadratic baseline (FP64)."""
    c, *_ = np.linalg.lstsq(quad_terms(xn), np.asarray(v, np.float64), rcond=None)
    return c


# ---- network --------------------------------------------------------------------------
class MLP:
    """Parameters as float32 arrays: W0 (H, D), b0, W1 (H, H), b1, W2 (H, H), b2, w3 (H,), b3."""

    def __init__(self, D, H, seed=0):
        rs = np.random.RandomState(seed)
        self.D, self.H = D, H

        def lin(o, i):
            return (rs.uniform(-1, 1, (o, i)) * np.sqrt(6.0 / (i + o))).astype(np.float32)
        self.p = dict(W0=lin(H, D), b0=np.zeros(H, np.float32), W1=lin(H, H), b1=np.zeros(H, np.float32),
                      W2=lin(H, H), b2=np.zeros(H, np.float32),
                      w3=(rs.uniform(-1, 1, H) * np.sqrt(6.0 / (H + 1))).astype(np.float32),
                      b3=np.zeros(1, np.float32))

    def forward(self, xn, rounding=True):
        """xn: normalised descriptors (B, D) float32.  Returns y (B,) and the cache.
        ``rounding=False`` drops the bfloat16 rounding (used to check the backward pass by
        central differences)."""
        p = self.p
        r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
        a = np.asarray(r(xn), np.float64)
        zs = []
        for l in range(3):
            W = np.asarray(r(p[f'W{l}']), np.float64)
            z = a @ W.T + p[f'b{l}'].astype(np.float64)
            zs.append(z)
            h = silu(z)
            a = np.asarray(r(h.astype(np.float32) if rounding else h), np.float64) if l < 2 else h
        y = a @ p['w3'].astype(np.float64) + float(p['b3'][0])
        return y, (xn, zs)

    def loss_grads(self, xn, t, rounding=True):
        """MSE mean((y - t)^2) and its gradients, the same rounding points as the GPU."""
        p = self.p
        rr = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
        y, (xn, zs) = self.forward(xn, rounding)
        B = len(t)
        r = y - t
        loss = float(np.mean(r * r))
        dy = 2.0 * r / B
        h2 = silu(zs[2])
        g = dict(w3=h2.T @ dy, b3=np.array([dy.sum()]))
        dh = np.outer(dy, p['w3'].astype(np.float64))
        for l in (2, 1, 0):
            dz = dh * dsilu(zs[l])
            if l == 0:
                hin = np.asarray(rr(xn), np.float64)
            else:
                hl = silu(zs[l - 1])
                hin = np.asarray(rr(hl.astype(np.float32) if rounding else hl), np.float64)
            dzr = np.asarray(rr(dz.astype(np.float32) if rounding else dz), np.float64)
            g[f'W{l}'] = dzr.T @ hin
            g[f'b{l}'] = dz.sum(axis=0)
            if l > 0:
                W = np.asarray(rr(p[f'W{l}']), np.float64)
                dh = dzr @ W
        return loss, y, g


class Adam:
    """Adam (Kingma & Ba 2015) with bias correction, on float32 parameters."""

    def __init__(self, params, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps
        self.m = {k: np.zeros_like(v, dtype=np.float32) for k, v in params.items()}
        self.v = {k: np.zeros_like(v, dtype=np.float32) for k, v in params.items()}
        self.t = 0

    def step(self, params, grads):
        self.t += 1
        c1 = 1.0 - self.b1 ** self.t
        c2 = 1.0 - self.b2 ** self.t
        for k in params:
            g = grads[k].astype(np.float32)
            self.m[k] = self.b1 * self.m[k] + (1 - self.b1) * g
            self.v[k] = self.b2 * self.v[k] + (1 - self.b2) * g * g
            mh = self.m[k] / c1
            vh = self.v[k] / c2
            params[k] = (params[k] - self.lr * mh / (np.sqrt(vh) + self.eps)).astype(np.float32)
