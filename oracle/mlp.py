"""NumPy restatement of the surrogate potential: descriptor, MLP forward/backward, Adam.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  The GPU trains and evaluates the
same network with BF16 MFMA operands and FP32 accumulation; this oracle rounds the same
operands to bfloat16 (round to nearest even) and accumulates in float64, so the two agree
up to accumulation order and the rare bfloat16 rounding flip that order causes.

Descriptor (reference ``simulation_utilities/tensorflow_descriptors/distance_descriptors.py``,
``DistIt`` with method 'coulomb': ``atm_atm_dists`` :102-113, ``sort_atoms`` :115-131,
``sort_groups`` :133-152, ``run`` :177-214): like atoms are permuted by descending
Coulomb-matrix row norm within each ``sorted_atoms`` sublist, then equal-size position
groups (``sorted_groups``) by descending summed norm; the features are the inverse distances
of the permuted atoms for position pairs (0,1), (0,2), ..., (1,2), ... -- the reference's
Coulomb values divided by z_i z_j.  Pinned by vectors the reference's DistIt wrote
(``tests/golden/distit.npz``), by the exact permutation invariance its own tests assert
(pyvibdmc/tests/test_cupy_dist.py:80-107, :110-142, on their fixtures) and by closed-form
distances.

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
def inverse_distances(cds):
    """1/r of every atom pair (i < j) in row order (distance_descriptors.py:102-113), (N, P)."""
    cds = np.asarray(cds, dtype=np.float64)
    m = cds.shape[1]
    return np.stack([1.0 / np.linalg.norm(cds[:, i] - cds[:, j], axis=1)
                     for i in range(m) for j in range(i + 1, m)], axis=1)


def coulomb_row_norms(cds, zs):
    """Row norms of the Coulomb matrix z_i z_j / r_ij with diagonal 0.5 z_i^2.4
    (distance_descriptors.py:154-166), (N, M)."""
    cds = np.asarray(cds, dtype=np.float64)
    zs = np.asarray(zs, dtype=np.float64)
    m = cds.shape[1]
    c = np.zeros((cds.shape[0], m, m))
    for i in range(m):
        c[:, i, i] = 0.5 * zs[i] ** 2.4
        for j in range(m):
            if j != i:
                c[:, i, j] = zs[i] * zs[j] / np.linalg.norm(cds[:, i] - cds[:, j], axis=1)
    return np.sqrt(np.sum(c * c, axis=2))


def atom_order(norms, sorted_atoms=None, sorted_groups=None):
    """Per geometry, the atom placed at each position: like atoms by descending norm inside
    each sublist (position sublist[r] gets the r-th largest, ties to the lower index), then
    position groups by descending summed norm (ties to the lower group), (N, M) int."""
    n, m = norms.shape
    perm = np.tile(np.arange(m), (n, 1))
    if sorted_atoms is not None:
        for sub in sorted_atoms:
            sub = list(sub)
            if len(sub) < 2:
                continue
            # stable sort by -norm: equal norms keep ascending atom order
            order = np.argsort(-norms[:, sub], axis=1, kind='stable')
            perm[:, sub] = np.asarray(sub)[order]
    if sorted_groups is not None and len(sorted_groups) > 1:
        g = np.asarray(sorted_groups)
        pn = np.take_along_axis(norms, perm, axis=1)          # norm of the atom at each position
        tot = np.stack([pn[:, grp].sum(axis=1) for grp in g], axis=1)
        rank_order = np.argsort(-tot, axis=1, kind='stable')  # k-th largest group
        new = perm.copy()
        for k in range(len(g)):
            src = g[rank_order[:, k]]                          # (N, group size) positions
            new[:, g[k]] = np.take_along_axis(perm, src, axis=1)
        perm = new
    return perm


def descriptor(cds, zs, sorted_atoms=None, sorted_groups=None):
    """Inverse distances of the permuted atoms, position pairs in row order, (N, P)."""
    cds = np.asarray(cds, dtype=np.float64)
    perm = atom_order(coulomb_row_norms(cds, zs), sorted_atoms, sorted_groups)
    return inverse_distances(np.take_along_axis(cds, perm[:, :, None], axis=1))


def like_atoms(elements):
    """One sublist per element, in order of first appearance."""
    return [[i for i, x in enumerate(elements) if x == e] for e in dict.fromkeys(elements)]


ZS = {'H': 1, 'D': 1, 'O': 8}


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
