"""Learned surrogate potential: permutation-sorted inverse-distance descriptor + MLP, trained and
evaluated by the library's BF16-MFMA kernels.

The reference's learned-PES convention is ``NN_Potential`` (a model handed to a user
function, simulation_utilities/potential_manager.py:177-214) with a pair-distance descriptor
(simulation_utilities/tensorflow_descriptors/distance_descriptors.py:102-152) and training
data dumped by ``DEBUG_save_training_every`` (pyvibdmc.py:800-804, 840-844).  Here the
descriptor, forward pass, backward pass and Adam all run on the GPU; ``SurrogatePotential``
plugs into ``DMC_Sim`` and is evaluated on the device straight from the walker planes,
with the analytic potential as the fallback outside the training range and as a judge on a
sample of walkers.
"""
import ctypes
import time

import numpy as np

from . import _native as N
from .potentials import DevicePotential

__all__ = ['MLPDesc', 'like_atoms', 'SurrogateModel', 'SurrogatePotential']


class MLPDesc(ctypes.Structure):
    _fields_ = [('D', ctypes.c_int32), ('H', ctypes.c_int32), ('n_atoms', ctypes.c_int32),
                ('z', ctypes.c_int32 * 6), ('sorted_atoms', ctypes.c_int32 * 6), ('sorted_atoms_len', ctypes.c_int32 * 6),
                ('n_sorted_atoms', ctypes.c_int32), ('sorted_groups', ctypes.c_int32 * 6), ('n_groups', ctypes.c_int32),
                ('group_size', ctypes.c_int32), ('use_range', ctypes.c_int32),
                ('xmean', ctypes.c_float * 32), ('xinv', ctypes.c_float * 32), ('lo', ctypes.c_float * 32),
                ('hi', ctypes.c_float * 32), ('ymean', ctypes.c_float), ('ystd', ctypes.c_float),
                ('quad', ctypes.c_float * 136), ('params', ctypes.c_void_p), ('packs', ctypes.c_void_p)]


_P = ctypes.c_void_p
_MSIGS = {
    'pvd_mlp_last_error': (ctypes.c_char_p, []),
    'pvd_mlp_param_count': (ctypes.c_int64, [ctypes.c_int32, ctypes.c_int32]),
    'pvd_mlp_pack_bytes': (ctypes.c_int64, [ctypes.c_int32]),
    'pvd_mlp_workspace_bytes': (ctypes.c_int64, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int64]),
    'pvd_mlp_pack': (ctypes.c_int, [ctypes.POINTER(MLPDesc), _P]),
    'pvd_mlp_descriptors': (ctypes.c_int, [ctypes.POINTER(MLPDesc), ctypes.c_int32, ctypes.c_int32, _P,
                                           ctypes.c_int64, ctypes.c_int32, _P, _P]),
    'pvd_mlp_energy': (ctypes.c_int, [ctypes.POINTER(MLPDesc), ctypes.c_int32, ctypes.c_int32, _P, ctypes.c_int64,
                                      _P, ctypes.c_int64, _P, ctypes.POINTER(N.PesDesc), _P]),
    'pvd_mlp_grads': (ctypes.c_int, [ctypes.POINTER(MLPDesc), _P, _P, ctypes.c_int64, _P, _P, _P, _P, _P]),
    'pvd_mlp_grads_noalloc': (ctypes.c_int, [ctypes.POINTER(MLPDesc), _P, _P, ctypes.c_int64, _P, _P, _P, _P]),
    'pvd_mlp_noalloc_scratch_bytes': (ctypes.c_int64, []),
    'pvd_mlp_adam': (ctypes.c_int, [_P, _P, _P, _P, ctypes.c_int64, ctypes.c_float, ctypes.c_float,
                                    ctypes.c_float, ctypes.c_float, ctypes.c_int64, _P]),
    'pvd_sim_device_view': (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_int64),
                                           ctypes.POINTER(_P)]),
    'pvd_mlp_judge': (ctypes.c_int, [ctypes.POINTER(N.PesDesc), ctypes.c_int32, ctypes.c_int32, _P, ctypes.c_int64,
                                     _P, ctypes.c_int64, _P, ctypes.c_int32, ctypes.c_int64, _P, _P]),
    'pvd_sim_set_surrogate': (ctypes.c_int, [_P, ctypes.POINTER(MLPDesc), ctypes.POINTER(N.PesDesc), ctypes.c_int32,
                                             ctypes.c_int32]),
    'pvd_sim_judge_log': (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), _P]),
}


def _lib():
    lib = N.load()
    if not getattr(lib, '_mlp_bound', False):
        for name, (res, args) in _MSIGS.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        lib._mlp_bound = True
    return lib


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f'{what}: {_lib().pvd_mlp_last_error().decode()}')


_Z = {e: z for z, e in enumerate('H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar'.split(), start=1)}
_Z.update(D=1, T=1)


def nuclear_charges(elements):
    """Nuclear charges of element symbols (isotopes D and T are hydrogen)."""
    try:
        return [_Z[e] for e in elements]
    except KeyError as err:
        raise ValueError(f'no nuclear charge for {err.args[0]!r}: pass zs explicitly') from None


def like_atoms(elements):
    """DistIt ``sorted_atoms`` that sorts every element's atoms among themselves: one sublist
    per element, in order of first appearance (a sublist of one atom keeps it in place)."""
    out = []
    for e in dict.fromkeys(elements):
        out.append([i for i, x in enumerate(elements) if x == e])
    return out


def _sort_spec(d, n_atoms, sorted_atoms, sorted_groups):
    """Fill the descriptor's sorting tables, with DistIt's argument checks
    (distance_descriptors.py:55-63, :133-152)."""
    d.n_sorted_atoms = 0
    d.n_groups = 0
    d.group_size = 0
    if sorted_atoms is not None:
        flat = [int(a) for sub in sorted_atoms for a in sub]
        if len(flat) != n_atoms or sorted(flat) != list(range(n_atoms)):
            raise ValueError("Please put all atoms in sorted_atoms list")
        o = 0
        for k, sub in enumerate(sorted_atoms):
            d.sorted_atoms_len[k] = len(sub)
            for a in sub:
                d.sorted_atoms[o] = int(a)
                o += 1
        d.n_sorted_atoms = len(sorted_atoms)
    if sorted_groups is not None and len(sorted_groups) > 1:
        size = len(sorted_groups[0])
        flat = [int(p) for g in sorted_groups for p in g]
        if any(len(g) != size for g in sorted_groups) or len(set(flat)) != len(flat) or \
                not all(0 <= p < n_atoms for p in flat):
            raise ValueError('sorted_groups must be equal-size groups of distinct atom positions')
        for k, p in enumerate(flat):
            d.sorted_groups[k] = p
        d.n_groups, d.group_size = len(sorted_groups), size


class SurrogateModel:
    """D -> H -> H -> H -> 1 SiLU network (H = 128 or 256) on the inverse distances of all
    atom pairs of ``elements`` after DistIt's permutation sorting
    (simulation_utilities/tensorflow_descriptors/distance_descriptors.py:115-152, sort key the
    Coulomb-matrix row norm): ``sorted_atoms`` (default: like atoms sorted among themselves,
    ``like_atoms``), ``sorted_groups`` (default none) and ``zs`` (default: the elements'
    nuclear charges) have DistIt's meaning.  FP32 master
    parameters and BF16 MFMA packs on the GPU."""

    def __init__(self, elements, hidden=128, seed=0, sorted_atoms='like', sorted_groups=None, zs=None):
        import torch
        self.elements = list(elements)
        self.n_atoms = len(self.elements)
        if not 2 <= self.n_atoms <= 6:
            raise ValueError('the descriptor supports 2 to 6 atoms')
        if isinstance(sorted_atoms, str) and sorted_atoms == 'like':
            sorted_atoms = like_atoms(self.elements)
        self.sorted_atoms = None if sorted_atoms is None else [list(map(int, s)) for s in sorted_atoms]
        self.sorted_groups = None if sorted_groups is None else [list(map(int, g)) for g in sorted_groups]
        self.zs = [int(z) for z in (nuclear_charges(self.elements) if zs is None else zs)]
        if len(self.zs) != self.n_atoms or min(self.zs) < 1:
            raise ValueError('zs: one positive nuclear charge per atom')
        self.D = self.n_atoms * (self.n_atoms - 1) // 2
        self.H = hidden
        lib = _lib()
        d = MLPDesc()
        d.D, d.H, d.n_atoms = self.D, hidden, self.n_atoms
        for k, z in enumerate(self.zs):
            d.z[k] = z
        _sort_spec(d, self.n_atoms, self.sorted_atoms, self.sorted_groups)
        for k in range(32):
            d.xmean[k] = 0.0
            d.xinv[k] = 1.0
            d.lo[k] = -1e30
            d.hi[k] = 1e30
        d.ymean, d.ystd = 0.0, 1.0
        self.desc = d
        npar = lib.pvd_mlp_param_count(self.D, hidden)
        rs = np.random.RandomState(seed)
        H, D = hidden, self.D

        def lin(o, i):
            return (rs.uniform(-1, 1, (o, i)) * np.sqrt(6.0 / (i + o))).astype(np.float32)
        flat = np.concatenate([lin(H, D).ravel(), np.zeros(H, np.float32), lin(H, H).ravel(), np.zeros(H, np.float32),
                               lin(H, H).ravel(), np.zeros(H, np.float32),
                               (rs.uniform(-1, 1, H) * np.sqrt(6.0 / (H + 1))).astype(np.float32),
                               np.zeros(1, np.float32)])
        assert flat.size == npar
        self.params = torch.from_numpy(flat).cuda()
        self.m1 = torch.zeros_like(self.params)
        self.m2 = torch.zeros_like(self.params)
        self.grads = torch.zeros_like(self.params)
        self.packs = torch.empty(int(lib.pvd_mlp_pack_bytes(hidden)), dtype=torch.uint8, device='cuda')
        d.params = self.params.data_ptr()
        d.packs = self.packs.data_ptr()
        self.t = 0
        self.history = []
        self.repack()

    # -- parameters as NumPy views (for checks) --------------------------------------
    def param_dict(self):
        p = self.params.cpu().numpy()
        H, D = self.H, self.D
        o = 0
        out = {}
        for k, sz, sh in (('W0', H * D, (H, D)), ('b0', H, (H,)), ('W1', H * H, (H, H)), ('b1', H, (H,)),
                          ('W2', H * H, (H, H)), ('b2', H, (H,)), ('w3', H, (H,)), ('b3', 1, (1,))):
            out[k] = p[o:o + sz].reshape(sh).copy()
            o += sz
        return out

    def set_params(self, pd):
        import torch
        flat = np.concatenate([np.asarray(pd[k], np.float32).ravel() for k in ('W0', 'b0', 'W1', 'b1', 'W2', 'b2',
                                                                            'w3', 'b3')])
        self.params.copy_(torch.from_numpy(flat).cuda())
        self.repack()

    def repack(self):
        _check(_lib().pvd_mlp_pack(ctypes.byref(self.desc), N.stream_handle()), 'pvd_mlp_pack')

    # -- descriptors --------------------------------------------------------------------
    def features(self, coords, normalise=True):
        """Descriptors of (n, M, 3) walkers (NumPy or device tensor) as a device tensor (D, n)."""
        import torch
        x = coords if isinstance(coords, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(coords, np.float64))
        x = x.cuda().contiguous()
        prec = N.PVD_FP64 if x.dtype == torch.float64 else N.PVD_FP32
        n = x.shape[0]
        ft = torch.empty((self.D, n), dtype=torch.float32, device='cuda')
        _check(_lib().pvd_mlp_descriptors(ctypes.byref(self.desc), prec, self.n_atoms, x.data_ptr(), n,
                                          1 if normalise else 0, ft.data_ptr(), N.stream_handle()),
               'pvd_mlp_descriptors')
        return ft

    def set_normalisation(self, coords, energies, range_margin=0.0):
        import torch
        f = self.features(coords, normalise=False)
        mean = f.mean(dim=1).cpu().numpy()
        std = f.std(dim=1).cpu().numpy()
        lo = f.min(dim=1).values.cpu().numpy()
        hi = f.max(dim=1).values.cpu().numpy()
        e = np.asarray(energies, np.float64)
        d = self.desc
        for k in range(self.D):
            d.xmean[k] = float(mean[k])
            d.xinv[k] = float(1.0 / max(std[k], 1e-12))
            span = hi[k] - lo[k]
            d.lo[k] = float(lo[k] - range_margin * span)
            d.hi[k] = float(hi[k] + range_margin * span)
        d.ymean = float(e.mean())
        d.ystd = float(e.std())

    # -- training ----------------------------------------------------------------------
    def _workspace(self, n):
        import torch
        nb = int(_lib().pvd_mlp_workspace_bytes(self.D, self.H, n))
        if getattr(self, '_ws', None) is None or self._ws.numel() < nb:
            self._ws = torch.empty(nb, dtype=torch.uint8, device='cuda')
            self._lp = torch.empty((n + 255) // 256 + 1, dtype=torch.float64, device='cuda')
        return self._ws, self._lp

    def loss_and_grads(self, x_t, targets):
        """MSE on normalised features (D, n) and normalised targets (n); fills self.grads."""
        import torch
        n = targets.shape[0]
        ws, lp = self._workspace(n)
        y = torch.empty(n, dtype=torch.float32, device='cuda')
        _check(_lib().pvd_mlp_grads(ctypes.byref(self.desc), x_t.data_ptr(), targets.data_ptr(), n,
                                    self.grads.data_ptr(), y.data_ptr(), lp.data_ptr(), ws.data_ptr(),
                                    N.stream_handle()), 'pvd_mlp_grads')
        return lp[:(n + 255) // 256].sum() / n, y

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.t += 1
        _check(_lib().pvd_mlp_adam(self.params.data_ptr(), self.m1.data_ptr(), self.m2.data_ptr(),
                                   self.grads.data_ptr(), self.params.numel(), lr, b1, b2, eps, self.t,
                                   N.stream_handle()), 'pvd_mlp_adam')
        self.repack()

    def quad_design(self, xn):
        """Design matrix (n, 1 + D + D(D+1)/2) of the quadratic baseline on normalised
        features xn (D, n): constant, x_i, then x_i x_j for i <= j in row order."""
        import torch
        x = xn.double()
        cols = [torch.ones(x.shape[1], dtype=torch.float64, device=x.device)]
        cols += [x[i] for i in range(self.D)]
        cols += [x[i] * x[j] for i in range(self.D) for j in range(i, self.D)]
        return torch.stack(cols, dim=1)

    def fit_baseline(self, xn, energies):
        """Least-squares quadratic baseline (FP64 normal equations on the device); returns
        the residual energies the network is trained on."""
        import torch
        e = energies.double()
        A = self.quad_design(xn)
        G = A.t() @ A
        G += torch.eye(G.shape[0], dtype=torch.float64, device=G.device) * (1e-12 * float(G.diagonal().mean()))
        c = torch.linalg.solve(G, A.t() @ e)
        cq = c.cpu().numpy()
        for k in range(136):
            self.desc.quad[k] = float(cq[k]) if k < len(cq) else 0.0
        # residual against the baseline as the kernels evaluate it (FP32 coefficients)
        cf = torch.from_numpy(cq.astype(np.float32).astype(np.float64)).to(A.device)
        return e - A @ cf

    def fit(self, coords, energies, epochs=200, batch=8192, lr=1e-3, seed=0, verbose=False, max_seconds=None,
            baseline=True):
        """Adam on the mean squared error of the normalised energies.  Data: (N, M, 3) walkers
        and (N,) energies in hartree.  With ``baseline`` the network learns the residual of a
        least-squares quadratic in the normalised descriptor (added back in FP32 at inference),
        which shrinks the target scale and with it the BF16 rounding error of the output.
        Returns the per-epoch training loss (normalised units)."""
        import torch
        self.set_normalisation(coords, energies)
        x = self.features(coords, normalise=True)
        e = torch.from_numpy(np.asarray(energies, np.float64)).cuda()
        for k in range(136):
            self.desc.quad[k] = 0.0
        if baseline:
            e = self.fit_baseline(x, e)
            self.desc.ymean = float(e.mean())
            self.desc.ystd = float(e.std())
        t = ((e - self.desc.ymean) / self.desc.ystd).float()
        n = t.shape[0]
        g = torch.Generator(device='cuda')
        g.manual_seed(seed)
        start = time.time()
        losses = []
        for ep in range(epochs):
            perm = torch.randperm(n, device='cuda', generator=g)
            tot = torch.zeros((), dtype=torch.float64, device='cuda')
            nb = 0
            for b0 in range(0, n - batch + 1 if n >= batch else 1, batch):
                idx = perm[b0:b0 + batch]
                xb = x[:, idx].contiguous()
                tb = t[idx].contiguous()
                loss, _ = self.loss_and_grads(xb, tb)
                self.adam_step(lr)
                tot += loss
                nb += 1
            losses.append(float(tot / max(nb, 1)))
            if verbose and (ep % 10 == 0 or ep == epochs - 1):
                print(f'epoch {ep}: loss {losses[-1]:.3e}', flush=True)
            if max_seconds and time.time() - start > max_seconds:
                break
        self.history += losses
        self.calibrate_offset(coords, energies)
        self.train_seconds = time.time() - start
        return losses

    def calibrate_offset(self, coords, energies, samples=262144):
        """Zero the mean error of the deployed network (the inference kernel, BF16 rounding
        included) on the training geometries by shifting the output offset.  A DMC energy
        estimate sees the surrogate's mean error over the sampled distribution directly, and
        the training data are that distribution."""
        n = len(energies)
        step = max(1, n // samples)
        c = np.asarray(coords)[::step]
        e = np.asarray(energies, np.float64)[::step]
        keep = self.desc.use_range
        self.desc.use_range = 0
        try:
            d = float(np.mean(self.energy(c) - e))
        finally:
            self.desc.use_range = keep
        self.desc.ymean = float(self.desc.ymean - d)
        return d

    # -- inference ----------------------------------------------------------------------
    def energy_planes(self, planes_ptr, stride, n_dev_ptr, n_max, out, precision, fallback=None):
        fb = fallback.desc if fallback is not None else None
        _check(_lib().pvd_mlp_energy(ctypes.byref(self.desc), precision, self.n_atoms, planes_ptr, stride,
                                     n_dev_ptr, n_max, out.data_ptr(), ctypes.byref(fb) if fb is not None else None,
                                     N.stream_handle()), 'pvd_mlp_energy')

    def energy(self, coords, fallback=None):
        """V (hartree, float64 NumPy) of (n, M, 3) host walkers."""
        import torch
        x = torch.from_numpy(np.ascontiguousarray(coords, np.float64)).cuda()
        n = x.shape[0]
        planes = x.reshape(n, -1).t().contiguous()
        out = torch.empty(n, dtype=torch.float64, device='cuda')
        self.energy_planes(planes.data_ptr(), n, None, n, out, N.PVD_FP64, fallback)
        return out.cpu().numpy()


class SurrogatePotential(DevicePotential):
    """A trained ``SurrogateModel`` as a DMC potential.  Walkers whose descriptor leaves the
    training range use ``fallback`` (an analytic device potential); every ``judge_every``
    steps the analytic potential is evaluated on every ``judge_stride``-th walker and the
    surrogate's error there is recorded (``judge_log``: (step, rms, max |error|, samples),
    hartree).  Inside ``DMC_Sim`` the library runs move, network, judge and branching on the
    device with no host synchronisation; ``attach`` copies the network description into
    the simulation, so a model trained further must be attached again."""

    def __init__(self, model, fallback, use_range=True, judge_every=100, judge_stride=100):
        super().__init__()
        self.model = model
        self.fallback = fallback
        self.n_atoms = model.n_atoms
        self.n_dim = 3
        self.desc.kind = N.PES_EXTERNAL
        self.model.desc.use_range = 1 if use_range else 0
        self.judge_every = judge_every
        self.judge_stride = judge_stride
        self._V = None
        self._sim = None

    def attach(self, sim):
        """Register the network with the simulation's library handle."""
        _check(_lib().pvd_sim_set_surrogate(sim._h, ctypes.byref(self.model.desc), ctypes.byref(self.fallback.desc),
                                            int(self.judge_every or 0), int(self.judge_stride or 1)),
               'pvd_sim_set_surrogate')
        self._sim = sim

    @property
    def judge_log(self):
        if self._sim is None or getattr(self._sim, '_h', None) is None:
            return []
        lib = _lib()
        n = ctypes.c_int32()
        N.check(lib.pvd_sim_judge_log(self._sim._h, None, 0, ctypes.byref(n), N.stream_handle()), 'pvd_sim_judge_log')
        if n.value == 0:
            return []
        out = np.empty((n.value, 4))
        N.check(lib.pvd_sim_judge_log(self._sim._h, out.ctypes.data, n.value, ctypes.byref(n), N.stream_handle()),
                'pvd_sim_judge_log')
        return [(int(r[0]), float(np.sqrt(r[1] / max(r[3], 1.0))), float(r[2]), int(r[3])) for r in out]

    def getpot(self, cds, timeit=False):
        start = time.time() if timeit else None
        v = self.model.energy(cds, self.fallback)
        return (v, time.time() - start) if timeit else v

    def device_energy(self, sim):
        """V of the sim's current walkers into a persistent device buffer (no host sync)."""
        import torch
        lib = _lib()
        planes = ctypes.c_void_p()
        stride = ctypes.c_int64()
        ndev = ctypes.c_void_p()
        N.check(lib.pvd_sim_device_view(sim._h, ctypes.byref(planes), ctypes.byref(stride), ctypes.byref(ndev)),
                'pvd_sim_device_view')
        cap = stride.value
        if self._V is None or self._V.numel() < cap:
            self._V = torch.empty(cap, dtype=torch.float64, device='cuda')
        prec = N.PVD_FP64 if sim.precision == 'fp64' else N.PVD_FP32
        self.model.energy_planes(planes.value, cap, ndev.value, cap, self._V, prec, self.fallback)
        return self._V
