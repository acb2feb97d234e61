"""NumPy FP64 potentials, trial wave functions and internal-coordinate derivatives.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Array convention everywhere: Cartesian walkers ``cds`` of shape (N, M, d), C order,
float64, as the reference passes them to a potential (``pyvibdmc/pyvibdmc.py:761,793``;
``simulation_utilities/potential_manager.py:71-99``).  Potentials return (N,) hartree.

Pinning (checked in ``tests/test_oracle_pes.py``):
* ``harmonic_x`` / ``ho_reference_style``: closed form; reference sample PES
  ``sample_potentials/PythonPots/harmonicOscillator1D.py:65-69``.
* ``morse_1d``: ``sample_potentials/PythonPots/morse_osc_1d.py:4-12``; pinned by
  V(0)=0, V(inf)=De and the harmonic limit.
* ``morse_bend`` / ``cluster6``: defined by the benchmark configurations (not by
  the reference); pinned by equilibrium value 0, separability checks and central
  finite differences of the analytic gradients.
* ``bond`` / ``angle`` derivatives and ``GaussProductTrial``: pinned by FP64
  central differences (h = 1e-4 bohr) in the tests.
* ``HarmonicTrial1D``: restates ``sample_potentials/PythonPots/harm_trial_wfn.py:6-40``;
  pinned by the closed forms psi'/psi = -alpha x, psi''/psi = alpha^2 x^2 - alpha.
"""
import numpy as np

# --- constants restated from simulation_utilities/Constants.py:3-4,65-69 ---------
WAVENUMBERS = 4.556335281212229e-6          # hartree per cm^-1   (Constants.py:66)
ANGSTROMS = 1 / 0.529177                    # bohr per angstrom    (Constants.py:67)
AMU = 1.000000000000000000 / 6.02213670000e23 / 9.10938970000e-28   # (Constants.py:68)
MASS_H_AMU = 1.00782503                      # Constants.py:3
MASS_O_AMU = 15.99491462                     # Constants.py:4


def reduced_mass_au(m1_amu, m2_amu):
    """Constants.reduced_mass (Constants.py:101-119): convert each mass, then m1 m2/(m1+m2)."""
    a = m1_amu * AMU
    b = m2_amu * AMU
    return a * b / (a + b)


# --- configuration constants (benchmark definitions) ------------------------------
M_H = 1837.15
M_O = 29156.95
HO_OMEGA = 0.013669
MB_DE, MB_A, MB_RE = 0.1916, 1.2, 1.81
MB_KB, MB_THETA0 = 0.16, np.deg2rad(104.5)
OO_DE, OO_A, OO_RE = 0.008, 0.8, 5.5
Q_O, Q_H = -0.8, 0.4


# --- potentials ---------------------------------------------------------------------
def harmonic_x(cds, k):
    """V = 0.5*k*x^2 on x = cds[:,0,0] (1-D oscillator on the first coordinate).

    Evaluation order (0.5*k) * (x*x), which is how NumPy evaluates ``0.5*k*x**2``."""
    x = cds[:, 0, 0]
    return (0.5 * k) * (x * x)


def ho_reference_style(cds, mass, omega):
    """``np.squeeze(0.5 * mass * omega ** 2 * cds ** 2)`` (harmonicOscillator1D.py:65-69).
    For (N,1,1) walkers."""
    return np.squeeze(0.5 * mass * omega ** 2 * cds ** 2)


def morse_1d(cds, De, alpha):
    """``De * (1 - exp(-alpha*x))**2`` (morse_osc_1d.py:4-12), x = cds[:,0,0]."""
    x = cds[:, 0, 0]
    t = 1 - np.exp(-alpha * x)
    return De * (t * t)


def _norm3(v):
    return np.sqrt(v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1] + v[..., 2] * v[..., 2])


def morse_bend(cds, h1=0, h2=1, o=2, De=MB_DE, a=MB_A, re=MB_RE, kb=MB_KB, theta0=MB_THETA0):
    """Triatomic Morse + harmonic bend.

    V = De*(1-exp(-a*(r1-re)))^2 + De*(1-exp(-a*(r2-re)))^2 + 0.5*kb*(theta-theta0)^2
    with r1=|H1-O|, r2=|H2-O|, theta = angle H1-O-H2 in radians.
    cos(theta) = (u.v)/(r1*r2) clipped to [-1, 1]."""
    u = cds[:, h1] - cds[:, o]
    v = cds[:, h2] - cds[:, o]
    r1 = _norm3(u)
    r2 = _norm3(v)
    dot = u[:, 0] * v[:, 0] + u[:, 1] * v[:, 1] + u[:, 2] * v[:, 2]
    c = np.clip(dot / (r1 * r2), -1.0, 1.0)
    th = np.arccos(c)
    m1 = 1.0 - np.exp(-a * (r1 - re))
    m2 = 1.0 - np.exp(-a * (r2 - re))
    dth = th - theta0
    return De * (m1 * m1) + De * (m2 * m2) + (0.5 * kb) * (dth * dth)


def cluster6(cds):
    """Two Morse+bend monomers [O,H,H,O,H,H] + O-O Morse + inter-monomer Coulomb.

    V = MB(A) + MB(B) + De_OO*(1-exp(-a_OO*(r_OO-re_OO)))^2 + sum_{i in A, j in B} q_i q_j / r_ij
    Monomer A = atoms (O0,H1,H2), B = (O3,H4,H5); q_O=-0.8, q_H=+0.4."""
    va = morse_bend(cds, h1=1, h2=2, o=0)
    vb = morse_bend(cds, h1=4, h2=5, o=3)
    roo = _norm3(cds[:, 0] - cds[:, 3])
    m = 1.0 - np.exp(-OO_A * (roo - OO_RE))
    v = va + vb + OO_DE * (m * m)
    q = np.array([Q_O, Q_H, Q_H, Q_O, Q_H, Q_H])
    for i in range(3):
        for j in range(3, 6):
            v = v + (q[i] * q[j]) / _norm3(cds[:, i] - cds[:, j])
    return v


# --- internal coordinates with Cartesian first and diagonal second derivatives -----
def bond(cds, i, j):
    """r = |x_i - x_j|.  Returns r (N,), dr (N,M,3), d2r (N,M,3) (diagonal second
    derivatives d^2 r / dx_{a,c}^2)."""
    n, m, _ = cds.shape
    u = cds[:, i] - cds[:, j]
    r = _norm3(u)
    e = u / r[:, None]
    dr = np.zeros((n, m, 3))
    d2 = np.zeros((n, m, 3))
    dr[:, i] = e
    dr[:, j] = -e
    h = (1.0 - e * e) / r[:, None]
    d2[:, i] = h
    d2[:, j] = h
    return r, dr, d2


def angle(cds, i, j, k):
    """theta = angle i-j-k (vertex j).  Returns theta, dtheta (N,M,3), d2theta (N,M,3)
    (diagonal second derivatives), from c = cos(theta):
      dtheta = -dc/s,  d2theta = -d2c/s - c*dc^2/s^3,  s = sqrt(1-c^2)."""
    n, m, _ = cds.shape
    u = cds[:, i] - cds[:, j]
    v = cds[:, k] - cds[:, j]
    ru = _norm3(u)
    rv = _norm3(v)
    uh = u / ru[:, None]
    vh = v / rv[:, None]
    c = np.clip(np.sum(uh * vh, axis=1), -1.0, 1.0)
    s = np.sqrt(1.0 - c * c)
    th = np.arccos(c)
    C = c[:, None]
    dcu = (vh - C * uh) / ru[:, None]                     # dc/du
    dcv = (uh - C * vh) / rv[:, None]                     # dc/dv
    d2uu = (-2.0 * uh * vh - C + 3.0 * C * uh * uh) / (ru * ru)[:, None]
    d2vv = (-2.0 * uh * vh - C + 3.0 * C * vh * vh) / (rv * rv)[:, None]
    d2uv = (1.0 - uh * uh - vh * vh + C * uh * vh) / (ru * rv)[:, None]
    dc = np.zeros((n, m, 3))
    d2c = np.zeros((n, m, 3))
    dc[:, i] = dcu
    dc[:, k] = dcv
    dc[:, j] = -(dcu + dcv)
    d2c[:, i] = d2uu
    d2c[:, k] = d2vv
    d2c[:, j] = d2uu + 2.0 * d2uv + d2vv
    S = s[:, None, None]
    dth = -dc / S
    d2th = -d2c / S - c[:, None, None] * dc * dc / (S * S * S)
    return th, dth, d2th


def morse_bend_grad(cds, h1=0, h2=1, o=2, De=MB_DE, a=MB_A, re=MB_RE, kb=MB_KB, theta0=MB_THETA0):
    """dV/dx (N,M,3) of ``morse_bend``."""
    r1, dr1, _ = bond(cds, h1, o)
    r2, dr2, _ = bond(cds, h2, o)
    th, dth, _ = angle(cds, h1, o, h2)
    e1 = np.exp(-a * (r1 - re))
    e2 = np.exp(-a * (r2 - re))
    g1 = 2.0 * De * (1.0 - e1) * a * e1
    g2 = 2.0 * De * (1.0 - e2) * a * e2
    gt = kb * (th - theta0)
    return g1[:, None, None] * dr1 + g2[:, None, None] * dr2 + gt[:, None, None] * dth


def cluster6_grad(cds):
    """dV/dx (N,6,3) of ``cluster6``."""
    g = morse_bend_grad(cds, 1, 2, 0) + morse_bend_grad(cds, 4, 5, 3)
    roo, droo, _ = bond(cds, 0, 3)
    e = np.exp(-OO_A * (roo - OO_RE))
    g = g + (2.0 * OO_DE * (1.0 - e) * OO_A * e)[:, None, None] * droo
    q = np.array([Q_O, Q_H, Q_H, Q_O, Q_H, Q_H])
    for i in range(3):
        for j in range(3, 6):
            r, dr, _ = bond(cds, i, j)
            g = g + (-(q[i] * q[j]) / (r * r))[:, None, None] * dr
    return g


# --- trial wave functions -------------------------------------------------------------
class HarmonicTrial1D:
    """harm_trial_wfn.py:6-40: psi = (alpha/pi)^0.25 exp(-alpha x^2/2), alpha = mu*omega,
    mu = reduced_mass('O-H'), omega = 3700 cm^-1.  ``derivs`` returns (psi'/psi, psi''/psi)
    formed as in ``derivative`` (harm_trial_wfn.py:36-40): each derivative divided by psi."""

    def __init__(self, omega_cm=3700.0):
        self.mass = reduced_mass_au(MASS_O_AMU, MASS_H_AMU)
        self.omega = omega_cm * WAVENUMBERS
        self.alpha = self.mass * self.omega

    def trial(self, x):
        a = self.alpha
        return ((a / np.pi) ** 0.25 * np.exp(-a * x ** 2 / 2)).squeeze()

    def derivs(self, x):
        a = self.alpha
        trl = self.trial(x)
        d1 = (a / np.pi) ** 0.25 * (-a * x) * np.exp(-a * x ** 2 / 2)
        d2 = (a / np.pi) ** 0.25 * (a ** 2 * x ** 2 - a) * np.exp(-a * x ** 2 / 2)
        return d1 / trl[:, None, None], d2 / trl[:, None, None]


class GaussProductTrial:
    """psi_T = prod_k (alpha_k/pi)^(1/4) exp(-alpha_k (q_k - q0_k)^2 / 2) over bonds and angles.

    ``bonds``  : list of (i, j, alpha, r0)
    ``angles`` : list of (i, j, k, alpha, theta0), vertex j.
    ``derivs`` returns (dpsi/psi, d2psi/psi), each (N,M,3), the per-Cartesian diagonal
    second derivative: d2psi/psi = (d ln psi)^2 + d2 ln psi.  Pattern of the reference's
    bend factor: ``call_trl_h2o.py:38-60``."""

    def __init__(self, bonds, angles):
        self.bonds = list(bonds)
        self.angles = list(angles)

    def _coords(self, cds):
        out = []
        for (i, j, a, q0) in self.bonds:
            q, dq, d2q = bond(cds, i, j)
            out.append((a, q0, q, dq, d2q))
        for (i, j, k, a, q0) in self.angles:
            q, dq, d2q = angle(cds, i, j, k)
            out.append((a, q0, q, dq, d2q))
        return out

    def trial(self, cds):
        psi = np.ones(len(cds))
        for (a, q0, q, _, _) in self._coords(cds):
            dq = q - q0
            psi = psi * ((a / np.pi) ** 0.25 * np.exp(-a * dq * dq / 2))
        return psi

    def log_trial(self, cds):
        s = np.zeros(len(cds))
        for (a, q0, q, _, _) in self._coords(cds):
            dq = q - q0
            s = s + (0.25 * np.log(a / np.pi) - a * dq * dq / 2)
        return s

    def derivs(self, cds):
        g = np.zeros(cds.shape)
        h = np.zeros(cds.shape)
        for (a, q0, q, dq, d2q) in self._coords(cds):
            x = (q - q0)[:, None, None]
            g += -a * x * dq
            h += -a * (dq * dq + x * d2q)
        return g, g * g + h

    def weighted_laplacian(self, cds, masses):
        """sum_a (1/m_a) lap_a ln psi in closed form (the FP32 device kernels' formula):
        bond r: |grad r|^2 = 1, lap r = 2/r at both atoms; angle theta with vertex j and arms
        u, v: |grad_i|^2 = 1/ru^2, |grad_k|^2 = 1/rv^2, |grad_j|^2 = 1/ru^2 + 1/rv^2 -
        2 cos/(ru rv); lap_i = cot/ru^2, lap_k = cot/rv^2, lap_j = cot (1/ru^2 + 1/rv^2) -
        2/(ru rv sin).  Pinned against ``derivs`` (which central differences pin):
        sum_p (1/m_p) d2psi/psi = sum_p (1/m_p) (d ln psi)^2 + weighted_laplacian."""
        im = 1.0 / np.asarray(masses, np.float64)
        lap = np.zeros(len(cds))
        for (i, j, al, r0) in self.bonds:
            u = cds[:, i] - cds[:, j]
            r = np.linalg.norm(u, axis=1)
            lap += -al * (1 + (r - r0) * 2 / r) * (im[i] + im[j])
        for (i, j, k, al, t0) in self.angles:
            u = cds[:, i] - cds[:, j]
            v = cds[:, k] - cds[:, j]
            ru = np.linalg.norm(u, axis=1)
            rv = np.linalg.norm(v, axis=1)
            c = np.clip((u * v).sum(1) / (ru * rv), -1, 1)
            s = np.sqrt(1 - c * c)
            cot = c / s
            dq = np.arccos(c) - t0
            iru2, irv2 = 1 / ru ** 2, 1 / rv ** 2
            lap += (-al * (iru2 + dq * cot * iru2)) * im[i] + (-al * (irv2 + dq * cot * irv2)) * im[k]
            lap += (-al * ((iru2 + irv2 - 2 * c / (ru * rv)) + dq * (cot * (iru2 + irv2) - 2 / (ru * rv * s)))) * im[j]
        return lap


def water_trial_params(m_h=M_H, m_o=M_O, De=MB_DE, a=MB_A, re=MB_RE, kb=MB_KB, theta0=MB_THETA0):
    """Harmonic Gaussian widths of the Morse+bend PES.

    alpha_r = mu_OH * omega_r, omega_r = a*sqrt(2 De/mu_OH);
    alpha_theta = omega_theta / G_thth, omega_theta = sqrt(kb*G_thth),
    G_thth as ``call_trl_h2o.py:38-40`` (with the configuration's masses)."""
    mu = m_h * m_o / (m_h + m_o)
    w_r = a * np.sqrt(2.0 * De / mu)
    alpha_r = mu * w_r
    g = 1 / m_h / re ** 2 + 1 / m_h / re ** 2 + 1 / m_o * (1 / re ** 2 + 1 / re ** 2 - 2 * np.cos(theta0) / (re * re))
    w_t = np.sqrt(kb * g)
    alpha_t = w_t / g
    return dict(alpha_r=alpha_r, alpha_t=alpha_t, omega_r=w_r, omega_t=w_t, g_tt=g, mu=mu)


def water_trial(h1=0, h2=1, o=2, **kw):
    p = water_trial_params(**kw)
    return GaussProductTrial(bonds=[(h1, o, p['alpha_r'], MB_RE), (h2, o, p['alpha_r'], MB_RE)],
                             angles=[(h1, o, h2, p['alpha_t'], MB_THETA0)])


def cluster6_trial():
    p = water_trial_params()
    ar, at = p['alpha_r'], p['alpha_t']
    return GaussProductTrial(bonds=[(1, 0, ar, MB_RE), (2, 0, ar, MB_RE), (4, 3, ar, MB_RE), (5, 3, ar, MB_RE)],
                             angles=[(1, 0, 2, at, MB_THETA0), (4, 3, 5, at, MB_THETA0)])


# --- geometries -------------------------------------------------------------------------
def water_eq():
    """Atoms [H, H, O]: O at origin, H1 on x at re, H2 at re*(cos 104.5, sin 104.5, 0)."""
    t = MB_THETA0
    return np.array([[MB_RE, 0.0, 0.0], [MB_RE * np.cos(t), MB_RE * np.sin(t), 0.0], [0.0, 0.0, 0.0]])


def cluster6_eq():
    """Atoms [O,H,H,O,H,H]; monomer B = monomer A translated 5.5 bohr along z."""
    w = water_eq()
    a = np.array([w[2], w[0], w[1]])
    b = a + np.array([0.0, 0.0, OO_RE])
    return np.concatenate([a, b], axis=0)


def finite_diff(cds, trial_func, dx=0.001):
    """3-point FD of psi (simulation_utilities/imp_samp.py:55-76): returns (first, sec, psi),
    NOT divided by psi (the managers divide, imp_samp_manager.py:119-120, 206-208).

    Like the reference, it perturbs ``cds`` IN PLACE (x - dx + 2dx - dx), so the caller's
    walkers keep the rounding of that round trip (imp_samp.py:67-71)."""
    num_atoms = cds.shape[1]
    num_dimz = cds.shape[-1]
    first = np.zeros(cds.shape)
    sec = np.zeros(cds.shape)
    tmp = np.zeros((len(cds), 3))
    tmp[:, 1] = trial_func(cds)
    for a in range(num_atoms):
        for d in range(num_dimz):
            cds[:, a, d] -= dx
            tmp[:, 0] = trial_func(cds)
            cds[:, a, d] += 2.0 * dx
            tmp[:, 2] = trial_func(cds)
            cds[:, a, d] -= dx
            first[:, a, d] = (tmp[:, 2] - tmp[:, 0]) / (2 * dx)
            sec[:, a, d] = ((tmp[:, 0] - 2.0 * tmp[:, 1] + tmp[:, 2]) / dx ** 2)
    return first, sec, tmp[:, 1]


# --- Partridge-Schwenke water surface ------------------------------------------------------
def _ps_table():
    """The surface's coefficient table, read from the header the device kernels compile
    (pyvibdmc_amd/csrc/pvd_ps_table.h, written by gen_ps_table.py from the reference's data)."""
    import os
    import re
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'pyvibdmc_amd', 'csrc',
                        'pvd_ps_table.h')
    s = open(path).read()
    par = {k: float(v) for k, v in re.findall(r'#define PVD_PS_([A-Z0-9]+) ([-0-9.e+]+)\n', s)}

    def block(name):
        m = re.search(r'#define ' + name + r' \\\n(.*?)(?:\n//|\Z)', s, re.S)
        return [t for t in re.split(r'[,\s\\]+', m.group(1)) if t]
    c = np.array([float(t) for t in block('PVD_PS_COEFFS')])
    e = np.array([int(t) for t in block('PVD_PS_EXPONENTS')]).reshape(-1, 3)
    return par, c, e


_PS = None


def ps_water(cds):
    """Partridge-Schwenke water surface (hartree) of [H, H, O] walkers (bohr), in the
    operation order of the reference's Fortran (calc_h2o_pot.f:12-27 for the internal
    coordinates, h2opes_v2.f:460-489 for the surface): r1, r2 the O-H distances, theta from
    acos(cos), x1 = (r1 - re)/re, x2 = (r2 - re)/re, x3 = cos(theta) - cos(theta_e);
    V = [sum_j c_j (x1^a x2^b + x2^a x1^b) x3^k] exp(-b1 ((r1-re)^2 + (r2-re)^2)) + c_0
        + V_OH(r1) + V_OH(r2) + V_HH(r_HH)."""
    global _PS
    if _PS is None:
        _PS = _ps_table()
    par, c, e = _PS
    h1, h2, o = cds[:, 0], cds[:, 1], cds[:, 2]
    d1 = o - h1
    d2 = o - h2
    r1s = ((0.0 + d1[:, 0] * d1[:, 0]) + d1[:, 1] * d1[:, 1]) + d1[:, 2] * d1[:, 2]
    r2s = ((0.0 + d2[:, 0] * d2[:, 0]) + d2[:, 1] * d2[:, 1]) + d2[:, 2] * d2[:, 2]
    ct = ((0.0 + d1[:, 0] * d2[:, 0]) + d1[:, 1] * d2[:, 1]) + d1[:, 2] * d2[:, 2]
    r1 = np.sqrt(r1s)
    r2 = np.sqrt(r2s)
    th = np.arccos(ct / r1 / r2)
    reoh = par['REOH']
    x1 = (r1 - reoh) / reoh
    x2 = (r2 - reoh) / reoh
    cth = np.cos(th)
    x3 = cth - par['CE']
    rhh = np.sqrt((r1 * r1 + r2 * r2) - 2.0 * r1 * r2 * cth)
    vhh = par['PHH1'] * np.exp(-par['PHH2'] * rhh)
    ex = np.exp(-par['ALPHAOH'] * (r1 - par['ROH']))
    voh1 = par['DEOH'] * ex * (ex - 2.0)
    ex = np.exp(-par['ALPHAOH'] * (r2 - par['ROH']))
    voh2 = par['DEOH'] * ex * (ex - 2.0)
    f1 = [np.ones_like(x1)]
    f2 = [np.ones_like(x1)]
    f3 = [np.ones_like(x1)]
    for _ in range(14):
        f1.append(f1[-1] * x1)
        f2.append(f2[-1] * x2)
        f3.append(f3[-1] * x3)
    v = np.zeros_like(x1)
    for j in range(1, len(c)):
        a, b, k = e[j]
        v = v + c[j] * (f1[a] * f2[b] + f1[b] * f2[a]) * f3[k]
    damp = np.exp(-par['B1'] * ((r1 - reoh) * (r1 - reoh) + (r2 - reoh) * (r2 - reoh)))
    return (((v * damp + c[0]) + voh1) + voh2) + vhh


def ps_eq():
    """The surface's equilibrium: r_e = 0.958649 angstrom, theta_e = 104.3475 degrees, [H, H, O]."""
    re_ = 0.958649 / 0.529177249
    t = np.deg2rad(104.3475)
    return np.array([[re_, 0.0, 0.0], [re_ * np.cos(t), re_ * np.sin(t), 0.0], [0.0, 0.0, 0.0]])
