// Potential-energy surfaces and trial wave functions evaluated per walker in registers.
//
// A walker is NP = M*d Cartesian components x[a*d + c] (atom a, component c), the same
// order as the reference's (N, M, d) arrays (pyvibdmc.py:156-169).  Each model is a
// small struct passed to kernels by value; `Real` is float (production) or double
// (parity / FP64 runs).  Operation order in the FP64 forms follows the NumPy
// expressions of the oracle so that FP64 results agree with it to a few ulps.
#pragma once
#include <utility>

#include "pvd_common.hpp"
#include "pvd_ps_table.h"

namespace pvd {

enum PesKind : int {
  PES_EXTERNAL = 0,    // energies supplied from outside (host callback or surrogate)
  PES_HARMONIC_X = 1,  // V = c*(x*x), x = first coordinate of atom 0
  PES_MORSE_1D = 2,    // V = De*(t*t), t = 1 - exp(-alpha*x)
  PES_MORSE_BEND = 3,  // two Morse O-H stretches + harmonic bend, atoms (h1, h2, o)
  PES_CLUSTER6 = 4,    // two Morse+bend monomers + O-O Morse + inter-monomer Coulomb
  PES_PS_WATER = 5,    // Partridge-Schwenke water surface, atoms (h1, h2, o)
};

struct PesParams {
  int kind;
  int idx[4];   // atom indices (MORSE_BEND: h1, h2, o)
  double p[12]; // HARMONIC_X: c | MORSE_1D: De, -alpha | MORSE_BEND: De, a, re, kb, theta0 |
                // CLUSTER6: De, a, re, kb, theta0, De_oo, a_oo, re_oo, qO, qH
};

template <typename R>
PVD_DEV R pvd_exp(R x);
// single precision: the hardware 2^x on x log2(e) (a few ulp; the FP32 surfaces and
// weights are far less sensitive than that)
template <>
PVD_DEV float pvd_exp<float>(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
template <>
PVD_DEV double pvd_exp<double>(double x) { return exp(x); }
template <typename R>
PVD_DEV R pvd_sqrt(R x) { return sqrt(x); }
template <>
PVD_DEV float pvd_sqrt<float>(float x) { return __fsqrt_rn(x); }
template <typename R>
PVD_DEV R pvd_acos(R x);
template <>
PVD_DEV float pvd_acos<float>(float x) { return acosf(x); }
template <>
PVD_DEV double pvd_acos<double>(double x) { return acos(x); }
template <typename R>
PVD_DEV R pvd_log(R x);
template <>
PVD_DEV float pvd_log<float>(float x) { return logf(x); }
template <>
PVD_DEV double pvd_log<double>(double x) { return log(x); }

// a / b with the reciprocal rb = 1 / b computed once by the caller: single precision
// multiplies by it; double precision keeps the literal division (the FP64 forms follow
// the oracle's arithmetic), and the unused reciprocal is dead code there.
template <typename R>
PVD_DEV R divr(R a, R b, R rb) {
  if constexpr (sizeof(R) == 4) {
    (void)b;
    return a * rb;
  } else {
    (void)rb;
    return a / b;
  }
}

template <typename R>
PVD_DEV R clamp1(R c) {
  return c > R(1) ? R(1) : (c < R(-1) ? R(-1) : c);
}

template <typename R>
PVD_DEV R norm3(R a, R b, R c) {
  return pvd_sqrt<R>(a * a + b * b + c * c);
}

// Morse + bend of one triatomic whose atoms start at component offsets h1, h2, o (d = 3).
template <typename R>
PVD_DEV R morse_bend_v(const R* x, int h1, int h2, int o, R De, R a, R re, R hk, R th0) {
  const R u0 = x[h1 + 0] - x[o + 0], u1 = x[h1 + 1] - x[o + 1], u2 = x[h1 + 2] - x[o + 2];
  const R v0 = x[h2 + 0] - x[o + 0], v1 = x[h2 + 1] - x[o + 1], v2 = x[h2 + 2] - x[o + 2];
  const R r1 = norm3(u0, u1, u2);
  const R r2 = norm3(v0, v1, v2);
  const R dot = u0 * v0 + u1 * v1 + u2 * v2;
  const R c = clamp1(dot / (r1 * r2));
  const R th = pvd_acos<R>(c);
  const R m1 = R(1) - pvd_exp<R>(-a * (r1 - re));
  const R m2 = R(1) - pvd_exp<R>(-a * (r2 - re));
  const R dt = th - th0;
  return De * (m1 * m1) + De * (m2 * m2) + hk * (dt * dt);
}

// gradient of morse_bend_v accumulated into g (NP components)
template <typename R>
PVD_DEV void morse_bend_g(const R* x, R* g, int h1, int h2, int o, R De, R a, R re, R kb, R th0) {
  R u[3], v[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    u[c] = x[h1 + c] - x[o + c];
    v[c] = x[h2 + c] - x[o + c];
  }
  const R r1 = norm3(u[0], u[1], u[2]);
  const R r2 = norm3(v[0], v[1], v[2]);
  const R e1 = pvd_exp<R>(-a * (r1 - re));
  const R e2 = pvd_exp<R>(-a * (r2 - re));
  const R g1 = R(2) * De * (R(1) - e1) * a * e1;
  const R g2 = R(2) * De * (R(1) - e2) * a * e2;
  R uh[3], vh[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    uh[c] = u[c] / r1;
    vh[c] = v[c] / r2;
  }
  const R cth = clamp1(uh[0] * vh[0] + uh[1] * vh[1] + uh[2] * vh[2]);
  const R s = pvd_sqrt<R>(R(1) - cth * cth);
  const R gt = kb * (pvd_acos<R>(cth) - th0);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const R dcu = (vh[c] - cth * uh[c]) / r1;
    const R dcv = (uh[c] - cth * vh[c]) / r2;
    const R dtu = -dcu / s, dtv = -dcv / s;
    g[h1 + c] += g1 * uh[c] + gt * dtu;
    g[h2 + c] += g2 * vh[c] + gt * dtv;
    g[o + c] += -g1 * uh[c] - g2 * vh[c] - gt * (dtu + dtv);
  }
}

// x components of atoms idx[0], idx[1], idx[2] into g[0..9) without dynamic register indexing
template <typename R, int NP>
PVD_DEV void gather3(const R* x, const int* idx, R* g) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int b = 3 * idx[a];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      R v = R(0);
#pragma unroll
      for (int p = 0; p < NP; ++p)
        if (p == b + c) v = x[p];
      g[3 * a + c] = v;
    }
  }
}

// Evaluate a device PES.  NP is the number of Cartesian components per walker.
// The [O, H, H, O, H, H] cluster: two Morse+bend monomers, an O-O Morse term and the
// Coulomb energy of the inter-monomer atom pairs.
template <typename R>
PVD_DEV R cluster6_v(const PesParams& P, const R* x) {
  const R De = R(P.p[0]), a = R(P.p[1]), re = R(P.p[2]), hk = R(0.5 * P.p[3]), th0 = R(P.p[4]);
  R v = morse_bend_v<R>(x, 3, 6, 0, De, a, re, hk, th0) + morse_bend_v<R>(x, 12, 15, 9, De, a, re, hk, th0);
  const R roo = norm3(x[0] - x[9], x[1] - x[10], x[2] - x[11]);
  const R m = R(1) - pvd_exp<R>(-R(P.p[6]) * (roo - R(P.p[7])));
  v = v + R(P.p[5]) * (m * m);
  const R q[3] = {R(P.p[8]), R(P.p[9]), R(P.p[9])};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const R r = norm3(x[3 * i] - x[9 + 3 * j], x[3 * i + 1] - x[10 + 3 * j], x[3 * i + 2] - x[11 + 3 * j]);
      v = v + (q[i] * q[j]) / r;
    }
  return v;
}

// ------------------------------------------------------------ Partridge-Schwenke water
// The surface of H. Partridge and D. W. Schwenke (J. Chem. Phys. 106, 4618 (1997)) as the
// reference evaluates it (sample_potentials/FortPots/Partridge_Schwenke_H2O/calc_h2o_pot.f,
// h2opes_v2.f:460-489):
//   V = [sum_j c_j (x1^a x2^b + x2^a x1^b) x3^k] exp(-b1 ((r1-re)^2 + (r2-re)^2)) + c_0
//       + D ex (ex - 2) for each O-H distance (ex = exp(-alpha (r - r0))) + A exp(-b r_HH),
//   x1 = (r1 - re)/re, x2 = (r2 - re)/re, x3 = cos(theta) - cos(theta_e).
// The 245-term table (pvd_ps_table.h) is unrolled at compile time, so the powers of x1, x2,
// x3 stay in registers and each coefficient is an immediate.  Double precision keeps the
// reference's order of operations (terms summed in table order); single precision fuses
// each term into one multiply-add.
inline constexpr double PS_C[PVD_PS_NTERMS] = {PVD_PS_COEFFS};
inline constexpr unsigned char PS_E[3 * PVD_PS_NTERMS] = {PVD_PS_EXPONENTS};

template <typename R, int J>
PVD_DEV R ps_term(const R* f1, const R* f2, const R* f3, R v) {
  constexpr int a = PS_E[3 * J], b = PS_E[3 * J + 1], k = PS_E[3 * J + 2];
  constexpr double c = PS_C[J];
  if constexpr (sizeof(R) == 8) {
    return v + (R(c) * (f1[a] * f2[b] + f1[b] * f2[a])) * f3[k];
  } else {
    return __builtin_fmaf(R(c) * (f1[a] * f2[b] + f1[b] * f2[a]), f3[k], v);
  }
}

template <typename R, int... J>
PVD_DEV R ps_poly(const R* f1, const R* f2, const R* f3, std::integer_sequence<int, J...>) {
  R v = R(0);
  ((v = ps_term<R, J + 1>(f1, f2, f3, v)), ...);  // terms 1 .. 244 in table order
  return v;
}

// x: [h1, h2, o] components 0..8
template <typename R>
PVD_DEV R ps_water_v(const R* x) {
  R r1s = R(0), r2s = R(0), ct = R(0);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const R d1 = x[6 + c] - x[c], d2 = x[6 + c] - x[3 + c];
    r1s = r1s + d1 * d1;
    r2s = r2s + d2 * d2;
    ct = ct + d1 * d2;
  }
  const R r1 = pvd_sqrt<R>(r1s), r2 = pvd_sqrt<R>(r2s);
  const R re = R(PVD_PS_REOH);
  R cth;
  if constexpr (sizeof(R) == 8) {
    cth = cos(acos((ct / r1) / r2));  // the reference's theta = acos(...), then cos(theta)
  } else {
    cth = ct / (r1 * r2);
  }
  const R x1 = (r1 - re) / re, x2 = (r2 - re) / re, x3 = cth - R(PVD_PS_CE);
  const R rhh = pvd_sqrt<R>((r1 * r1 + r2 * r2) - ((R(2) * r1) * r2) * cth);
  const R vhh = R(PVD_PS_PHH1) * pvd_exp<R>(-(R(PVD_PS_PHH2) * rhh));
  R ex = pvd_exp<R>(-(R(PVD_PS_ALPHAOH) * (r1 - R(PVD_PS_ROH))));
  const R voh1 = (R(PVD_PS_DEOH) * ex) * (ex - R(2));
  ex = pvd_exp<R>(-(R(PVD_PS_ALPHAOH) * (r2 - R(PVD_PS_ROH))));
  const R voh2 = (R(PVD_PS_DEOH) * ex) * (ex - R(2));
  R f1[15], f2[15], f3[15];
  f1[0] = f2[0] = f3[0] = R(1);
#pragma unroll
  for (int j = 1; j < 15; ++j) {
    f1[j] = f1[j - 1] * x1;
    f2[j] = f2[j - 1] * x2;
    f3[j] = f3[j - 1] * x3;
  }
  const R poly = ps_poly<R>(f1, f2, f3, std::make_integer_sequence<int, PVD_PS_NTERMS - 1>());
  const R d1 = r1 - re, d2 = r2 - re;
  const R damp = pvd_exp<R>(-(R(PVD_PS_B1) * (d1 * d1 + d2 * d2)));
  return (((poly * damp + R(PS_C[0])) + voh1) + voh2) + vhh;
}

template <typename R, int NP>
PVD_DEV R pes_eval(const PesParams& P, const R* x) {
  switch (P.kind) {
    case PES_HARMONIC_X: {
      const R x0 = x[0];
      return R(P.p[0]) * (x0 * x0);
    }
    case PES_MORSE_1D: {
      const R t = R(1) - pvd_exp<R>(R(P.p[1]) * x[0]);
      return R(P.p[0]) * (t * t);
    }
    case PES_MORSE_BEND:
      if constexpr (NP >= 9) {
        if (P.idx[0] == 0 && P.idx[1] == 1 && P.idx[2] == 2)  // [H, H, O], the usual order
          return morse_bend_v<R>(x, 0, 3, 6, R(P.p[0]), R(P.p[1]), R(P.p[2]), R(0.5 * P.p[3]), R(P.p[4]));
        // other orders: gather (h1, h2, o) with selects so the walker stays in registers
        R g[9];
        gather3<R, NP>(x, P.idx, g);
        return morse_bend_v<R>(g, 0, 3, 6, R(P.p[0]), R(P.p[1]), R(P.p[2]), R(0.5 * P.p[3]), R(P.p[4]));
      }
      break;
    case PES_CLUSTER6:
      if constexpr (NP == 18) return cluster6_v<R>(P, x);
      break;
    case PES_PS_WATER:
      if constexpr (NP == 9) {
        if (P.idx[0] == 0 && P.idx[1] == 1 && P.idx[2] == 2) return ps_water_v<R>(x);
        R g[9];
        gather3<R, NP>(x, P.idx, g);
        return ps_water_v<R>(g);
      }
      break;
    default:
      break;
  }
  return R(0) / R(0);  // unsupported: NaN trips the weight guard
}

// dV/dx for the analytic surfaces (forces are -g).
template <typename R, int NP>
PVD_DEV void pes_grad(const PesParams& P, const R* x, R* g) {
#pragma unroll
  for (int p = 0; p < NP; ++p) g[p] = R(0);
  switch (P.kind) {
    case PES_HARMONIC_X:
      g[0] = R(2) * R(P.p[0]) * x[0];
      break;
    case PES_MORSE_1D: {
      const R e = pvd_exp<R>(R(P.p[1]) * x[0]);
      g[0] = R(2) * R(P.p[0]) * (R(1) - e) * (-R(P.p[1])) * e;
      break;
    }
    case PES_MORSE_BEND:
      if constexpr (NP >= 9) {
        R xg[9], gg[9];
        gather3<R, NP>(x, P.idx, xg);
#pragma unroll
        for (int q = 0; q < 9; ++q) gg[q] = R(0);
        morse_bend_g<R>(xg, gg, 0, 3, 6, R(P.p[0]), R(P.p[1]), R(P.p[2]), R(P.p[3]), R(P.p[4]));
        // scatter back
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int p = 0; p < NP; ++p)
              if (p == 3 * P.idx[a] + c) g[p] += gg[3 * a + c];
      }
      break;
    case PES_CLUSTER6:
      if constexpr (NP == 18) {
        const R De = R(P.p[0]), a = R(P.p[1]), re = R(P.p[2]), kb = R(P.p[3]), th0 = R(P.p[4]);
        morse_bend_g<R>(x, g, 3, 6, 0, De, a, re, kb, th0);
        morse_bend_g<R>(x, g, 12, 15, 9, De, a, re, kb, th0);
        {
          R d[3];
#pragma unroll
          for (int c = 0; c < 3; ++c) d[c] = x[c] - x[9 + c];
          const R r = norm3(d[0], d[1], d[2]);
          const R e = pvd_exp<R>(-R(P.p[6]) * (r - R(P.p[7])));
          const R s = R(2) * R(P.p[5]) * (R(1) - e) * R(P.p[6]) * e / r;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            g[c] += s * d[c];
            g[9 + c] -= s * d[c];
          }
        }
        const R q[3] = {R(P.p[8]), R(P.p[9]), R(P.p[9])};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            R d[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) d[c] = x[3 * i + c] - x[9 + 3 * j + c];
            const R r = norm3(d[0], d[1], d[2]);
            const R s = -(q[i] * q[j]) / (r * r * r);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              g[3 * i + c] += s * d[c];
              g[9 + 3 * j + c] -= s * d[c];
            }
          }
      }
      break;
    default:
      break;
  }
}

// ------------------------------------------------------------------------------ trial psi
enum TrialKind : int {
  TRIAL_NONE = 0,
  TRIAL_HARMONIC_1D = 1,  // psi = c exp(-alpha x^2/2) on coordinate 0
  TRIAL_GAUSS = 2,        // product of Gaussians in bond lengths and bond angles
  TRIAL_HOST = 3,         // evaluated by the caller (split importance-sampled move)
};

#define PVD_MAX_TB 8
#define PVD_MAX_TA 4

struct TrialParams {
  int kind;
  int nb, na;
  int bi[PVD_MAX_TB], bj[PVD_MAX_TB];
  int ai[PVD_MAX_TA], aj[PVD_MAX_TA], ak[PVD_MAX_TA];
  double balpha[PVD_MAX_TB], br0[PVD_MAX_TB];
  double aalpha[PVD_MAX_TA], ath0[PVD_MAX_TA];
  double c0, c1, c2;  // HARMONIC_1D: (alpha/pi)^0.25, alpha, alpha^2 | GAUSS: log normalisation
};

// Result of a trial evaluation at one walker: psi, d ln psi / dx and d2 psi / dx2 / psi
// (diagonal, per Cartesian component), i.e. exactly what a reference ``deriv_function``
// returns (imp_samp.py:21-26).
template <typename R, int NP>
struct TrialOut {
  R psi;
  R lnpsi;
  R d1[NP];
  R d2[NP];
};

// Gradient / diagonal-Hessian contributions of one Gaussian factor in a bond length
// r = |x_i - x_j| (component offsets i, j): ln psi -= al (r - r0)^2 / 2.
template <typename R, int NP>
PVD_DEV void bond_term(const R* x, int i, int j, R al, R r0, R* g, R* h, R& q2) {
  R u[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    R xi = R(0), xj = R(0);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      if (p == i + c) xi = x[p];
      if (p == j + c) xj = x[p];
    }
    u[c] = xi - xj;
  }
  const R r = norm3(u[0], u[1], u[2]);
  const R rr = R(1) / r;
  const R dq = r - r0;
  q2 += al * (dq * dq);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const R e = divr(u[c], r, rr);
    const R d2 = divr(R(1) - e * e, r, rr);
    const R t1 = -al * dq * e;
    const R t2 = -al * (e * e + dq * d2);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      if (p == i + c) {
        g[p] += t1;
        h[p] += t2;
      }
      if (p == j + c) {
        g[p] -= t1;
        h[p] += t2;
      }
    }
  }
}

// One Gaussian factor in the bond angle i-j-k (vertex j), through c = cos(theta):
// dtheta = -dc/s, d2theta = -d2c/s - c dc^2/s^3, s = sin(theta).
template <typename R, int NP>
PVD_DEV void angle_term(const R* x, int i, int j, int kk, R al, R th0, R* g, R* h, R& q2) {
  R u[3], v[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    R xi = R(0), xj = R(0), xk = R(0);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      if (p == i + c) xi = x[p];
      if (p == j + c) xj = x[p];
      if (p == kk + c) xk = x[p];
    }
    u[c] = xi - xj;
    v[c] = xk - xj;
  }
  const R ru = norm3(u[0], u[1], u[2]), rv = norm3(v[0], v[1], v[2]);
  const R rru = R(1) / ru, rrv = R(1) / rv;
  R uh[3], vh[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    uh[c] = divr(u[c], ru, rru);
    vh[c] = divr(v[c], rv, rrv);
  }
  const R cth = clamp1(uh[0] * vh[0] + uh[1] * vh[1] + uh[2] * vh[2]);
  const R s = pvd_sqrt<R>(R(1) - cth * cth);
  const R rs = R(1) / s, rs3 = rs * rs * rs;
  const R dq = pvd_acos<R>(cth) - th0;
  q2 += al * (dq * dq);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const R dcu = divr(vh[c] - cth * uh[c], ru, rru);
    const R dcv = divr(uh[c] - cth * vh[c], rv, rrv);
    const R duu = divr(R(-2) * uh[c] * vh[c] - cth + R(3) * cth * uh[c] * uh[c], ru * ru, rru * rru);
    const R dvv = divr(R(-2) * uh[c] * vh[c] - cth + R(3) * cth * vh[c] * vh[c], rv * rv, rrv * rrv);
    const R duv = divr(R(1) - uh[c] * uh[c] - vh[c] * vh[c] + cth * uh[c] * vh[c], ru * rv, rru * rrv);
    const R dcs[3] = {dcu, dcv, -(dcu + dcv)};
    const R d2cs[3] = {duu, dvv, duu + R(2) * duv + dvv};
    const int at[3] = {i + c, kk + c, j + c};
#pragma unroll
    for (int z = 0; z < 3; ++z) {
      const R dth = divr(-dcs[z], s, rs);
      const R d2th = divr(-d2cs[z], s, rs) - divr(cth * dcs[z] * dcs[z], s * s * s, rs3);
      const R t1 = -al * dq * dth;
      const R t2 = -al * (dth * dth + dq * d2th);
#pragma unroll
      for (int p = 0; p < NP; ++p)
        if (p == at[z]) {
          g[p] += t1;
          h[p] += t2;
        }
    }
  }
}

template <typename R, int NP>
PVD_DEV void trial_eval(const TrialParams& T, const R* x, TrialOut<R, NP>& o) {
  if (T.kind == TRIAL_HARMONIC_1D) {
    // harm_trial_wfn.py:6-40: trl = c*e, derv = c*(-alpha x)*e, sderv = c*(alpha^2 x^2 - alpha)*e
    const R c = R(T.c0), al = R(T.c1), al2 = R(T.c2);
    const R x0 = x[0];
    const R e = pvd_exp<R>(((-al) * (x0 * x0)) / R(2));
    const R trl = c * e;
    const R derv = (c * ((-al) * x0)) * e;
    const R sderv = (c * (al2 * (x0 * x0) - al)) * e;
    o.psi = trl;
    o.lnpsi = pvd_log<R>(c) + ((-al) * (x0 * x0)) / R(2);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      o.d1[p] = R(0);
      o.d2[p] = R(0);
    }
    o.d1[0] = derv / trl;
    o.d2[0] = sderv / trl;
    return;
  }
  // TRIAL_GAUSS: ln psi = lnnorm - sum_k alpha_k (q_k - q0_k)^2 / 2
  R g[NP], h[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    g[p] = R(0);
    h[p] = R(0);
  }
  R q2 = R(0);
  if constexpr (NP % 3 == 0) {
    if (NP == 9 && T.nb == 2 && T.na == 1 && T.bi[0] == 0 && T.bj[0] == 2 && T.bi[1] == 1 && T.bj[1] == 2 &&
        T.ai[0] == 0 && T.aj[0] == 2 && T.ak[0] == 1) {
      // [H, H, O] water pattern with constant indices
      bond_term<R, NP>(x, 0, 6, R(T.balpha[0]), R(T.br0[0]), g, h, q2);
      bond_term<R, NP>(x, 3, 6, R(T.balpha[1]), R(T.br0[1]), g, h, q2);
      angle_term<R, NP>(x, 0, 6, 3, R(T.aalpha[0]), R(T.ath0[0]), g, h, q2);
    } else {
      for (int k = 0; k < T.nb; ++k)
        bond_term<R, NP>(x, 3 * T.bi[k], 3 * T.bj[k], R(T.balpha[k]), R(T.br0[k]), g, h, q2);
      for (int k = 0; k < T.na; ++k)
        angle_term<R, NP>(x, 3 * T.ai[k], 3 * T.aj[k], 3 * T.ak[k], R(T.aalpha[k]), R(T.ath0[k]), g, h, q2);
    }
  }
  o.lnpsi = R(T.c0) - R(0.5) * q2;
  o.psi = pvd_exp<R>(o.lnpsi);
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    o.d1[p] = g[p];
    o.d2[p] = g[p] * g[p] + h[p];
  }
}

// Compile-time model specialisations of the step kernels.  The host selects them when the
// model matches exactly, so the kernels carry no runtime index selects for that layout:
//   MS_PES_WATER      PES_MORSE_BEND with atoms (h1, h2, o) = (0, 1, 2), 3 atoms x 3 dims
//   MS_TRIAL_WATER    TRIAL_GAUSS with bonds (0-2), (1-2) and the angle 0-2-1, same layout
//   MS_PES_CLUSTER    PES_CLUSTER6, atoms [O, H, H, O, H, H] in 3 dims
//   MS_TRIAL_CLUSTER  TRIAL_GAUSS with bonds (1-0), (2-0), (4-3), (5-3) and the angles
//                     1-0-2, 4-3-5, same layout
enum ModelSpec : int { MS_ANY = 0, MS_PES_WATER = 1, MS_TRIAL_WATER = 2, MS_PES_CLUSTER = 4, MS_TRIAL_CLUSTER = 8 };
constexpr int MS_PES_ANY_SPEC = MS_PES_WATER | MS_PES_CLUSTER;
constexpr int MS_TRIAL_ANY_SPEC = MS_TRIAL_WATER | MS_TRIAL_CLUSTER;

template <typename R, int NP, int MS>
PVD_DEV R pes_eval_s(const PesParams& P, const R* x) {
  if constexpr ((MS & MS_PES_WATER) != 0 && NP == 9) {
    return morse_bend_v<R>(x, 0, 3, 6, R(P.p[0]), R(P.p[1]), R(P.p[2]), R(0.5 * P.p[3]), R(P.p[4]));
  } else if constexpr ((MS & MS_PES_CLUSTER) != 0 && NP == 18) {
    return cluster6_v<R>(P, x);
  } else {
    return pes_eval<R, NP>(P, x);
  }
}

// Any Gaussian-product trial (runtime bond and angle lists, 3-D atoms) in the same closed
// form as trial_water_fast: ln psi, g = grad ln psi and the mass-weighted Laplacian
// lap = sum_a (1/m_a) lap_a ln psi (inv_m: 1/m per component, atom a at 3a).  Coordinates
// are gathered and gradients scattered with selects over the NP components, so the walker
// stays in registers.
template <typename R, int NP>
PVD_DEV void trial_gauss_fast(const TrialParams& T, const double* inv_m, const R* x, R& lnpsi, R* g, R& lap) {
#pragma unroll
  for (int p = 0; p < NP; ++p) g[p] = R(0);
  R q2 = R(0);
  lap = R(0);
  auto gather = [&](int a, R* o) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      R v = R(0);
#pragma unroll
      for (int p = 0; p < NP; ++p)
        if (p == 3 * a + c) v = x[p];
      o[c] = v;
    }
  };
  auto scatter = [&](int a, const R* d) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int p = 0; p < NP; ++p)
        if (p == 3 * a + c) g[p] += d[c];
  };
  for (int k = 0; k < T.nb; ++k) {
    const int i = T.bi[k], j = T.bj[k];
    R xi[3], xj[3], u[3];
    gather(i, xi);
    gather(j, xj);
#pragma unroll
    for (int c = 0; c < 3; ++c) u[c] = xi[c] - xj[c];
    const R r = norm3(u[0], u[1], u[2]), ir = R(1) / r;
    const R al = R(T.balpha[k]), dr = r - R(T.br0[k]);
    q2 += al * (dr * dr);
    R gi[3], gj[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      gi[c] = -al * dr * (u[c] * ir);
      gj[c] = -gi[c];
    }
    scatter(i, gi);
    scatter(j, gj);
    lap += -al * (R(1) + dr * R(2) * ir) * (R)(inv_m[3 * i] + inv_m[3 * j]);
  }
  for (int k = 0; k < T.na; ++k) {
    const int i = T.ai[k], j = T.aj[k], kk = T.ak[k];
    R xi[3], xj[3], xk[3], u[3], v[3];
    gather(i, xi);
    gather(j, xj);
    gather(kk, xk);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      u[c] = xi[c] - xj[c];
      v[c] = xk[c] - xj[c];
    }
    const R ru = norm3(u[0], u[1], u[2]), rv = norm3(v[0], v[1], v[2]);
    const R iru = R(1) / ru, irv = R(1) / rv;
    R uh[3], vh[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      uh[c] = u[c] * iru;
      vh[c] = v[c] * irv;
    }
    const R cth = clamp1(uh[0] * vh[0] + uh[1] * vh[1] + uh[2] * vh[2]);
    const R sth = pvd_sqrt<R>(R(1) - cth * cth), isth = R(1) / sth;
    const R al = R(T.aalpha[k]), dq = pvd_acos<R>(cth) - R(T.ath0[k]);
    q2 += al * (dq * dq);
    R gi[3], gk[3], gj[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      gi[c] = -al * dq * (-(vh[c] - cth * uh[c]) * (iru * isth));
      gk[c] = -al * dq * (-(uh[c] - cth * vh[c]) * (irv * isth));
      gj[c] = -(gi[c] + gk[c]);
    }
    scatter(i, gi);
    scatter(kk, gk);
    scatter(j, gj);
    const R iru2 = iru * iru, irv2 = irv * irv, cot = cth * isth;
    const R la_i = -al * (iru2 + dq * cot * iru2);
    const R la_k = -al * (irv2 + dq * cot * irv2);
    const R la_j = -al * ((iru2 + irv2 - R(2) * cth * iru * irv) + dq * (cot * (iru2 + irv2) - R(2) * iru * irv * isth));
    lap += la_i * (R)inv_m[3 * i] + la_k * (R)inv_m[3 * kk] + la_j * (R)inv_m[3 * j];
  }
  lnpsi = R(T.c0) - R(0.5) * q2;
}

// The [H, H, O] Gaussian-product trial (bonds 0-2 and 1-2, angle 0-2-1) for importance
// sampling: ln psi, g = grad ln psi (9 components) and the mass-weighted Laplacian
// lap = sum_a (1/m_a) lap_a ln psi in closed form, so the local kinetic energy is
// -1/2 (sum_p (1/m_p) g_p^2 + lap) without per-component second derivatives:
//   lap_a ln psi = -sum_k alpha_k (|grad_a q_k|^2 + (q_k - q0_k) lap_a q_k),
//   bond r:  |grad r|^2 = 1, lap r = 2/r at both atoms;
//   angle theta (vertex j, arms u = x_i - x_j, v = x_k - x_j): |grad_i|^2 = 1/ru^2,
//   |grad_k|^2 = 1/rv^2, |grad_j|^2 = 1/ru^2 + 1/rv^2 - 2 cos/(ru rv); lap_i = cot/ru^2,
//   lap_k = cot/rv^2, lap_j = cot (1/ru^2 + 1/rv^2) - 2/(ru rv sin).
// im_h, im_o: 1/m of the hydrogens and of the oxygen.
template <typename R>
PVD_DEV void trial_water_fast(const TrialParams& T, const R* x, R im_h, R im_o, R& lnpsi, R* g, R& lap) {
  R u[3], v[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    u[c] = x[c] - x[6 + c];
    v[c] = x[3 + c] - x[6 + c];
  }
  const R ru = norm3(u[0], u[1], u[2]), rv = norm3(v[0], v[1], v[2]);
  const R iru = R(1) / ru, irv = R(1) / rv;
  const R a1 = R(T.balpha[0]), a2 = R(T.balpha[1]), at = R(T.aalpha[0]);
  const R d1 = ru - R(T.br0[0]), d2 = rv - R(T.br0[1]);
  R uh[3], vh[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    uh[c] = u[c] * iru;
    vh[c] = v[c] * irv;
  }
  const R cth = clamp1(uh[0] * vh[0] + uh[1] * vh[1] + uh[2] * vh[2]);
  const R sth = pvd_sqrt<R>(R(1) - cth * cth);
  const R isth = R(1) / sth;
  const R dq = pvd_acos<R>(cth) - R(T.ath0[0]);
  lnpsi = R(T.c0) - R(0.5) * ((a1 * (d1 * d1) + a2 * (d2 * d2)) + at * (dq * dq));
  // gradients: bonds -alpha dr r_hat; angle -alpha dq grad theta
  const R gb1 = -a1 * d1, gb2 = -a2 * d2, ga = -at * dq;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const R tu = -(vh[c] - cth * uh[c]) * (iru * isth);  // d theta / d u
    const R tv = -(uh[c] - cth * vh[c]) * (irv * isth);  // d theta / d v
    g[c] = gb1 * uh[c] + ga * tu;
    g[3 + c] = gb2 * vh[c] + ga * tv;
    g[6 + c] = -(gb1 * uh[c] + gb2 * vh[c]) - ga * (tu + tv);
  }
  const R iru2 = iru * iru, irv2 = irv * irv, cot = cth * isth;
  const R lb1 = -a1 * (R(1) + d1 * R(2) * iru), lb2 = -a2 * (R(1) + d2 * R(2) * irv);
  const R la_i = -at * (iru2 + dq * cot * iru2);
  const R la_k = -at * (irv2 + dq * cot * irv2);
  const R la_j = -at * ((iru2 + irv2 - R(2) * cth * iru * irv) + dq * (cot * (iru2 + irv2) - R(2) * iru * irv * isth));
  lap = im_h * (lb1 + lb2 + la_i + la_k) + im_o * (lb1 + lb2 + la_j);
}

// One monomer of a Gaussian-product trial in the closed form of trial_water_fast: bonds
// h1-o and h2-o (parameters KB, KB + 1) and the angle h1-o-h2 (parameter KA), atoms at the
// compile-time component offsets H1, H2, O.  Adds to q2 = sum alpha (q - q0)^2, to the
// gradient g of ln psi and to the mass-weighted Laplacian lap.
template <typename R, int H1, int H2, int O, int KB, int KA>
PVD_DEV void trial_mono_acc(const TrialParams& T, const R* x, R im_h, R im_o, R& q2, R* g, R& lap) {
  R u[3], v[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    u[c] = x[H1 + c] - x[O + c];
    v[c] = x[H2 + c] - x[O + c];
  }
  const R ru = norm3(u[0], u[1], u[2]), rv = norm3(v[0], v[1], v[2]);
  const R iru = R(1) / ru, irv = R(1) / rv;
  const R a1 = R(T.balpha[KB]), a2 = R(T.balpha[KB + 1]), at = R(T.aalpha[KA]);
  const R d1 = ru - R(T.br0[KB]), d2 = rv - R(T.br0[KB + 1]);
  R uh[3], vh[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    uh[c] = u[c] * iru;
    vh[c] = v[c] * irv;
  }
  const R cth = clamp1(uh[0] * vh[0] + uh[1] * vh[1] + uh[2] * vh[2]);
  const R sth = pvd_sqrt<R>(R(1) - cth * cth);
  const R isth = R(1) / sth;
  const R dq = pvd_acos<R>(cth) - R(T.ath0[KA]);
  q2 += (a1 * (d1 * d1) + a2 * (d2 * d2)) + at * (dq * dq);
  const R gb1 = -a1 * d1, gb2 = -a2 * d2, ga = -at * dq;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const R tu = -(vh[c] - cth * uh[c]) * (iru * isth);
    const R tv = -(uh[c] - cth * vh[c]) * (irv * isth);
    g[H1 + c] += gb1 * uh[c] + ga * tu;
    g[H2 + c] += gb2 * vh[c] + ga * tv;
    g[O + c] += -(gb1 * uh[c] + gb2 * vh[c]) - ga * (tu + tv);
  }
  const R iru2 = iru * iru, irv2 = irv * irv, cot = cth * isth;
  const R lb1 = -a1 * (R(1) + d1 * R(2) * iru), lb2 = -a2 * (R(1) + d2 * R(2) * irv);
  const R la_i = -at * (iru2 + dq * cot * iru2);
  const R la_k = -at * (irv2 + dq * cot * irv2);
  const R la_j = -at * ((iru2 + irv2 - R(2) * cth * iru * irv) + dq * (cot * (iru2 + irv2) - R(2) * iru * irv * isth));
  lap += im_h * (lb1 + lb2 + la_i + la_k) + im_o * (lb1 + lb2 + la_j);
}

// The [O, H, H, O, H, H] cluster trial (bonds 1-0, 2-0, 4-3, 5-3; angles 1-0-2, 4-3-5) as
// two monomers; im_h, im_o: 1/m of the hydrogens and of the oxygens.
template <typename R>
PVD_DEV void trial_cluster_fast(const TrialParams& T, const R* x, R im_h, R im_o, R& lnpsi, R* g, R& lap) {
#pragma unroll
  for (int p = 0; p < 18; ++p) g[p] = R(0);
  R q2 = R(0);
  lap = R(0);
  trial_mono_acc<R, 3, 6, 0, 0, 0>(T, x, im_h, im_o, q2, g, lap);
  trial_mono_acc<R, 12, 15, 9, 2, 1>(T, x, im_h, im_o, q2, g, lap);
  lnpsi = R(T.c0) - R(0.5) * q2;
}

template <typename R, int NP, int MS>
PVD_DEV void trial_eval_s(const TrialParams& T, const R* x, TrialOut<R, NP>& o) {
  if constexpr ((MS & MS_TRIAL_WATER) != 0 && NP == 9) {
    R g[NP], h[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      g[p] = R(0);
      h[p] = R(0);
    }
    R q2 = R(0);
    bond_term<R, NP>(x, 0, 6, R(T.balpha[0]), R(T.br0[0]), g, h, q2);
    bond_term<R, NP>(x, 3, 6, R(T.balpha[1]), R(T.br0[1]), g, h, q2);
    angle_term<R, NP>(x, 0, 6, 3, R(T.aalpha[0]), R(T.ath0[0]), g, h, q2);
    o.lnpsi = R(T.c0) - R(0.5) * q2;
    o.psi = pvd_exp<R>(o.lnpsi);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      o.d1[p] = g[p];
      o.d2[p] = g[p] * g[p] + h[p];
    }
  } else {
    trial_eval<R, NP>(T, x, o);
  }
}

}  // namespace pvd
