"""BBKS transfer-function table in the CLASS format that loadTransferFunctions
reads (ic_basic.hpp:494-717), for the initial conditions of the target
workloads ("ICs ... from a BBKS transfer table written to a tk file").

Transfer function: Bardeen, Bond, Kaiser & Szalay (1986), eq. G3,
    T(q) = ln(1 + 2.34 q) / (2.34 q) [1 + 3.89 q + (16.1 q)^2 + (5.46 q)^3 + (6.71 q)^4]^(-1/4),
    q = k / (Gamma h Mpc^-1),  Gamma = Omega_m h  (k in h/Mpc, baryons folded into the matter).

Normalisation and gauge follow the CLASS table the reference ships
(class_tk.dat: Newtonian gauge, initial curvature zeta = 1, k in h/Mpc,
theta in 1/Mpc): in matter domination the potential is Phi = (3/5) zeta T(k),
constant in time, and the Newtonian-gauge density contrast and velocity
divergence of the (single) matter fluid are
    delta = -2 Phi - (2/3) (k / Hc)^2 Phi,     theta = (2/3) k^2 Phi / Hc,
with Hc = a H the conformal Hubble rate in 1/Mpc.  With these,
generateIC_basic's potential, (1.5 Hc^2 delta + rescale theta / k^2) with
rescale = 4.5 Hc^3 in matter domination, is -k^2 Phi exactly.
"""
import numpy as np

C_KMS = 299792.458  # speed of light in km/s


def bbks_T(k_hmpc, Omega_m, h):
    """BBKS transfer function at k in h/Mpc."""
    q = np.asarray(k_hmpc, dtype=np.float64) / (Omega_m * h)
    x = 2.34 * q
    lnf = np.where(x > 1e-8, np.log1p(x) / np.where(x > 0, x, 1.0), 1.0 - x / 2.0)
    return lnf * (1.0 + 3.89 * q + (16.1 * q) ** 2 + (5.46 * q) ** 3 + (6.71 * q) ** 4) ** -0.25


def conformal_hubble(z, Omega_m, Omega_L, h, Omega_rad=0.0):
    """a H(a) in 1/Mpc for a flat matter + Lambda (+ radiation) background."""
    a = 1.0 / (1.0 + z)
    H0 = 100.0 * h / C_KMS  # 1/Mpc
    return a * H0 * np.sqrt(Omega_m / a ** 3 + Omega_L + Omega_rad / a ** 4)


def bbks_table(k_hmpc, z, Omega_m, h, Omega_L=None):
    """Columns (k, d_tot, t_tot) of a single matter fluid (d_cdm = d_tot,
    t_cdm = t_tot with the baryons folded in): k in h/Mpc, theta in 1/Mpc."""
    if Omega_L is None:
        Omega_L = 1.0 - Omega_m
    k = np.asarray(k_hmpc, dtype=np.float64)
    Hc = conformal_hubble(z, Omega_m, Omega_L, h)
    kk = k * h  # 1/Mpc
    Phi = 0.6 * bbks_T(k, Omega_m, h)
    d = -2.0 * Phi - (2.0 / 3.0) * (kk / Hc) ** 2 * Phi
    t = (2.0 / 3.0) * kk ** 2 * Phi / Hc
    return k, d, t


def write_class_tk(path, z, Omega_m, h, kmin=1e-5, kmax=300.0, n=400, Omega_L=None):
    """Writes a CLASS-format table (columns k, d_cdm, d_tot, t_cdm, t_tot) on
    n log-spaced wave numbers in h/Mpc; returns the arrays written."""
    k = np.logspace(np.log10(kmin), np.log10(kmax), n)
    k, d, t = bbks_table(k, z, Omega_m, h, Omega_L)
    with open(path, "w") as f:
        f.write("# BBKS transfer functions T_i(k) (normalized to initial curvature=1) at redshift z=%g\n" % z)
        f.write("# Newtonian gauge; single matter fluid (Omega_m = %g, h = %g)\n" % (Omega_m, h))
        f.write("# for k=%.6e to %.6e h/Mpc,\n# number of wavenumbers equal to %d\n" % (k[0], k[-1], n))
        f.write("#    1:k (h/Mpc)              2:d_cdm                  3:d_tot                  "
                "4:t_cdm                  5:t_tot\n")
        for row in zip(k, d, d, t, t):
            f.write("  " + "  ".join("%.15e" % v for v in row) + "\n")
    return k, d, t
