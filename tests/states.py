"""Seeded synthetic box states shared by the oracle and GPU tests.

Particles start on the sc1 lattice of the reference's crystal template
(one particle per cell at the cell centre, sc1_crystal.dat) displaced by a
smooth Gaussian random field; fields are smooth Gaussian random fields.  IDs
are the lattice index, so storage order is cell order.
"""
import numpy as np


def smooth_field(N, rng, amp, ncomp=None, kcut=None):
    """Real field(s) with Gaussian-filtered white noise, rms = amp."""
    shape = (N, N, N) if ncomp is None else (ncomp, N, N, N)
    white = rng.standard_normal(shape)
    k = np.fft.fftfreq(N) * N
    kz, ky, kx = np.meshgrid(k, k, k[: N // 2 + 1], indexing="ij")
    kcut = kcut if kcut is not None else max(2.0, N / 8.0)
    filt = np.exp(-(kx ** 2 + ky ** 2 + kz ** 2) / (2 * kcut ** 2))
    axes = (-3, -2, -1)
    f = np.fft.irfftn(np.fft.rfftn(white, axes=axes) * filt, s=(N, N, N), axes=axes)
    f *= amp / f.std()
    return np.ascontiguousarray(f)


def lattice_particles(N, rng, disp_cells=0.05, qamp=2e-6, npc_side=1):
    """npc_side^3 particles per cell on a lattice, displaced smoothly."""
    M = N * npc_side
    g = (np.arange(M) + 0.5) / M
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    lat = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    disp = smooth_field(M, rng, 1.0, ncomp=3)
    d = np.stack([disp[0].ravel(), disp[1].ravel(), disp[2].ravel()], axis=1)
    pos = np.mod(lat + d * disp_cells / N, 1.0)
    pos[pos >= 1.0] = 0.0
    q = d * qamp + rng.standard_normal(d.shape) * qamp * 0.1
    ids = np.arange(M ** 3, dtype=np.int64)
    return np.ascontiguousarray(pos), np.ascontiguousarray(q), ids


def gr_state(N, seed, npc_side=1):
    rng = np.random.default_rng(seed)
    pos, q, ids = lattice_particles(N, rng, npc_side=npc_side)
    st = dict(pos=pos, q=q, ids=ids, pmass=0.31 / len(ids))
    st["phi"] = smooth_field(N, rng, 2e-5)
    st["chi"] = smooth_field(N, rng, 3e-6)
    st["Bi"] = smooth_field(N, rng, 1e-7, ncomp=3)
    # k-space state consistent with the unnormalised c2r: c2r(BiFT) = Bi
    st["BiFT"] = np.fft.rfftn(st["Bi"], axes=(-3, -2, -1)) / N ** 3
    st["scalarFT"] = np.fft.rfftn(st["phi"]) / N ** 3
    return st


def sim_params(N, boxsize, gr=True, vector_elliptic=False, z_in=100.0):
    """Settings of the reference's config for a box (Cf 48, steplimit 0.04)."""
    C = 2997.92458
    return dict(N=N, gr_flag=1 if gr else 0, vector_flag=1 if vector_elliptic else 0,
                Cf=48.0, steplimit=0.04, z_relax=z_in, fourpiG=1.5 * boxsize * boxsize / C / C,
                z_in=z_in)
