"""Synthetic workloads: the uniform Stromgren sphere and the clumpy log-normal
medium, plus layout helpers between the global lattice and the reference's
subgrid-major cell order.

The reference builds its density on the fly (a uniform 1e8 m^-3 default,
DensitySubGrid.hpp:546, or a sum of ten random cosine modes,
testDensitySubGrid_omp.cpp:1881-1936); the log-normal Gaussian-random-field
medium below is this project's own workload generator.
"""
import numpy as np

PC = 3.086e16          # m (scripts/plot_neutral_fraction_profile.py:82)
SIGMA_H = 6.3e-22      # m^2 (testDensitySubGrid_omp.cpp:589)
Q_STROMGREN = 4.26e49  # s^-1 (testDensitySubGrid_omp.cpp:2011)
ALPHA_B = 2.7e-19      # m^3 s^-1, case B at 8000 K
ALPHA_REF = 4.0e-19    # m^3 s^-1, DensitySubGrid.hpp:1637
N_H = 1.0e8            # m^-3
XN0 = 1.0e-6           # DensitySubGrid.hpp:549


def box_10pc():
    """[-5 pc, 5 pc]^3 as anchor + sides (params.txt:3-5)."""
    return np.array([-5 * PC, -5 * PC, -5 * PC, 10 * PC, 10 * PC, 10 * PC])


def global_to_subgrid_major(a, nsub):
    """(nx, ny, nz) lattice -> flat subgrid-major reference order."""
    nx, ny, nz = a.shape
    sx, sy, sz = nx // nsub[0], ny // nsub[1], nz // nsub[2]
    b = a.reshape(nsub[0], sx, nsub[1], sy, nsub[2], sz)
    return np.ascontiguousarray(b.transpose(0, 2, 4, 1, 3, 5)).reshape(-1)


def subgrid_major_to_global(v, ncell, nsub):
    """Flat subgrid-major reference order -> (nx, ny, nz) lattice."""
    sx, sy, sz = (ncell[0] // nsub[0], ncell[1] // nsub[1],
                  ncell[2] // nsub[2])
    b = np.asarray(v).reshape(nsub[0], nsub[1], nsub[2], sx, sy, sz)
    return np.ascontiguousarray(b.transpose(0, 3, 1, 4, 2, 5)).reshape(
        ncell[0], ncell[1], ncell[2])


def cell_centres(box, ncell):
    """Cell-centre coordinates of the global lattice (3 arrays)."""
    ax = [box[a] + (np.arange(ncell[a]) + 0.5) * box[3 + a] / ncell[a]
          for a in range(3)]
    return np.meshgrid(*ax, indexing="ij")


def gaussian_random_field(n, slope=-11.0 / 3.0, seed=2, dtype=np.float64):
    """Zero-mean, unit-variance Gaussian random field on an n^3 periodic
    lattice with power spectrum P(k) ~ k^slope (k = 0 removed)."""
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((n, n, n))
    f = np.fft.rfftn(white)
    kx = np.fft.fftfreq(n)[:, None, None]
    ky = np.fft.fftfreq(n)[None, :, None]
    kz = np.fft.rfftfreq(n)[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz
    k2[0, 0, 0] = 1.0
    amp = k2 ** (slope / 4.0)   # sqrt(P) = k^(slope/2)
    amp[0, 0, 0] = 0.0
    g = np.fft.irfftn(f * amp, s=(n, n, n))
    g -= g.mean()
    g /= g.std()
    return g.astype(dtype)


def lognormal_density(n, sigma_ln=1.0, mean=N_H, seed=2):
    """n_H = mean * exp(sigma g - sigma^2 / 2): log-normal with the given
    mean (exactly, in expectation) and log-dispersion."""
    g = gaussian_random_field(n, seed=seed)
    return mean * np.exp(sigma_ln * g - 0.5 * sigma_ln * sigma_ln)


def random_sources(nsrc, box, seed=3, lum_seed=4, lmin=1e48, lmax=1e49):
    """Uniform-random positions in the box, log-uniform luminosities."""
    rp = np.random.default_rng(seed)
    rl = np.random.default_rng(lum_seed)
    pos = np.empty((nsrc, 3))
    for a in range(3):
        pos[:, a] = box[a] + rp.uniform(0.02, 0.98, nsrc) * box[3 + a]
    L = 10 ** rl.uniform(np.log10(lmin), np.log10(lmax), nsrc)
    return pos, L


def stromgren_radius(Q, n, alpha):
    """Rs = (3 Q / (4 pi n^2 alpha))^(1/3)
    (scripts/plot_neutral_fraction_profile.py:68)."""
    return (3.0 * Q / (4.0 * np.pi * n * n * alpha)) ** (1.0 / 3.0)


def ionized_radius(xn, box, ncell):
    """Volume-equivalent radius of the ionized region: V = sum (1 - x_n) dV,
    R = (3 V / (4 pi))^(1/3)."""
    dv = (box[3] / ncell[0]) * (box[4] / ncell[1]) * (box[5] / ncell[2])
    v = float(np.sum(1.0 - np.asarray(xn))) * dv
    return (3.0 * v / (4.0 * np.pi)) ** (1.0 / 3.0)


def analytic_profile(Q, n, alpha, sigma, r):
    """1-D neutral fraction of a Stromgren sphere with attenuation
    (restated from scripts/plot_neutral_fraction_profile.py:70-79): at radius
    r, A = Q sigma exp(-tau(r)) / (8 pi n alpha r^2), x = 1 + A -
    sqrt(2A + A^2), tau' = n sigma x."""
    r = np.asarray(r, dtype=np.float64)
    x = np.ones_like(r)
    tau = 0.0
    factor = 0.125 * Q * sigma / (np.pi * n * alpha)
    prev_r, prev_x = 0.0, 0.0
    for i, ri in enumerate(r):
        if ri <= 0:
            x[i] = 0.0
            continue
        # trapezoid in x from the previous radius (first step: x ~ 0)
        A = factor * np.exp(-tau) / ri ** 2
        xi = 1.0 + A - np.sqrt(2.0 * A + A * A)
        tau += 0.5 * n * sigma * (prev_x + xi) * (ri - prev_r)
        x[i] = xi
        prev_r, prev_x = ri, xi
    return x
