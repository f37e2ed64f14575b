"""Synthetic workloads of the batch-SOM benchmark configurations.

  mixture_dataset      10x10 square / D=8 CPU case: a seeded mixture of 5 isotropic Gaussians.
  systematics_dataset  one row per HEALPix pixel (nested order) of synthetic systematics:
                       lognormal(0, 0.5) features, N(0, 1) features and sin/cos of the pixel-centre
                       colatitude, every column standardised (FP64) then cast to FP32.

The pixel-centre colatitude uses the closed form of the HEALPix nested scheme (Gorski et al. 2005,
ApJ 622:759, eqs. 4-5 for the polar caps and the equatorial zone), which is what the reference
evaluates with u = v = 0.5 in decontamination/hp/__init__.py:376-419.
"""
import numpy as np

# ring index (in units of nside) of the centre of each of the 12 base faces
_JRLL = np.array([2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4], dtype=np.int64)


def _compress_bits(v):
    """Keep the even bits of v (de-interleave a Morton code)."""
    v = v & 0x5555555555555555
    v = (v | (v >> 1)) & 0x3333333333333333
    v = (v | (v >> 2)) & 0x0F0F0F0F0F0F0F0F
    v = (v | (v >> 4)) & 0x00FF00FF00FF00FF
    v = (v | (v >> 8)) & 0x0000FFFF0000FFFF
    v = (v | (v >> 16)) & 0x00000000FFFFFFFF
    return v


def nest_pixel_z(nside, pixels):
    """cos(colatitude) of the centres of nested-scheme pixels."""
    pixels = np.asarray(pixels, dtype=np.int64)
    npface = nside * nside
    face = pixels // npface
    ipf = pixels & (npface - 1)
    ix = _compress_bits(ipf)
    iy = _compress_bits(ipf >> 1)
    jr = _JRLL[face] * nside - ix - iy - 1  # ring number, 1 .. 4 nside - 1
    z = np.empty(pixels.shape, dtype=np.float64)
    north = jr < nside
    south = jr > 3 * nside
    equ = ~(north | south)
    fact = 1.0 / (3.0 * nside * nside)
    z[north] = 1.0 - jr[north].astype(np.float64) ** 2 * fact
    nr = (4 * nside - jr[south]).astype(np.float64)
    z[south] = -(1.0 - nr ** 2 * fact)
    z[equ] = (2 * nside - jr[equ]).astype(np.float64) * (2.0 / (3.0 * nside))
    return z


def standardise(a):
    a = np.asarray(a, dtype=np.float64)
    mu = a.mean(axis=0)
    sd = a.std(axis=0)
    sd[sd == 0] = 1.0
    return (a - mu) / sd


def systematics_dataset(nside, n_lognormal, n_gauss, seed=0, block=1 << 22):
    """(12 nside^2, n_lognormal + n_gauss + 2) FP32 synthetic systematics, standardised per column.

    Columns: n_lognormal lognormal(mu=0, sigma=0.5), n_gauss N(0, 1), sin(theta), cos(theta) of the
    nested pixel centre. Generated block by block to bound host memory."""
    npix = 12 * nside * nside
    dim = n_lognormal + n_gauss + 2
    rng = np.random.default_rng(seed)
    out = np.empty((npix, dim), dtype=np.float32)
    # first pass: raw values in FP32 storage plus FP64 column moments
    s1 = np.zeros(dim)
    s2 = np.zeros(dim)
    for b0 in range(0, npix, block):
        b1 = min(npix, b0 + block)
        cols = np.empty((b1 - b0, dim), dtype=np.float64)
        cols[:, :n_lognormal] = rng.lognormal(0.0, 0.5, (b1 - b0, n_lognormal))
        cols[:, n_lognormal:n_lognormal + n_gauss] = rng.standard_normal((b1 - b0, n_gauss))
        z = nest_pixel_z(nside, np.arange(b0, b1, dtype=np.int64))
        cols[:, -2] = np.sqrt(np.maximum(0.0, 1.0 - z * z))
        cols[:, -1] = z
        s1 += cols.sum(0)
        s2 += (cols * cols).sum(0)
        out[b0:b1] = cols
    mu = s1 / npix
    sd = np.sqrt(np.maximum(s2 / npix - mu * mu, 0.0))
    sd[sd == 0] = 1.0
    for b0 in range(0, npix, block):
        b1 = min(npix, b0 + block)
        out[b0:b1] = ((out[b0:b1].astype(np.float64) - mu) / sd).astype(np.float32)
    return out


def mixture_dataset(n=100_000, dim=8, k=5, sigma=0.5, seed=0):
    """Seeded mixture of k isotropic Gaussians with means uniform in [-3, 3]^dim."""
    rng = np.random.default_rng(seed)
    means = rng.uniform(-3.0, 3.0, (k, dim))
    labels = rng.integers(0, k, n)
    x = means[labels] + sigma * rng.standard_normal((n, dim))
    return x.astype(np.float32)


def uniform_weights(mn, dim, low=-1.0, high=1.0, seed=0):
    return np.random.default_rng(seed).uniform(low, high, (mn, dim)).astype(np.float32)


def catalogue_rows(n, n_lognormal, n_gauss, seed=0, block=1 << 22):
    """(n, n_lognormal + n_gauss + 2) FP32 catalogue rows from the same feature distributions as
    systematics_dataset, at uniformly random sky positions (cos(theta) uniform in [-1, 1]), each column
    standardised with the moments of the distribution it is drawn from."""
    rng = np.random.default_rng(seed)
    dim = n_lognormal + n_gauss + 2
    out = np.empty((n, dim), dtype=np.float32)
    # analytic moments: lognormal(0, s): mean exp(s^2/2), var (exp(s^2)-1) exp(s^2); sin(theta) with
    # z uniform: mean pi/4, E[sin^2] = 2/3; z: mean 0, var 1/3
    s2 = 0.25
    ln_mu = np.exp(s2 / 2.0)
    ln_sd = np.sqrt((np.exp(s2) - 1.0) * np.exp(s2))
    sin_mu = np.pi / 4.0
    sin_sd = np.sqrt(2.0 / 3.0 - sin_mu ** 2)
    z_sd = np.sqrt(1.0 / 3.0)
    for b0 in range(0, n, block):
        b1 = min(n, b0 + block)
        k = b1 - b0
        out[b0:b1, :n_lognormal] = (rng.lognormal(0.0, 0.5, (k, n_lognormal)) - ln_mu) / ln_sd
        out[b0:b1, n_lognormal:n_lognormal + n_gauss] = rng.standard_normal((k, n_gauss))
        z = rng.uniform(-1.0, 1.0, k)
        out[b0:b1, -2] = (np.sqrt(1.0 - z * z) - sin_mu) / sin_sd
        out[b0:b1, -1] = z / z_sd
    return out
