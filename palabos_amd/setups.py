"""Set-up helpers for the benchmark configurations: Taylor-Green fields,
channel inlet profiles, sphere voxelisation, Brinkman porosity fields built
from seeded Gaussian blobs, seeded porous media and uneven block tilings.
(The reference sets these up with its data initialisers and domain
functionals, dataProcessors/dataInitializerWrapper3D.h.)"""
import numpy as np


def taylor_green_velocity(n, u0, x0=0, x1=None):
    """u = u0 (sin kx cos ky cos kz, -cos kx sin ky cos kz, 0), k = 2 pi / n, on x slab [x0, x1)."""
    x1 = n if x1 is None else x1
    k = 2.0 * np.pi / n
    c = np.arange(n) * k
    s, co = np.sin(c), np.cos(c)
    xs = slice(x0, x1)
    u = np.zeros((x1 - x0, n, n, 3))
    u[..., 0] = u0 * s[xs, None, None] * co[None, :, None] * co[None, None, :]
    u[..., 1] = -u0 * co[xs, None, None] * s[None, :, None] * co[None, None, :]
    return u


def poiseuille_profile(ny, nz, u_max):
    """u_x on the inlet face interior (y = 1..ny-2, z = 1..nz-2), walls half-way
    at -0.5 and n - 0.5: 16 u_max y'(Hy - y') z'(Hz - z') / (Hy^2 Hz^2)."""
    y = np.arange(1, ny - 1, dtype=np.float64) + 0.5
    z = np.arange(1, nz - 1, dtype=np.float64) + 0.5
    py = y * (ny - y) / float(ny * ny)
    pz = z * (nz - z) / float(nz * nz)
    return 16.0 * u_max * py[:, None] * pz[None, :]


def sphere_mask(shape, center, radius):
    x, y, z = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    return (x - center[0]) ** 2 + (y - center[1]) ** 2 + (z - center[2]) ** 2 <= radius * radius


def brinkman_blobs(shape, n_blobs=32, sigma=(4.0, 10.0), amplitude=(2.0, 8.0), shift=3.0, seed=4):
    """alpha = sigmoid(sum_b A_b exp(-|x - c_b|^2 / (2 sigma_b^2)) - shift), centres,
    widths and amplitudes drawn from numpy's RandomState(seed)."""
    rng = np.random.RandomState(seed)
    centres = rng.rand(n_blobs, 3) * np.array(shape)
    sig = rng.uniform(sigma[0], sigma[1], size=n_blobs)
    amp = rng.uniform(amplitude[0], amplitude[1], size=n_blobs)
    x, y, z = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    s = np.zeros(shape)
    for b in range(n_blobs):
        r2 = (x - centres[b, 0]) ** 2 + (y - centres[b, 1]) ** 2 + (z - centres[b, 2]) ** 2
        s += amp[b] * np.exp(-r2 / (2 * sig[b] ** 2))
    return 1.0 / (1.0 + np.exp(-(s - shift)))


def random_spheres(shape, count, r_range, seed, periodic=(True, True, True), max_tries=200000):
    """`count` non-overlapping spheres (radii uniform in r_range) by seeded
    rejection sampling; returns (centres [count,3], radii [count])."""
    rng = np.random.RandomState(seed)
    shape = np.array(shape, dtype=np.float64)
    centres, radii = [], []
    tries = 0
    while len(centres) < count and tries < max_tries:
        tries += 1
        r = rng.uniform(r_range[0], r_range[1])
        c = rng.rand(3) * shape
        ok = True
        for c2, r2 in zip(centres, radii):
            d = np.abs(c - c2)
            for a in range(3):
                if periodic[a]:
                    d[a] = min(d[a], shape[a] - d[a])
            if (d * d).sum() < (r + r2) ** 2:
                ok = False
                break
        if ok:
            centres.append(c)
            radii.append(r)
    if len(centres) < count:
        raise RuntimeError(f"placed only {len(centres)} of {count} spheres")
    return np.array(centres), np.array(radii)


def spheres_mask(shape, centres, radii, scale=1.0, periodic=(True, True, True)):
    """Solid mask of spheres with radii * scale (periodic images included)."""
    mask = np.zeros(shape, dtype=bool)
    for c, r in zip(centres, radii):
        r = r * scale
        lo = np.floor(c - r).astype(int)
        hi = np.ceil(c + r).astype(int)
        xs = [np.arange(lo[a], hi[a] + 1) for a in range(3)]
        X, Y, Z = np.meshgrid(*xs, indexing="ij")
        inside = (X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2 <= r * r
        idx = [X[inside], Y[inside], Z[inside]]
        keep = np.ones(idx[0].shape, dtype=bool)
        for a in range(3):
            if periodic[a]:
                idx[a] = idx[a] % shape[a]
            else:
                keep &= (idx[a] >= 0) & (idx[a] < shape[a])
        mask[idx[0][keep], idx[1][keep], idx[2][keep]] = True
    return mask


def tile_blocks(shape, cuts):
    """Tiles a domain with boxes: cuts[a] lists the interior cut positions along
    axis a (a block ends at cut - 1). Returns inclusive (x0,x1,y0,y1,z0,z1) boxes."""
    edges = []
    for a in range(3):
        e = [0] + sorted(cuts[a]) + [shape[a]]
        edges.append([(e[k], e[k + 1] - 1) for k in range(len(e) - 1)])
    return [(x0, x1, y0, y1, z0, z1) for (x0, x1) in edges[0] for (y0, y1) in edges[1] for (z0, z1) in edges[2]]


def uneven_blocks(shape, counts, lo, hi, seed):
    """Tiling of the domain by counts[0] x counts[1] x counts[2] blocks whose
    edges are drawn in [lo, hi] (seeded rejection sampling, summing to the extent)."""
    rng = np.random.RandomState(seed)
    cuts = []
    for a in range(3):
        n, k = shape[a], counts[a]
        if not (k * lo <= n <= k * hi):
            raise ValueError(f"axis {a}: {k} pieces in [{lo}, {hi}] cannot sum to {n}")
        while True:
            lengths = list(rng.randint(lo, hi + 1, size=k - 1))
            last = n - sum(lengths)
            if lo <= last <= hi:
                lengths.append(last)
                break
        cuts.append([int(c) for c in np.cumsum(lengths)[:-1]])
    return tile_blocks(shape, cuts)
