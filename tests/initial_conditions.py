"""Initial conditions of the advection workloads, evaluated at cell centres
exactly as the reference's init loops do (center = origin + length*(i+0.5),
cell.hpp:130-139, with origin 0 and length 2**-level on the unit box)."""
import numpy as np


def centers(cells):
    """cells: (n, >=4) rows (level, x, y, z, ...) -> (n, 3) float64 centres."""
    length = np.ldexp(1.0, -cells[:, 0].astype(np.int64))
    return 0.0 + length[:, None] * (cells[:, 1:4].astype(np.float64) + 0.5)


def disc_2d(cells, xc=0.3, yc=0.3, radius=0.2):
    """demos/FiniteVolume/advection_2d.cpp:22-45."""
    c = centers(cells)
    d = (c[:, 0] - xc) * (c[:, 0] - xc) + (c[:, 1] - yc) * (c[:, 1] - yc)
    return np.where(d <= radius * radius, 1.0, 0.0)


def box(cells, dim, lo=0.4, hi=0.6):
    """u = 1 inside [lo, hi]^dim (README.md:111-118 with the box of config 1)."""
    c = centers(cells)
    inside = np.ones(len(cells), dtype=bool)
    for d in range(dim):
        inside &= (c[:, d] >= lo) & (c[:, d] <= hi)
    return np.where(inside, 1.0, 0.0)


def sphere_3d(cells, center=(0.5, 0.5, 0.5), radius=0.15):
    """advection_3d-like sphere of the 3D workload."""
    c = centers(cells)
    d = ((c[:, 0] - center[0]) * (c[:, 0] - center[0]) + (c[:, 1] - center[1]) * (c[:, 1] - center[1])
         + (c[:, 2] - center[2]) * (c[:, 2] - center[2]))
    return np.where(d <= radius * radius, 1.0, 0.0)


# exp(x) from IEEE-754 +, -, * and exact scaling only, so that the seeded
# fields are bit-identical on every host (the libm and SIMD exp of numpy may
# differ in the last place between CPUs).  Cody-Waite reduction
# x = k ln2 + r, |r| <= ln2 / 2, with ln2 split so that k * LN2_HI is exact,
# then the Taylor polynomial of degree 16 in Horner form (truncation error
# below 1e-21 relative) and the scaling by 2^k.
_LN2_HI = 6.93147180369123816490e-01
_LN2_LO = 1.90821492927058770002e-10
_INV_LN2 = 1.44269504088896338700e+00
_TAYLOR = [1.0 / np.prod(np.arange(1, k + 1, dtype=np.float64)) for k in range(17)]


def det_exp(x):
    x = np.asarray(x, dtype=np.float64)
    xc = np.clip(x, -745.0, 709.0)
    k = np.rint(xc * _INV_LN2)
    r = (xc - k * _LN2_HI) - k * _LN2_LO
    p = np.full_like(r, _TAYLOR[16])
    for c in _TAYLOR[15::-1]:
        p = p * r + c
    return np.where(x < -745.0, 0.0, np.ldexp(p, k.astype(np.int64)))


def seeded_2d(cells, seed, n_boxes=3, n_gauss=2):
    """Seeded sum of indicator boxes and Gaussians (the large 2D workload).

    Draw order (numpy Generator PCG64(seed)): for each box, centre x, centre y,
    side; then for each Gaussian, centre x, centre y, sigma, amplitude.  The
    Gaussians use det_exp, so the field is the same bit for bit on every host.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    c = centers(cells)
    u = np.zeros(len(cells))
    for _ in range(n_boxes):
        cx, cy = rng.uniform(0.2, 0.8, size=2)
        side = rng.uniform(0.05, 0.2)
        inside = (np.abs(c[:, 0] - cx) <= 0.5 * side) & (np.abs(c[:, 1] - cy) <= 0.5 * side)
        u = u + np.where(inside, 1.0, 0.0)
    for _ in range(n_gauss):
        cx, cy = rng.uniform(0.2, 0.8, size=2)
        sigma = rng.uniform(0.02, 0.08)
        amp = rng.uniform(0.5, 1.0)
        r2 = (c[:, 0] - cx) * (c[:, 0] - cx) + (c[:, 1] - cy) * (c[:, 1] - cy)
        u = u + amp * det_exp(-r2 / (2.0 * sigma * sigma))
    return u


def batch_instance(seed, levels=(7, 10)):
    """Parameters of one instance of the batched 2D workload: max_level drawn in
    levels (inclusive), a unit velocity at an angle in [0, 2 pi), and 1 to 5
    shapes split into indicator boxes and Gaussians (drawn as in seeded_2d with
    the seed `seed + 1000003`).  Returns (max_level, (ax, ay), n_boxes, n_gauss)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    max_level = int(rng.integers(levels[0], levels[1] + 1))
    theta = rng.uniform(0.0, 2.0 * np.pi)
    k = int(rng.integers(1, 6))
    n_boxes = int(rng.integers(0, k + 1))
    return max_level, (float(np.cos(theta)), float(np.sin(theta))), n_boxes, k - n_boxes


def batch_field(cells, seed, n_boxes, n_gauss):
    return seeded_2d(cells, seed + 1000003, n_boxes, n_gauss)
