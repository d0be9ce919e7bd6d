"""Seeded synthetic inputs: coarse meshes and rough coefficient fields.

The coarse mesh is the reference's ``hyper_cube(0,1)`` refined ``g`` times
(source/Q/q_global.cc:92-102): nc = 2^g axis-aligned cells per direction, here
numbered lexicographically (x fastest).  Coefficients are sampled on the
global fine grid of nc*2^l cells per direction and cut into per-cell blocks
(fine cells lexicographic inside each coarse cell), the layout the C-ABI takes.
"""
from __future__ import annotations

import numpy as np


def coarse_vertices(dim: int, nc: int) -> np.ndarray:
    """(nc^dim, 2^dim, dim) vertex coordinates of the structured coarse mesh."""
    nv = 1 << dim
    h = 1.0 / nc
    cells = np.stack(np.meshgrid(*[np.arange(nc)] * dim, indexing="ij"), -1)
    cells = cells.transpose(*range(dim - 1, -1, -1), dim).reshape(-1, dim)  # x fastest
    out = np.empty((cells.shape[0], nv, dim))
    for v in range(nv):
        for d in range(dim):
            out[:, v, d] = (cells[:, d] + ((v >> d) & 1)) * h
    return out


def to_cell_blocks(field: np.ndarray, dim: int, nc: int, N: int) -> np.ndarray:
    """Cut a global fine-cell field (array indexed [z][y][x] or [y][x], with
    nc*N cells per direction) into (nc^dim, N^dim) per-coarse-cell blocks."""
    if dim == 2:
        f = field.reshape(nc, N, nc, N)             # [cy][fy][cx][fx]
        return np.ascontiguousarray(f.transpose(0, 2, 1, 3)).reshape(nc * nc, N * N)
    f = field.reshape(nc, N, nc, N, nc, N)          # [cz][fz][cy][fy][cx][fx]
    return np.ascontiguousarray(f.transpose(0, 2, 4, 1, 3, 5)).reshape(nc ** 3, N ** 3)


def from_cell_blocks(blocks: np.ndarray, dim: int, nc: int, N: int) -> np.ndarray:
    if dim == 2:
        return blocks.reshape(nc, nc, N, N).transpose(0, 2, 1, 3).reshape(nc * N, nc * N)
    return blocks.reshape(nc, nc, nc, N, N, N).transpose(0, 3, 1, 4, 2, 5).reshape(
        nc * N, nc * N, nc * N)


def gaussian_field(shape, sigma: float, corr_len_cells: float, seed: int) -> np.ndarray:
    """Stationary Gaussian field, mean 0, standard deviation ``sigma``, squared
    exponential covariance with correlation length ``corr_len_cells`` (in grid
    cells), sampled by FFT on the periodic grid and normalised to the target
    standard deviation.  corr_len_cells <= 0 gives iid N(0, sigma^2)."""
    rng = np.random.default_rng(seed)
    white = rng.standard_normal(shape)
    if corr_len_cells <= 0:
        return sigma * white
    freqs = np.meshgrid(*[np.fft.fftfreq(n) for n in shape], indexing="ij")
    k2 = sum(f * f for f in freqs)
    # spectral density of exp(-r^2 / (2 l^2)) up to a constant
    amp = np.exp(-(np.pi ** 2) * 2.0 * (corr_len_cells ** 2) * k2 / 1.0)
    g = np.real(np.fft.ifftn(np.fft.fftn(white) * np.sqrt(amp)))
    g -= g.mean()
    g *= sigma / g.std()
    return g


def lognormal_coefficient(dim: int, nc: int, n_refine: int, sigma: float, corr_len: float,
                          seed: int) -> np.ndarray:
    """A = exp(g) per fine cell, g Gaussian (sigma, correlation length in units
    of the domain); returns (nc^dim, N^dim) per-cell blocks."""
    N = 1 << n_refine
    n = nc * N
    g = gaussian_field((n,) * dim, sigma, corr_len * n if corr_len > 0 else 0.0, seed)
    return to_cell_blocks(np.exp(g), dim, nc, N)


def inclusion_coefficient(dim: int, nc: int, n_refine: int, n_inclusions: int, radius: float,
                          value: float, seed: int, background: float = 1.0) -> np.ndarray:
    """Background coefficient with seeded spherical inclusions of ``value``
    (fine cell centres inside a ball), as per-cell blocks."""
    rng = np.random.default_rng(seed)
    N = 1 << n_refine
    n = nc * N
    centres = rng.random((n_inclusions, dim))
    x = (np.arange(n) + 0.5) / n
    grids = np.meshgrid(*[x] * dim, indexing="ij")[::-1]  # grids[d] varies along axis d
    field = np.full((n,) * dim, background)
    for c in centres:
        r2 = sum((grids[d] - c[d]) ** 2 for d in range(dim))
        field[r2 <= radius * radius] = value
    return to_cell_blocks(field, dim, nc, N)


def gauss_points(dim: int) -> np.ndarray:
    """QGauss(2) points on [0,1]^dim, lexicographic (x fastest)."""
    g = np.array([0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0)])
    return np.array([[g[(q >> d) & 1] for d in range(dim)] for q in range(1 << dim)])


def qp_coordinates(vertices: np.ndarray, n_refine: int) -> np.ndarray:
    """Physical QGauss(2) points of every fine cell: (n_cells, N^dim, 2^dim, dim)."""
    n_cells, nv, dim = vertices.shape
    N = 1 << n_refine
    lo = vertices[:, 0, :]
    h = (vertices[:, nv - 1, :] - lo) / N
    fc = np.stack(np.meshgrid(*[np.arange(N)] * dim, indexing="ij"), -1)
    fc = fc.transpose(*range(dim - 1, -1, -1), dim).reshape(-1, dim)  # x fastest
    gp = gauss_points(dim)
    pts = (fc[None, :, None, :] + gp[None, None, :, :]) * h[:, None, None, :] + lo[:, None, None, :]
    return pts


def _euler_rotation(rotate=True) -> np.ndarray:
    """The fixed ZXZ Euler rotation of eqn_coeff_A.cc:25-41 (pi/3, pi/6, pi/4)."""
    a, b, g = np.pi / 3, np.pi / 6, np.pi / 4
    if not rotate:
        return np.eye(3)
    return np.array([
        [np.cos(a) * np.cos(g) - np.sin(a) * np.cos(b) * np.sin(g),
         -np.cos(a) * np.sin(g) - np.sin(a) * np.cos(b) * np.cos(g), np.sin(a) * np.sin(b)],
        [np.sin(a) * np.cos(g) + np.cos(a) * np.cos(b) * np.sin(g),
         -np.sin(a) * np.sin(g) + np.cos(a) * np.cos(b) * np.cos(g), -np.cos(a) * np.sin(b)],
        [np.sin(b) * np.sin(g), np.sin(b) * np.cos(g), np.cos(b)]])


def _sym6(A: np.ndarray) -> np.ndarray:
    return np.stack([A[..., 0, 0], A[..., 0, 1], A[..., 0, 2], A[..., 1, 1], A[..., 1, 2],
                     A[..., 2, 2]], -1)


def diffusion_A_tensor(points: np.ndarray, freq=(17, 16, 15), scale=(1.0, 1.0, 1.0),
                       alpha=(0.9, 0.9, 0.9), rotate=True) -> np.ndarray:
    """Diffusion_A of source/equation_data/eqn_coeff_A.cc:163-191 at 3-D points,
    returned as the symmetric 6-vector (xx,xy,xz,yy,yz,zz) per point."""
    R = _euler_rotation(rotate)
    D = np.stack([scale[d] * (1.0 - alpha[d] * np.sin(2 * np.pi * freq[d] * points[..., d]))
                  for d in range(3)], -1)
    return _sym6(np.einsum("ik,...k,jk->...ij", R, D, R))


def diffusion_inverse_A_tensor(points: np.ndarray, freq=(17, 16, 15), scale=(1.0, 1.0, 1.0),
                               alpha=(0.9, 0.9, 0.9), rotate=True) -> np.ndarray:
    """DiffusionInverse_A (eqn_coeff_A.cc:194-242): the rotated diagonal of
    reciprocals, as the symmetric 6-vector per point."""
    R = _euler_rotation(rotate)
    D = np.stack([1.0 / (scale[d] * (1.0 - alpha[d] * np.sin(2 * np.pi * freq[d] * points[..., d])))
                  for d in range(3)], -1)
    return _sym6(np.einsum("ik,...k,jk->...ij", R, D, R))


def diffusion_B(points: np.ndarray, freq=14, scale=1.0, alpha=0.9) -> np.ndarray:
    """Diffusion_B with the expression of the Ned_RT examples
    (example_parameters/prm_ned_rt_test-01.prm: scale (1 - alpha sin(2 pi f x)))."""
    return scale * (1.0 - alpha * np.sin(2 * np.pi * freq * points[..., 0]))


def rt_dq_rhs_3d(points: np.ndarray) -> np.ndarray:
    """The right-hand side of example_parameters/prm_rt_dq_test-01.prm,
    sin(2 pi x) sin(2 pi y) sin(2 pi z), per point."""
    return np.prod(np.sin(2 * np.pi * points), axis=-1)


def ned_rt_rhs(points: np.ndarray, scale=100.0) -> np.ndarray:
    """The right-hand side of prm_ned_rt_test-01.prm: scale times the gradient
    of x y z (x-1) (y-1) (z-1)."""
    x, y, z = points[..., 0], points[..., 1], points[..., 2]
    return scale * np.stack([(2 * x - 1) * (y * y - y) * (z * z - z),
                             (2 * y - 1) * (x * x - x) * (z * z - z),
                             (2 * z - 1) * (x * x - x) * (y * y - y)], -1)


def gauss_points_3d(verts: np.ndarray, l: int) -> np.ndarray:
    """Physical QGauss(2) points of every fine cell of every box: verts
    (n, 8, 3) in deal.II vertex order -> (n, 8^l, 8, 3), fine cells x fastest,
    points x fastest."""
    verts = np.asarray(verts, dtype=np.float64).reshape(-1, 8, 3)
    m = 1 << l
    g = np.array([0.5 - 0.5 / np.sqrt(3), 0.5 + 0.5 / np.sqrt(3)])
    c = np.arange(m ** 3)
    cell = np.stack([c % m, (c // m) % m, c // (m * m)], -1).astype(np.float64)   # (m^3, 3)
    q = np.arange(8)
    qp = np.stack([g[q & 1], g[(q >> 1) & 1], g[q >> 2]], -1)                    # (8, 3)
    ref = (cell[:, None, :] + qp[None, :, :]) / m                                   # (m^3, 8, 3)
    lo = verts[:, 0, :]
    H = np.stack([verts[:, 1, 0] - lo[:, 0], verts[:, 2, 1] - lo[:, 1], verts[:, 4, 2] - lo[:, 2]], -1)
    return lo[:, None, None, :] + ref[None] * H[:, None, None, :]


def box_grid_3d(n: int) -> np.ndarray:
    """The n^3 boxes of the unit cube (x fastest), vertices (n^3, 8, 3) in deal.II order."""
    h = 1.0 / n
    out = np.empty((n ** 3, 8, 3))
    for c in range(n ** 3):
        i, j, k = c % n, (c // n) % n, c // (n * n)
        for v in range(8):
            out[c, v] = [(i + (v & 1)) * h, (j + ((v >> 1) & 1)) * h, (k + ((v >> 2) & 1)) * h]
    return out


def inclusion_centres(dim: int, n_inclusions: int, seed: int) -> np.ndarray:
    """Seeded inclusion centres in the unit cube (the generator of
    inclusion_coefficient)."""
    return np.random.default_rng(seed).random((n_inclusions, dim))


def uneven_levels(dim: int, nc: int, centres: np.ndarray, seed: int, fine: int = 4,
                  coarse=(2, 3)) -> np.ndarray:
    """Refinement level per coarse cell: ``fine`` for the cells that hold an
    inclusion centre, otherwise one of ``coarse`` by a seeded coin."""
    rng = np.random.default_rng(seed)
    lv = rng.choice(np.asarray(coarse), size=nc ** dim)
    idx = np.minimum((centres * nc).astype(int), nc - 1)
    flat = idx[:, 0] + nc * idx[:, 1] + (nc * nc * idx[:, 2] if dim == 3 else 0)
    lv[np.unique(flat)] = fine
    return lv


def inclusion_blocks(dim: int, nc: int, levels: np.ndarray, centres: np.ndarray, radius: float,
                     value: float, background: float = 1.0) -> list:
    """Per-cell coefficient blocks at each cell's own refinement: ``value`` on
    the fine cells whose centre lies in a ball of ``radius`` around a centre,
    ``background`` elsewhere (lexicographic, x fastest)."""
    h = 1.0 / nc
    out = []
    r2 = radius * radius
    for c in range(nc ** dim):
        ix = [c % nc, (c // nc) % nc, c // (nc * nc)][:dim]
        N = 1 << int(levels[c])
        g = (np.arange(N) + 0.5) / N
        axes = [(ix[d] + g) * h for d in range(dim)]
        grids = np.meshgrid(*axes[::-1], indexing="ij")[::-1]   # grids[d] varies along axis d
        blk = np.full((N,) * dim, background)
        lo = np.array([ix[d] * h for d in range(dim)]) - radius
        hi = lo + h + 2 * radius
        near = np.all((centres >= lo) & (centres <= hi), axis=1)
        for cc in centres[near]:
            d2 = sum((grids[d] - cc[d]) ** 2 for d in range(dim))
            blk[d2 <= r2] = value
        out.append(blk.reshape(-1))
    return out
