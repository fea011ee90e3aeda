"""Synthetic interferometer datasets and source meshes.

The reference ships no data for this path (its ``dataset/data.npz`` is not in
the repository) and builds meshes upstream in autolens.  This module makes the
inputs the likelihood needs, with the reference's array conventions:

* grid: ``grid_radians_slim[m] = (y, x)`` in radians, slim order =
  ``argwhere(~mask)`` (``tests.py:279-302``), centred on the mask centre;
* mask: ``numba.py:50-86`` (unmasked where r <= radius);
* mapping: ``pix_indexes_for_sub_slim_index`` i8[M, B] with -1 padding and
  ``pix_weights_for_sub_slim_index`` f8[M, B] (``numba.py:355-360``);
* neighbours: i8[S, P] sorted and -1 padded, ``neighbors_sizes`` i8[S]
  (``tests.py:83-132``, ``numba.py:672-677``);
* noise: complex ``noise_map = sigma (1 + 1j)`` so that the real noise used by
  W-tilde equals the noise of both components in chi^2.

Five workloads are defined (seeds fixed: 0 for the first, ``c`` for the c-th):

1. 30x30 grid at 0.1", r = 1.5" (716 px); 2,000 uv uniform in a 1e5-lambda
   disc; 20x20 rectangular source, one-hot mapping; lambda = 1.
2. 100x100 grid at 0.05", r = 2.5" (7,860 px); 1e5 uv log-uniform in
   [1e3, 5e5]; 1,000-vertex Delaunay source in a 1" disc, Dirichlet(1,1,1)
   barycentric weights on a random triangle per pixel.
3. as 2 with 10,000 px and 1e6 uv (precision study).
4. as 2 with lambda log-uniform in [0.01, 100] (gradients).
5. 5,000 px, 1e6 uv, 64 samples with S ~ U{300..1500}, own mesh and
   mapping each, lambda log-uniform in [0.01, 100] (batched throughput).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

ARCSEC = np.pi / 648000.0


def mask_circular(shape_native, pixel_scale, radius):
    """True = masked.  Same rule as ``numba.py:82-86`` with centre (0, 0)."""
    cy = 0.5 * (shape_native[0] - 1)
    cx = 0.5 * (shape_native[1] - 1)
    ys, xs = np.indices(shape_native)
    return (radius * radius) < (np.square((ys - cy) * pixel_scale) + np.square((xs - cx) * pixel_scale))


def circular_grid(n_side, pixel_scale, radius, n_pix=None):
    """Returns (grid_radians_slim [M,2] (y,x), native_index_for_slim_index [M,2]).

    With ``n_pix`` the slim list is truncated to its first ``n_pix`` entries
    (the grid must hold at least that many unmasked pixels).
    """
    mask = mask_circular((n_side, n_side), pixel_scale, radius)
    native = np.ascontiguousarray(np.argwhere(~mask)).astype(np.int64)
    if n_pix is not None:
        if native.shape[0] < n_pix:
            raise ValueError(f"mask holds {native.shape[0]} pixels < {n_pix}")
        native = np.ascontiguousarray(native[:n_pix])
    d = pixel_scale * ARCSEC
    c = 0.5 * (n_side - 1)
    grid = np.empty((native.shape[0], 2))
    grid[:, 0] = (c - native[:, 0]) * d
    grid[:, 1] = (native[:, 1] - c) * d
    return grid, native


def grid_for_pixel_count(n_pix, pixel_scale):
    """Smallest centred circular mask at ``pixel_scale`` holding >= n_pix pixels, truncated to n_pix."""
    r = np.sqrt(n_pix / np.pi) * pixel_scale
    while True:
        n_side = int(np.ceil(2 * r / pixel_scale)) + 3
        grid, native = circular_grid(n_side, pixel_scale, r)
        if grid.shape[0] >= n_pix:
            return circular_grid(n_side, pixel_scale, r, n_pix)
        r *= 1.002


def uv_disc_uniform(rng, K, rmax):
    r = rmax * np.sqrt(rng.random(K))
    th = 2.0 * np.pi * rng.random(K)
    return np.ascontiguousarray(np.stack([r * np.cos(th), r * np.sin(th)], axis=1))


def uv_loguniform(rng, K, rmin, rmax):
    r = np.exp(rng.uniform(np.log(rmin), np.log(rmax), K))
    th = 2.0 * np.pi * rng.random(K)
    return np.ascontiguousarray(np.stack([r * np.cos(th), r * np.sin(th)], axis=1))


def visibilities(rng, K, sigma_lo, sigma_hi):
    """Complex N(0,1) data and noise_map = sigma (1+1j), sigma ~ U(lo, hi)."""
    data = rng.standard_normal(K) + 1j * rng.standard_normal(K)
    sigma = rng.uniform(sigma_lo, sigma_hi, K)
    return np.ascontiguousarray(data), np.ascontiguousarray(sigma * (1.0 + 1.0j))


def neighbors_from_edges(S, edges):
    """Sorted, -1 padded neighbour lists from an undirected edge list [E,2]."""
    nb = [set() for _ in range(S)]
    for a, b in edges:
        if a != b:
            nb[a].add(int(b))
            nb[b].add(int(a))
    P = max(1, max(len(x) for x in nb))
    out = np.full((S, P), -1, dtype=np.int64)
    sizes = np.zeros(S, dtype=np.int64)
    for i, x in enumerate(nb):
        v = sorted(x)
        out[i, : len(v)] = v
        sizes[i] = len(v)
    return out, sizes


def rectangular_source(n):
    """n x n rectangular source, 4-connected neighbours."""
    S = n * n
    edges = []
    for r in range(n):
        for c in range(n):
            i = r * n + c
            if c + 1 < n:
                edges.append((i, i + 1))
            if r + 1 < n:
                edges.append((i, i + n))
    return neighbors_from_edges(S, np.array(edges, dtype=np.int64))


@dataclass
class Mesh:
    vertices: np.ndarray  # [S,2]
    triangles: np.ndarray  # [T,3] int64
    neighbors: np.ndarray  # [S,P] int64
    neighbors_sizes: np.ndarray  # [S] int64


def delaunay_source(rng, S, radius_arcsec=1.0):
    """S vertices uniform in a disc, Delaunay triangulated (scipy.spatial)."""
    from scipy.spatial import Delaunay

    r = radius_arcsec * ARCSEC * np.sqrt(rng.random(S))
    th = 2.0 * np.pi * rng.random(S)
    v = np.stack([r * np.sin(th), r * np.cos(th)], axis=1)  # (y, x)
    tri = Delaunay(v).simplices.astype(np.int64)
    edges = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    nb, sizes = neighbors_from_edges(S, edges)
    return Mesh(np.ascontiguousarray(v), np.ascontiguousarray(tri), nb, sizes)


def barycentric_mapping(rng, M, triangles):
    """Each image pixel -> the 3 vertices of a random triangle, Dirichlet(1,1,1) weights."""
    t = rng.integers(0, triangles.shape[0], M)
    idx = np.ascontiguousarray(triangles[t]).astype(np.int64)
    w = np.ascontiguousarray(rng.dirichlet(np.ones(3), M))
    return idx, w


def one_hot_mapping(rng, M, S):
    idx = rng.integers(0, S, (M, 1)).astype(np.int64)
    return np.ascontiguousarray(idx), np.ones((M, 1))


def dirty_image_numpy(data, noise_map, uv_wavelengths, grid_radians_slim, k_block=8192):
    """Consistent dirty image sum_k (Re d cos phi + Im d sin phi) / sigma^2.

    phi_ik = 2 pi (x_i u_k0 + y_i u_k1) with the W-tilde pairing of
    ``numba.py:204``; with this dirty image the reference's log-likelihood is
    the exact Gaussian evidence of the 2K real data.  CPU version for small
    sizes; the library computes the same sum on the GPU.
    """
    M = grid_radians_slim.shape[0]
    a = data.real / noise_map.real**2
    b = data.imag / noise_map.real**2
    out = np.zeros(M)
    for k0 in range(0, data.size, k_block):
        uv = uv_wavelengths[k0 : k0 + k_block]
        ph = 2.0 * np.pi * (grid_radians_slim[:, 1:2] * uv[None, :, 0] + grid_radians_slim[:, 0:1] * uv[None, :, 1])
        out += np.cos(ph) @ a[k0 : k0 + k_block] + np.sin(ph) @ b[k0 : k0 + k_block]
    return out


@dataclass
class Sample:
    """One lens-model sample: mapping, source neighbour graph, coefficient."""

    pix_indexes_for_sub_slim_index: np.ndarray
    pix_weights_for_sub_slim_index: np.ndarray
    neighbors: np.ndarray
    neighbors_sizes: np.ndarray
    coefficient: float = 1.0
    mesh: Mesh | None = None

    @property
    def pixels(self) -> int:
        return int(self.neighbors_sizes.size)


@dataclass
class Dataset:
    name: str
    seed: int
    grid_radians_slim: np.ndarray
    native_index_for_slim_index: np.ndarray
    pixel_scale: float
    n_side: int
    uv_wavelengths: np.ndarray
    data: np.ndarray
    noise_map: np.ndarray
    samples: list = field(default_factory=list)
    dirty_image: np.ndarray | None = None
    grid_radians_2d: np.ndarray | None = None  # [N, N, 2] (y, x) radians, when the file convention gives one

    @property
    def M(self):
        return self.grid_radians_slim.shape[0]

    @property
    def K(self):
        return self.uv_wavelengths.shape[0]

    @property
    def noise_map_real(self):
        return np.ascontiguousarray(self.noise_map.real)

    def ensure_dirty_image(self):
        if self.dirty_image is None:
            self.dirty_image = dirty_image_numpy(self.data, self.noise_map, self.uv_wavelengths, self.grid_radians_slim)
        return self.dirty_image


def config1(seed=0, K=2000):
    rng = np.random.default_rng(seed)
    grid, native = circular_grid(30, 0.1, 1.5)
    uv = uv_disc_uniform(rng, K, 1e5)
    data, noise = visibilities(rng, K, 0.5, 1.5)
    nb, sizes = rectangular_source(20)
    idx, w = one_hot_mapping(rng, grid.shape[0], 400)
    ds = Dataset("cfg1", seed, grid, native, 0.1, 30, uv, data, noise)
    ds.samples.append(Sample(idx, w, nb, sizes, 1.0))
    return ds


def _delaunay_dataset(name, seed, n_pix, K, n_samples, s_range, lam_range, grid_side=None, radius=None):
    rng = np.random.default_rng(seed)
    if grid_side is not None:
        grid, native = circular_grid(grid_side, 0.05, radius)
        n_side = grid_side
    else:
        grid, native = grid_for_pixel_count(n_pix, 0.05)
        n_side = int(native.max()) + 1
    uv = uv_loguniform(rng, K, 1e3, 5e5)
    data, noise = visibilities(rng, K, 0.5, 2.0)
    ds = Dataset(name, seed, grid, native, 0.05, n_side, uv, data, noise)
    for _ in range(n_samples):
        S = int(rng.integers(s_range[0], s_range[1] + 1)) if s_range[0] != s_range[1] else s_range[0]
        mesh = delaunay_source(rng, S)
        idx, w = barycentric_mapping(rng, grid.shape[0], mesh.triangles)
        lam = 1.0 if lam_range is None else float(np.exp(rng.uniform(np.log(lam_range[0]), np.log(lam_range[1]))))
        ds.samples.append(Sample(idx, w, mesh.neighbors, mesh.neighbors_sizes, lam, mesh))
    return ds


def config2(seed=2, K=100_000):
    return _delaunay_dataset("cfg2", seed, None, K, 1, (1000, 1000), None, grid_side=100, radius=2.5)


def config3(seed=3, K=1_000_000):
    return _delaunay_dataset("cfg3", seed, 10_000, K, 1, (1000, 1000), None)


def config4(seed=4, K=100_000):
    return _delaunay_dataset("cfg4", seed, None, K, 1, (1000, 1000), (0.01, 100.0), grid_side=100, radius=2.5)


def config5(seed=5, K=1_000_000, n_samples=64):
    return _delaunay_dataset("cfg5", seed, 5000, K, n_samples, (300, 1500), (0.01, 100.0))


def small_delaunay(seed, n_pix, K, S, n_samples=1, lam_range=None, uv_max=5e5):
    """Reduced-size instance of the Delaunay generator for tests."""
    rng = np.random.default_rng(seed)
    grid, native = grid_for_pixel_count(n_pix, 0.05)
    uv = uv_loguniform(rng, K, 1e3, uv_max)
    data, noise = visibilities(rng, K, 0.5, 2.0)
    ds = Dataset(f"small{seed}", seed, grid, native, 0.05, int(native.max()) + 1, uv, data, noise)
    for _ in range(n_samples):
        s = int(rng.integers(S[0], S[1] + 1)) if isinstance(S, tuple) else S
        mesh = delaunay_source(rng, s)
        idx, w = barycentric_mapping(rng, grid.shape[0], mesh.triangles)
        lam = 1.0 if lam_range is None else float(np.exp(rng.uniform(np.log(lam_range[0]), np.log(lam_range[1]))))
        ds.samples.append(Sample(idx, w, mesh.neighbors, mesh.neighbors_sizes, lam, mesh))
    return ds


# ---------------------------------------------------------------------------
# The reference's dataset file (tests.py:396-452, DataLoaded): data.npz with the arrays below; the
# grid, mask and pixel-to-grid indexes are not stored but fixed by tests.py:155-306.
# ---------------------------------------------------------------------------
NPZ_KEYS = ("dirty_image", "data", "noise_map", "uv_wavelengths", "neighbors", "neighbors_sizes",
            "pix_indexes_for_sub_slim_index", "pix_weights_for_sub_slim_index")
NPZ_GRID_SIZE = 30  # tests.py:26 (AUTOJAX_GRID_SIZE default)
NPZ_PIXEL_SCALE = 0.2  # tests.py:158
NPZ_CORNER_ARCSEC = 9.9  # tests.py:296, the dataset's first grid coordinate


def reference_grid_radians_2d(n_side, pixel_scale, corner_arcsec=NPZ_CORNER_ARCSEC):
    """tests.py:291-302: g[I, J] = (-d I + c, d J - c) in radians, d = pixel_scale arcsec, c = corner."""
    d = pixel_scale * ARCSEC
    c = corner_arcsec * ARCSEC
    I, J = np.mgrid[:n_side, :n_side]
    g = np.empty((n_side, n_side, 2))
    g[:, :, 0] = -d * I + c
    g[:, :, 1] = d * J - c
    return g


def reference_mask(n_side, pixel_scale, radius, centre=(0.0, 0.0)):
    """original.py:8-95 mask_2d_circular_from on an n x n grid (True = masked): unmasked where the
    distance of the pixel centre from `centre`, sqrt(x^2 + y^2), is at most radius."""
    cy = (n_side - 1) / 2.0 - centre[0] / pixel_scale
    cx = (n_side - 1) / 2.0 + centre[1] / pixel_scale
    ys, xs = np.indices((n_side, n_side))
    r = np.sqrt(((xs - cx) * pixel_scale) ** 2 + ((ys - cy) * pixel_scale) ** 2)
    return ~(r <= radius)


def load_npz(path, grid_size=NPZ_GRID_SIZE, pixel_scale=NPZ_PIXEL_SCALE, coefficient=1.0):
    """A Dataset from a file holding the reference's data.npz arrays (NPZ_KEYS), with the grid the
    reference's harness builds for it: grid_size x grid_size at pixel_scale, circular mask of radius
    ((grid_size + 1) // 2) * pixel_scale (tests.py:210-212), slim order argwhere(~mask).  Loaded with
    allow_pickle=False.  Raises ValueError when a key is missing or the mapping's row count differs
    from the mask's pixel count (another grid_size / pixel_scale)."""
    with np.load(path, allow_pickle=False) as z:
        missing = [k for k in NPZ_KEYS if k not in z.files]
        if missing:
            raise ValueError(f"{path}: missing arrays {missing}")
        a = {k: np.ascontiguousarray(z[k]) for k in NPZ_KEYS}
    radius = ((grid_size + 1) // 2) * pixel_scale
    mask = reference_mask(grid_size, pixel_scale, radius)
    g2 = reference_grid_radians_2d(grid_size, pixel_scale)
    native = np.ascontiguousarray(np.argwhere(~mask)).astype(np.int64)
    grid = np.ascontiguousarray(g2[~mask])
    M = a["pix_indexes_for_sub_slim_index"].shape[0]
    if grid.shape[0] != M:
        raise ValueError(f"{path}: mapping has {M} image pixels, the {grid_size}x{grid_size} mask at "
                         f"{pixel_scale}\" has {grid.shape[0]}")
    sample = Sample(a["pix_indexes_for_sub_slim_index"].astype(np.int64),
                    a["pix_weights_for_sub_slim_index"].astype(np.float64), a["neighbors"].astype(np.int64),
                    a["neighbors_sizes"].astype(np.int64), float(coefficient))
    ds = Dataset("npz", -1, grid, native, pixel_scale, grid_size, a["uv_wavelengths"].astype(np.float64),
                 a["data"].astype(np.complex128), a["noise_map"].astype(np.complex128), [sample],
                 a["dirty_image"].astype(np.float64), g2)
    return ds


def save_npz(ds, path, sample=0):
    """Write a Dataset (its `sample`-th lens-model sample) with the reference's data.npz keys."""
    s = ds.samples[sample]
    np.savez(path, dirty_image=ds.ensure_dirty_image(), data=ds.data, noise_map=ds.noise_map,
             uv_wavelengths=ds.uv_wavelengths, neighbors=s.neighbors, neighbors_sizes=s.neighbors_sizes,
             pix_indexes_for_sub_slim_index=s.pix_indexes_for_sub_slim_index,
             pix_weights_for_sub_slim_index=s.pix_weights_for_sub_slim_index)


def reference_convention_dataset(seed=0, K=1024, S=64, grid_size=NPZ_GRID_SIZE, pixel_scale=NPZ_PIXEL_SCALE):
    """A small dataset on the reference harness's grid (reference_grid_radians_2d + reference_mask):
    uv uniform in a 1e5-lambda disc, N(0,1) data, sigma ~ U(0.5, 1.5), a Delaunay source of S vertices
    spanning the grid with barycentric mapping (B = 3)."""
    rng = np.random.default_rng(seed)
    radius = ((grid_size + 1) // 2) * pixel_scale
    mask = reference_mask(grid_size, pixel_scale, radius)
    g2 = reference_grid_radians_2d(grid_size, pixel_scale)
    native = np.ascontiguousarray(np.argwhere(~mask)).astype(np.int64)
    grid = np.ascontiguousarray(g2[~mask])
    uv = uv_disc_uniform(rng, K, 1e5)
    data, noise = visibilities(rng, K, 0.5, 1.5)
    mesh = delaunay_source(rng, S, radius_arcsec=radius)
    idx, w = barycentric_mapping(rng, grid.shape[0], mesh.triangles)
    ds = Dataset("refgrid", seed, grid, native, pixel_scale, grid_size, uv, data, noise,
                 [Sample(idx, w, mesh.neighbors, mesh.neighbors_sizes, 1.0, mesh)], None, g2)
    return ds
