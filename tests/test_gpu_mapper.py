"""GPU source-mesh mapper (Delaunay mesh, neighbour lists, barycentric mapping) against the scipy-based
oracle (oracle/mapper.py), and the batched log-evidence over its output."""

import numpy as np
import pytest

from autojax_amd import datasets as ds
from oracle import mapper as MP
from oracle import restated as R

pytestmark = pytest.mark.gpu


def _tri_set(T, canon=None):
    T = np.asarray(T)
    if canon is not None:
        T = canon[T]
    return sorted(tuple(sorted(int(v) for v in t)) for t in T)


def test_mapper_random_samples_match_scipy(gpu):
    """Eight samples from 3 to 3,072 vertices (source-plane scale, 1e-5 rad), 2,000 shared pixel positions
    reaching past every hull: the triangle sets equal scipy's, the neighbour lists are equal, and the
    dense mappings agree to 1e-12 (pixels on a shared edge may name either triangle)."""
    from autojax_amd import mapper

    rng = np.random.default_rng(50)
    sizes = [3, 4, 57, 300, 777, 1500, 3072, 1000]
    verts = [rng.random((n, 2)) * 2e-5 - 1e-5 for n in sizes]
    pos = rng.uniform(-1.2e-5, 1.2e-5, (2000, 2))
    res = mapper.build_mappers(verts, pos)
    for s, V in enumerate(verts):
        got = res.sample(s)
        T_ref = MP.delaunay_triangles(V)
        print(f"sample {s}: {len(V)} vertices, {len(got['triangles'])} triangles (scipy {len(T_ref)})", flush=True)
        assert _tri_set(got["triangles"]) == _tri_set(T_ref)
        nb_ref, sz_ref = MP.neighbors_from_triangles(len(V), T_ref)
        np.testing.assert_array_equal(got["neighbors_sizes"], sz_ref)
        np.testing.assert_array_equal(got["neighbors"], nb_ref)
        idx_ref, w_ref = MP.barycentric_mapping(V, T_ref, pos)
        Tg = MP.mapping_dense(got["pix_indexes_for_sub_slim_index"], got["pix_weights_for_sub_slim_index"], len(V))
        Tr = MP.mapping_dense(idx_ref, w_ref, len(V))
        np.testing.assert_allclose(Tg, Tr, rtol=0, atol=1e-12)


def test_mapper_per_sample_positions_and_degenerate_inputs(gpu):
    """Per-sample positions; a regular 20 x 20 integer lattice (cocircular quadruples, the predicates exact:
    any valid triangulation, checked by its empty circumcircles and the Euler count 2n - h - 2); repeated
    vertices (left out,
    the rest equal to scipy's with each repeat named by its first occurrence); collinear vertices
    (rejected)."""
    from scipy.spatial import ConvexHull

    from autojax_amd import mapper

    rng = np.random.default_rng(51)
    lat = np.stack(np.meshgrid(np.arange(20.0), np.arange(20.0), indexing="ij"), -1).reshape(-1, 2)
    base = rng.random((200, 2)) * 1e-5
    dup = np.concatenate([base, base[[3, 17, 150]]])
    pos = [rng.uniform(-1.0, 20.0, (500, 2)), rng.uniform(-1e-6, 1.1e-5, (500, 2))]
    res = mapper.build_mappers([lat, dup], pos)
    g = res.sample(0)
    T = g["triangles"]
    # h counts every vertex on the hull's boundary, the collinear ones included
    eq = ConvexHull(lat).equations
    h = int(np.any(np.abs(lat @ eq[:, :2].T + eq[:, 2]) < 1e-9, axis=1).sum())
    assert h == 76 and len(T) == 2 * len(lat) - h - 2
    for t in T:
        a, b, c = lat[t]
        ax, ay, bx, by, cx, cy = a[1], a[0], b[1], b[0], c[1], c[0]
        d = 2 * (ax * (by - cy) + bx * (cy - ay) + cx * (ay - by))
        ux = ((ax**2 + ay**2) * (by - cy) + (bx**2 + by**2) * (cy - ay) + (cx**2 + cy**2) * (ay - by)) / d
        uy = ((ax**2 + ay**2) * (cx - bx) + (bx**2 + by**2) * (ax - cx) + (cx**2 + cy**2) * (bx - ax)) / d
        r2 = (ax - ux) ** 2 + (ay - uy) ** 2
        assert np.all((lat[:, 1] - ux) ** 2 + (lat[:, 0] - uy) ** 2 >= r2 * (1 - 1e-9))
    idx0, w0 = MP.barycentric_mapping(lat, T, pos[0])
    np.testing.assert_allclose(
        MP.mapping_dense(g["pix_indexes_for_sub_slim_index"], g["pix_weights_for_sub_slim_index"], len(lat)),
        MP.mapping_dense(idx0, w0, len(lat)), rtol=0, atol=1e-12)
    d = res.sample(1)
    canon = np.arange(len(dup))
    canon[200:] = [3, 17, 150]
    assert _tri_set(d["triangles"]) == _tri_set(MP.delaunay_triangles(dup), canon)
    assert np.all(d["neighbors_sizes"][200:] == 0)
    with pytest.raises(ValueError, match="line"):
        t = np.arange(30.0)
        mapper.build_mappers([np.stack([t, 2 * t], 1)], pos[0])


def test_batched_evidence_over_device_mapper(gpu):
    """Six lens-model samples mapped on the device feed the batched log-evidence without a host copy of
    the mapping; each sample's log-evidence equals the oracle's with the oracle's (scipy) mapping."""
    import torch

    from autojax_amd import evidence as E
    from autojax_amd import mapper
    from oracle import evidence as ev

    from autojax_amd import api

    rng = np.random.default_rng(52)
    d = ds.small_delaunay(52, 600, 3000, 200, n_samples=1)
    grid = d.grid_radians_slim
    r = np.abs(grid).max()
    verts = [rng.uniform(-r, r, (n, 2)) for n in (120, 250, 400, 90, 333, 512)]
    lams = 10.0 ** rng.uniform(-2, 2, len(verts))
    res = mapper.build_mappers(verts, grid)
    batch = E.SampleBatch.from_mappers(res, lams)
    W = api.w_tilde_curvature_interferometer_from(d.noise_map_real, d.uv_wavelengths, grid, precision="fp64")
    y = d.ensure_dirty_image()
    out = batch.log_evidence(torch.from_numpy(W).cuda(), torch.from_numpy(y).cuda(),
                             E.chi2_noise_norm(d.data, d.noise_map), check=True).cpu().numpy()
    chi = d.data.view(np.float64) / d.noise_map.view(np.float64)
    nn = R.noise_normalization_complex_from(d.noise_map)
    for b, V in enumerate(verts):
        T_ref = MP.delaunay_triangles(V)
        idx, w = MP.barycentric_mapping(V, T_ref, grid)
        nb, sz = MP.neighbors_from_triangles(len(V), T_ref)
        To = R.mapping_matrix_from(idx, w, len(V))
        F = R.curvature_matrix_via_w_tilde_from(W, To)
        D = R.data_vector_from(To, y)
        H = R.constant_regularization_matrix_from(lams[b], nb, sz)
        ldh = ev.logdet_deflated(H, lams[b] ** 2 * sz.mean())
        want = R.log_evidence_from_parts(F, D, chi @ chi, nn, nb, sz, lams[b], logdet_h=ldh)[0]
        print(f"sample {b}: S={len(V)} lam={lams[b]:.3g} {out[b]:.12g} vs {want:.12g}", flush=True)
        assert abs(out[b] - want) < 1e-9 * max(1.0, abs(want))


def test_mapper_reports_degree_above_max_neighbors(gpu):
    """A wheel (a centre vertex inside a ring of 40) gives the centre 40 neighbours: above max_neighbors = 32
    the sample fails with a ValueError naming the limit; with max_neighbors = 48 it succeeds and the
    centre's neighbour list is the whole ring (as scipy's triangulation gives)."""
    from autojax_amd import mapper

    th = 2.0 * np.pi * np.arange(40) / 40
    wheel = np.concatenate([[[0.0, 0.0]], np.stack([np.sin(th), np.cos(th)], 1)]) * 1e-5
    pos = np.zeros((10, 2))
    with pytest.raises(ValueError, match="max_neighbors"):
        mapper.build_mappers([wheel], pos)
    res = mapper.build_mappers([wheel], pos, max_neighbors=48)
    got = res.sample(0)
    nb_ref, sz_ref = MP.neighbors_from_triangles(len(wheel), MP.delaunay_triangles(wheel))
    assert got["neighbors_sizes"][0] == 40 == sz_ref[0]
    np.testing.assert_array_equal(got["neighbors"][0, :40], np.arange(1, 41))
