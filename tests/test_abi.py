"""The C-ABI library loads and exports every symbol include/autojax_amd.h declares (CPU only)."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "autojax_amd.h"


def declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^(?:int|size_t|const char\*)\s+(aj_\w+)\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "aj_w_tilde_curvature_interferometer_from" in names
    assert "aj_log_likelihood_function_via_w_tilde_from" in names
    assert len(names) >= 20


def test_library_exports_all_declared_symbols():
    from autojax_amd import _lib

    lib = _lib.load()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared()) == set(_lib.SIGNATURES), set(declared()) ^ set(_lib.SIGNATURES)


def declared_arity():
    """Parameter count of every declared entry point, from the header text."""
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    out = {}
    for m in re.finditer(r"^(?:int|size_t|const char\*)\s+(aj_\w+)\(([^)]*)\)\s*;", text, re.M):
        params = m.group(2).strip()
        out[m.group(1)] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def test_ctypes_signatures_match_header_arity():
    """A short argtypes list makes ctypes pass the trailing arguments (the stream handle) as C int."""
    from autojax_amd import _lib

    arity = declared_arity()
    assert set(arity) == set(declared())
    bad = {n: (len(_lib.SIGNATURES[n][1]), a) for n, a in arity.items() if len(_lib.SIGNATURES[n][1]) != a}
    assert not bad, bad


def test_workspace_queries_without_gpu():
    from autojax_amd import _lib

    assert _lib.query("aj_w_tilde_workspace_bytes", 2000, 716, 0) > 716 * 8192 * 8
    assert _lib.query("aj_w_tilde_workspace_bytes", 2000, 716, 99) == 0
    # LU workspace: the factors (S^2 doubles) and the pivots (S int64), each 256-byte aligned
    assert _lib.query("aj_solve_workspace_bytes", 100) >= 800 * 100 + 8 * 100
    assert _lib.load().aj_abi_version() == 1


def test_bad_arguments_rejected_before_any_gpu_work():
    from autojax_amd import _lib

    with pytest.raises(ValueError):
        _lib.call("aj_w_tilde_curvature_interferometer_from", None, None, None, 10, 0, 0, None, None, 0, None)
    with pytest.raises(ValueError):
        _lib.call("aj_w_tilde_curvature_interferometer_from", None, None, None, 10, 5, 7, None, None, 0, None)


def test_new_entry_points_reject_bad_arguments_without_gpu():
    """Argument checks of the residue-mode tuning, the preload form and the mapper run before any HIP
    call: each returns AJ_EINVAL (ValueError) on the CPU."""
    from autojax_amd import _lib

    for variant, chunk in ((5, 0), (0, 1000), (0, 65536), (-2, 0)):
        with pytest.raises(ValueError):
            _lib.call("aj_w_tilde_crt_configure", variant, chunk)
    _lib.call("aj_w_tilde_crt_configure", -1, 0)  # defaults
    # preload table: corner larger than the grid, negative K
    with pytest.raises(ValueError):
        _lib.call("aj_w_tilde_curvature_preload_interferometer_from", None, None, 0, 31, 30, 1, 30, 30, 1, None, 0,
                  None)
    with pytest.raises(ValueError):
        _lib.call("aj_w_tilde_curvature_preload_interferometer_from", None, None, -1, 30, 30, 1, 30, 30, 1, None, 0,
                  None)
    with pytest.raises(ValueError):
        _lib.call("aj_curvature_matrix_via_w_tilde_curvature_preload_interferometer_from", 1, 60, 60, 1, 1, 1, 0, 3,
                  10, 1, 1, 1, None)
    # mapper: no samples; a sample of two vertices (host offsets); no pixels
    offs = np.array([0, 2], dtype=np.int64)
    with pytest.raises(ValueError):
        _lib.call("aj_delaunay_mesh_batched", 1, 1, offs.ctypes.data, 0, 32, 1, 1, 1, 1, 1, None)
    with pytest.raises(ValueError, match="vertices"):
        _lib.call("aj_delaunay_mesh_batched", 1, 1, offs.ctypes.data, 1, 32, 1, 1, 1, 1, 1, None)
    with pytest.raises(ValueError):
        _lib.call("aj_delaunay_locate_batched", 1, 1, 1, 1, 1, 1, 0, 0, 1, 1, None)
    # MFMA probe: unknown kind, no workgroups, too many iterations
    for args in ((2, 1, 1), (0, 0, 1), (1, 1, 1 << 19)):
        with pytest.raises(ValueError):
            _lib.call("aj_mfma_peak_probe", *args, 1, None)
    # CRT constants: host-only query, 11 or 12 moduli
    buf = [np.zeros(12) for _ in range(3)] + [np.zeros(4)]
    with pytest.raises(ValueError):
        _lib.call("aj_crt_constants", 10, *(b.ctypes.data for b in buf))
    _lib.call("aj_crt_constants", 11, *(b.ctypes.data for b in buf))
    assert buf[0][:11].tolist() == [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223]
    assert buf[3][3] == 60.0 * 2.0**30


def test_mapper_rejects_bad_vertex_counts_on_host():
    from autojax_amd import mapper

    with pytest.raises(ValueError, match="3..3072"):
        mapper.build_mappers([np.zeros((2, 2))], np.zeros((4, 2)))
    with pytest.raises(ValueError, match="3..3072"):
        mapper.build_mappers([np.zeros((3073, 2))], np.zeros((4, 2)))
    with pytest.raises(ValueError, match="no samples"):
        mapper.build_mappers([], np.zeros((4, 2)))


def test_drop_in_signatures_match_reference_names():
    """Parameter names and order of the drop-in equal those of src/autojax/numba.py and, for the
    preload form, src/autojax/original.py (listed here, since those modules cannot be imported without
    numba)."""
    import inspect

    from autojax_amd import api

    ref = {
        "w_tilde_curvature_interferometer_from": ["noise_map_real", "uv_wavelengths", "grid_radians_slim"],
        "w_tilde_data_interferometer_from": ["visibilities_real", "noise_map_real", "uv_wavelengths", "grid_radians_slim"],
        "mapping_matrix_from": ["pix_indexes_for_sub_slim_index", "pix_weights_for_sub_slim_index", "pixels"],
        "data_vector_from": ["mapping_matrix", "dirty_image"],
        "curvature_matrix_via_w_tilde_from": ["w_tilde", "mapping_matrix"],
        "sparse_mapping_matrix_transpose_matmul": ["matrix", "pix_indexes_for_sub_slim_index", "pix_weights_for_sub_slim_index", "pixels"],
        "curvature_matrix_via_w_wilde_sparse_mapping_matrix_from": ["w_tilde", "pix_indexes_for_sub_slim_index", "pix_weights_for_sub_slim_index", "pixels"],
        "constant_regularization_matrix_from": ["coefficient", "neighbors", "neighbors_sizes"],
        "reconstruction_positive_negative_from": ["data_vector", "curvature_reg_matrix"],
        "noise_normalization_complex_from": ["noise_map"],
        "log_likelihood_function_via_w_tilde_from": ["dirty_image", "data", "noise_map", "w_tilde", "pix_indexes_for_sub_slim_index", "pix_weights_for_sub_slim_index", "neighbors", "neighbors_sizes"],
        "w_compact_curvature_interferometer_from": ["grid_size", "noise_map_real", "uv_wavelengths", "pixel_scale"],
        "w_tilde_via_compact_from": ["w_compact", "native_index_for_slim_index"],
        "w_tilde_via_preload_from": ["w_tilde_preload", "native_index_for_slim_index"],
        "curvature_matrix_via_w_compact_sparse_mapping_matrix_direct_from": ["w_compact", "native_index_for_slim_index", "pix_indexes_for_sub_slim_index", "pix_weights_for_sub_slim_index", "pixels"],
        "curvature_matrix_via_w_compact_sparse_mapping_matrix_from": ["w_compact", "native_index_for_slim_index", "pix_indexes_for_sub_slim_index", "pix_weights_for_sub_slim_index", "pixels"],
        "log_likelihood_function_via_w_compact_from": ["dirty_image", "data", "noise_map", "w_compact", "native_index_for_slim_index", "pix_indexes_for_sub_slim_index", "pix_weights_for_sub_slim_index", "neighbors", "neighbors_sizes"],
        "mask_2d_centres_from": ["shape_native", "pixel_scales", "centre"],
        "mask_2d_circular_from": ["shape_native", "pixel_scales", "radius", "centre"],
        # original.py:230-235 and :600-607 (the preload interchange form)
        "w_tilde_curvature_preload_interferometer_from": ["noise_map_real", "uv_wavelengths", "shape_masked_pixels_2d", "grid_radians_2d"],
        "curvature_matrix_via_w_tilde_curvature_preload_interferometer_from": ["curvature_preload", "pix_indexes_for_sub_slim_index", "pix_size_for_sub_slim_index", "pix_weights_for_sub_slim_index", "native_index_for_slim_index", "pix_pixels"],
    }
    for name, params in ref.items():
        sig = inspect.signature(getattr(api, name))
        pos = [p.name for p in sig.parameters.values() if p.kind == p.POSITIONAL_OR_KEYWORD]
        assert pos == params, (name, pos)


def test_host_validation():
    import numpy as np

    from autojax_amd import api

    with pytest.raises(ValueError):
        api.validate_mapping(np.array([[0, 5]]), 5)
    api.validate_mapping(np.array([[1, 1]]), 5)  # repeats are accepted (assigned or summed as the reference)
    with pytest.raises(ValueError):
        api.validate_mapping(np.array([[1, 1]]), 5, unique=True)
    api.validate_mapping(np.array([[1, -1, 1]]), 5, unique=True)  # repeat after the -1 is ignored
    with pytest.raises(ValueError):
        api.validate_neighbors(np.array([[1], [-1]]), np.array([1, 0]), symmetric=True)
    api.validate_neighbors(np.array([[1], [-1]]), np.array([1, 0]))  # one-sided lists are the reference's input too
    api.validate_neighbors(np.array([[1], [0]]), np.array([1, 1]), symmetric=True)
    with pytest.raises(ValueError):
        api.validate_neighbors(np.array([[2], [0]]), np.array([1, 1]))  # out of range


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus 2 exits with status 2 and says so, before touching a device."""
    import subprocess
    import sys

    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_backward_workspace_holds_per_sample_vectors_when_sources_outnumber_pixels():
    """The backward workspace (host-only query) covers the inverse blocks (sum S^2), the per-source-pixel
    column norms (sum S) and two per-image-pixel vectors per sample (2 M n), each 256-byte aligned, also
    when S exceeds M."""
    import ctypes

    from autojax_amd import _lib

    def align(x):
        return (x + 255) // 256 * 256

    for M, S in ((64, [200]), (400, [600, 30]), (5000, [300, 1500, 1041])):
        Sa = np.array(S, np.int64)
        got = _lib.query("aj_log_evidence_backward_workspace_bytes", M, len(S), Sa.ctypes.data_as(ctypes.c_void_p))
        need = align(8 * int((Sa**2).sum())) + align(8 * int(Sa.sum())) + 2 * align(8 * M * len(S))
        assert got >= need, (M, S, got, need)


def test_assignment_weights_and_neighbour_symmetry():
    """Host logic of the drop-in likelihood: a row naming a source pixel twice keeps only the last entry's
    weight (numba.py:375 assigns), entries after a row's -1 are ignored; symmetric and one-sided neighbour
    lists are told apart."""
    from autojax_amd import api

    idx = np.array([[3, 1, 3], [2, 2, 2], [0, -1, 0], [4, 5, 6]], np.int64)
    w = np.array([[0.1, 0.2, 0.3], [0.4, 0.5, 0.6], [0.7, 0.8, 0.9], [1.0, 1.1, 1.2]])
    assert api.has_repeated_indices(idx)
    got = api.assignment_weights(idx, w)
    np.testing.assert_array_equal(got, [[0.0, 0.2, 0.3], [0.0, 0.0, 0.6], [0.7, 0.8, 0.9], [1.0, 1.1, 1.2]])
    # summing the kept weights reproduces the dense assignment of every row
    dense = np.zeros((4, 7))
    for m in range(4):
        for b in range(3):
            if idx[m, b] == -1:
                break
            dense[m, idx[m, b]] = w[m, b]
    summed = np.zeros((4, 7))
    for m in range(4):
        for b in range(3):
            if idx[m, b] == -1:
                break
            summed[m, idx[m, b]] += got[m, b]
    np.testing.assert_array_equal(summed, dense)
    plain = np.array([[0, 1], [2, -1]], np.int64)
    assert not api.has_repeated_indices(plain)
    assert api.assignment_weights(plain, w[:2, :2]) is w[:2, :2] or np.array_equal(api.assignment_weights(plain, w[:2, :2]), w[:2, :2])
    nb = np.array([[1, 2], [0, -1], [0, -1]], np.int64)
    sz = np.array([2, 1, 1], np.int64)
    assert api.neighbors_symmetric(nb, sz)
    one_sided = np.array([[1, 2], [0, -1], [-1, -1]], np.int64)
    assert not api.neighbors_symmetric(one_sided, np.array([2, 1, 0], np.int64))
    with pytest.raises(ValueError, match="symmetric"):
        api.validate_neighbors(one_sided, np.array([2, 1, 0], np.int64), symmetric=True)
    api.validate_neighbors(one_sided, np.array([2, 1, 0], np.int64))
