"""CPU oracle for the interferometer-inversion likelihood path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library and the
``autojax_amd`` package) imports, links or calls anything in this folder.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use it, and only as the checker or as the timed CPU baseline.

Contents
--------
``restated``
    FP64 numpy restatement of the functions of ``src/autojax/numba.py`` on
    the likelihood path (same names, same parameter names and order), plus the
    two algebraically equal log-likelihood forms of ``original.py`` and
    ``jax.py``.  Every function cites the reference lines it follows.
``evidence``
    Independent checks that pin the restatement: brute-force Gaussian
    evidence, the closed-form spectrum of the rectangular-mesh regulariser,
    the deflated log-determinant of the regulariser, a dense FP64 torch
    restatement used for autograd cross-checks.
``wtilde_cpu.c``
    C/OpenMP restatement of ``numba.py:197-207`` (the W-tilde loop), used as
    the timed CPU baseline and for sub-sampled checks at full grid sizes.

Parity status (what pins each function; tests/test_oracle.py)
-------------------------------------------------------------------
w_tilde_curvature_interferometer_from   brute-force fsum triple loop, closed
                                        forms (diag = sum w, K=1 rank-2),
                                        translation invariance, compact-lattice
                                        formulation (numba.py:210-255)
mapping_matrix_from                     hand-built small cases, -1 break rule
data_vector_from                        brute-force loop
curvature_matrix_via_w_tilde_from       one-hot block sums, sparse two-stage
                                        formulation (numba.py:552-614)
constant_regularization_matrix_from     closed-form grid spectrum, zero row
                                        sums, original.py:703-743 loop form
reconstruction_positive_negative_from   LU vs Cholesky vs residual check
noise_normalization_complex_from        original.py:781-798 form, fsum
log_likelihood_function_via_w_tilde_from brute-force Gaussian evidence of the
                                        2K-dimensional data (tiny sizes), the
                                        original.py and jax.py forms
w_tilde_data_interferometer_from        brute-force fsum loop
"""
