/* C ABI of libautojax_amd: MI355X (gfx950) interferometer-inversion likelihood.
 *
 * Every pointer argument is DEVICE memory owned by the caller unless the
 * parameter name ends in _host.  Sizes are int64_t.  `stream` is a
 * hipStream_t (NULL = the default stream).  Functions return AJ_OK (0) or an
 * error status; aj_last_error() gives the text of the last error of the
 * calling thread.  Functions that need scratch memory take a caller-provided
 * `workspace` of at least the size returned by the matching *_workspace_bytes
 * query; the library allocates nothing.
 *
 * Each entry point cites the function of the reference (ickc/python-autojax,
 * src/autojax/numba.py) whose arguments and result it reproduces.  Array
 * layouts are the reference's C-contiguous ones: f8[M,2] grid (y, x),
 * f8[K,2] uv, complex128 arrays as interleaved (re, im) doubles, i8[M,B]
 * mapping indices with -1 terminating a row, i8[S,P] neighbour lists with -1
 * terminating a row.
 */
#ifndef AUTOJAX_AMD_H
#define AUTOJAX_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AJ_OK 0
#define AJ_EINVAL 1  /* bad argument (size, null pointer, precision) */
#define AJ_ENOTPD 2  /* matrix not positive definite (see info output) */
#define AJ_EHIP 3    /* HIP runtime error */
#define AJ_ESPACE 4  /* workspace too small */

/* W-tilde arithmetic modes. */
#define AJ_PREC_FP64 0   /* FP64 features, FP64 MFMA, FP64 accumulation */
#define AJ_PREC_FP32 1   /* FP32 features and MFMA, FP32 accumulation per 4096-visibility chunk */
#define AJ_PREC_BF16X3 2 /* split-BF16 features, 3 products, FP32 accumulation per chunk */
#define AJ_PREC_INT8X3 3 /* 3 signed 8-bit digits per feature, exact int32 accumulation */
#define AJ_PREC_INT8X4 4 /* 4 signed 8-bit digits per feature, exact int32 accumulation */
#define AJ_PREC_INT8X5 5 /* 5 signed 8-bit digits per feature, exact int32 accumulation */
#define AJ_PREC_INT8CRT 6 /* features on the 5-digit fixed-point grid, exact integer Gram per chunk of
                             visibilities from its residues modulo 12 moduli <= 256 (one int8 product
                             each), recombined in FP64 (Chinese remainder theorem) */
#define AJ_PREC_INT8CRT11 7 /* as AJ_PREC_INT8CRT on a grid of 60 * 2^30 (36 bits) with 11 moduli:
                               11 int8 products per FP64-equivalent product, chunks up to 16384 */

const char* aj_last_error(void);
int aj_abi_version(void);

/* Kernel timing (instrumentation, no reference counterpart).  When enabled,
 * HIP events are recorded around the launches named by the timer ids below,
 * on the caller's stream; aj_timing_read synchronises on them and returns the
 * summed duration and the launch count since the last aj_timing_enable. */
#define AJ_TIMER_WT_SYRK 0     /* W-tilde MFMA SYRK launches */
#define AJ_TIMER_WT_FEATURES 1 /* W-tilde feature launches */
#define AJ_TIMER_OMEGA 2       /* Omega = W T gather (batched) */
#define AJ_TIMER_FGATHER 3     /* F = T^T Omega gather (batched) */
#define AJ_TIMER_CHOL 4        /* batched Cholesky of A and H */
#define AJ_TIMER_SOLVE 5       /* triangular solves + log-evidence */
#define AJ_TIMER_WT_COMBINE 6  /* W-tilde residue recombination (AJ_PREC_INT8CRT) */
#define AJ_TIMER_BATCH 7       /* a whole batched evaluation, first to last launch */
#define AJ_TIMER_COUNT 8
int aj_timing_enable(int on);
int aj_timing_read(int timer, double* total_ms, int64_t* count);

/* Device characterisation (no reference counterpart): register-only MFMA loops giving the matrix
 * units' achievable rate on this device.  kind 0: v_mfma_i32_16x16x64_i8 on 32 accumulators per wave
 * (out: int64 per thread); kind 1: v_mfma_f64_16x16x4f64 on 16 accumulators per wave (out: double per
 * thread).  Operands are small integers, so out for iters iterations is exactly iters times out for one.
 * workgroups of 512 threads, iters <= 2^18; MFMAs per launch: workgroups * 8 * iters * (32 or 16). */
int aj_mfma_peak_probe(int kind, int64_t workgroups, int64_t iters, void* out, void* stream);

/* numba.py:155-207 w_tilde_curvature_interferometer_from(noise_map_real f8[K],
 * uv_wavelengths f8[K,2], grid_radians_slim f8[M,2]) -> f8[M,M].
 * Writes the full symmetric matrix w_tilde[M*M]. */
size_t aj_w_tilde_workspace_bytes(int64_t K, int64_t M, int precision);
int aj_w_tilde_curvature_interferometer_from(const double* noise_map_real, const double* uv_wavelengths,
                                             const double* grid_radians_slim, int64_t K, int64_t M, int precision,
                                             double* w_tilde, void* workspace, size_t workspace_bytes, void* stream);


/* Tuning of the W-tilde build, process-wide, no reference counterpart.
 * aj_w_tilde_set_variant: kernel variant of the integer-mode SYRK, 0 .. count-1, or -1 (default)
 *   for the fastest measured variant of each digit count.
 * aj_w_tilde_set_group_bytes: cap on the feature matrix of one SYRK launch (0 = default 256 MiB);
 * changes aj_w_tilde_workspace_bytes, so query the workspace after setting it. */
int aj_w_tilde_set_variant(int variant);
int aj_w_tilde_variant_count(void);
int aj_w_tilde_set_group_bytes(size_t bytes);

/* Constants of the AJ_PREC_INT8CRT (nmod = 12) or AJ_PREC_INT8CRT11 (nmod = 11) recombination, as the
 * kernels use them (host-only query, no reference counterpart): moduli[nmod]; the CRT basis
 * y_i = (P/m_i) ((P/m_i)^-1 mod m_i) mod P split as y_hi[i] (a multiple of 2^53) + y_lo[i] (an integer
 * below 2^53); pq = {P_hi, P_lo, 1/P, Q} with Q the largest feature integer of the mode. */
int aj_crt_constants(int nmod, double* moduli, double* y_hi, double* y_lo, double* pq);
/* The persistent residue SYRK's schedule (host-only query for checks, no reference counterpart): for
 * ntiles tiles, nm moduli, `pairs` stage pairs per (tile, modulus) unit and G workgroups (a multiple of
 * 8, at most 256), the segments each workgroup runs (segs: 8 ints each, {tile, modulus, first stage,
 * stages, slot or -1, 0, 0, 0}; beg: G + 1 offsets), the units split between workgroups (fix: 4 ints
 * each, {tile, modulus, first slot, slots}) and counts = {segments, split units, slots}. */
int aj_crt_sk_schedule(int ntiles, int nm, int pairs, int G, int* segs, int64_t seg_cap, int* beg, int* fix,
                       int64_t fix_cap, int64_t* counts);
/* Tuning of AJ_PREC_INT8CRT and AJ_PREC_INT8CRT11 (process-wide): residue SYRK kernel variant 0..3
 * (-1 = default 3, the persistent form with one workgroup per CU; all compute identical residues) and visibilities per exact integer chunk (a
 * multiple of 4096 up to 32768, capped at 16384 with 11 moduli; 0 = default 16384).  The chunk size
 * changes aj_w_tilde_workspace_bytes for these modes. */
int aj_w_tilde_crt_configure(int variant, int64_t chunk_visibilities);

/* numba.py:210-234 w_compact_curvature_interferometer_from(grid_size, noise_map_real f8[K],
 * uv_wavelengths f8[K,2], pixel_scale [arcsec]) -> f8[N, 2N-1], N = grid_size: the distinct values
 * of W-tilde on a regular N x N pixel grid, w_compact[m, n] = sum_k n_k^-2 cos(2 pi d ((n - N + 1) u_k0
 * - m u_k1)), d = pixel_scale in radians (FP64 MFMA over per-visibility phasors). */
size_t aj_w_compact_workspace_bytes(int64_t K, int64_t grid_size);
int aj_w_compact_curvature_interferometer_from(const double* noise_map_real, const double* uv_wavelengths, int64_t K,
                                               int64_t grid_size, double pixel_scale, double* w_compact,
                                               void* workspace, size_t workspace_bytes, void* stream);

/* numba.py:237-255 w_tilde_via_compact_from(w_compact f8[N,2N-1], native_index_for_slim_index i8[M,2])
 * -> f8[M,M] (native indices are the (row, column) grid position of each image pixel, in [0, N)). */
int aj_w_tilde_via_compact_from(const double* w_compact, int64_t grid_size,
                                const int64_t* native_index_for_slim_index, int64_t M, double* w_tilde, void* stream);

/* numba.py:259-291 w_tilde_via_preload_from(w_tilde_preload f8[Py,Px], native_index_for_slim_index i8[M,2])
 * -> f8[M,M], W[i, j] = preload[y_j - y_i, x_j - x_i] for every pair (i, j), negative offsets wrapping as
 * numpy indexes (numba.py:282-291; the caller guarantees |offset| < Py, Px). */
int aj_w_tilde_via_preload_from(const double* w_tilde_preload, int64_t preload_rows, int64_t preload_cols,
                                const int64_t* native_index_for_slim_index, int64_t M, double* w_tilde, void* stream);

/* original.py:230-403 w_tilde_curvature_preload_interferometer_from(noise_map_real f8[K],
 * uv_wavelengths f8[K,2], shape_masked_pixels_2d (Y, X), grid_radians_2d f8[Gy,Gx,2]) -> f8[2Y,2X]: for
 * the Y x X corner of the grid, entry [i, j] = sum_k n_k^-2 cos(2 pi (dx u_k0 + dy u_k1)) with (dy, dx)
 * = g[0,0] - g[i,j]; [i, -j], [-i, j], [-i, -j] (numpy indexes) take the offsets from the corners
 * g[0,X-1], g[Y-1,0], g[Y-1,X-1] to g[i,X-1-j], g[Y-1-i,j], g[Y-1-i,X-1-j]; row Y and column X are 0.
 * Y <= Gy, X <= Gx.  Workspace from aj_w_tilde_preload_workspace_bytes. */
size_t aj_w_tilde_preload_workspace_bytes(int64_t K, int64_t y_shape, int64_t x_shape);
int aj_w_tilde_curvature_preload_interferometer_from(const double* noise_map_real, const double* uv_wavelengths,
                                                     int64_t K, int64_t y_shape, int64_t x_shape,
                                                     const double* grid_radians_2d, int64_t grid_rows,
                                                     int64_t grid_cols, double* curvature_preload, void* workspace,
                                                     size_t workspace_bytes, void* stream);

/* numba.py:95-152 w_tilde_data_interferometer_from(visibilities_real f8[K],
 * noise_map_real f8[K], uv_wavelengths f8[K,2], grid_radians_slim f8[M,2]) -> f8[M].
 * Same weight as the reference, (n_k^2 / v_k)^2 (numba.py:149). */
size_t aj_visibility_sum_workspace_bytes(int64_t K, int64_t M);
int aj_w_tilde_data_interferometer_from(const double* visibilities_real, const double* noise_map_real,
                                        const double* uv_wavelengths, const double* grid_radians_slim, int64_t K,
                                        int64_t M, double* w_tilde_data, void* workspace, size_t workspace_bytes,
                                        void* stream);

/* Dirty image consistent with W-tilde: sum_k (Re d_k cos phi_ik + Im d_k sin phi_ik) / n_k^2,
 * phi_ik = 2 pi (g_i1 u_k0 + g_i0 u_k1), n_k = Re noise_map_k.  No reference counterpart (the
 * reference takes the dirty image as an input, numba.py:762); data preparation for its callers.
 * Uses aj_visibility_sum_workspace_bytes. */
int aj_dirty_image_from(const double* data, const double* noise_map, const double* uv_wavelengths,
                        const double* grid_radians_slim, int64_t K, int64_t M, double* dirty_image, void* workspace,
                        size_t workspace_bytes, void* stream);

/* W-tilde and the data-side sums in one pass over the visibilities: no single reference counterpart; it
 * fuses numba.py:155-207 w_tilde_curvature_interferometer_from with numba.py:95-152
 * w_tilde_data_interferometer_from and the dirty image (the sums the reference evaluates in loops of
 * their own), so that each cos/sin of a (pixel, visibility) phase is evaluated once.  Outputs as the three
 * separate calls: w_tilde[M*M] as aj_w_tilde_curvature_interferometer_from at this precision;
 * dirty_image[M] as aj_dirty_image_from with n_k = noise_map_real[k] (data: K complex visibilities,
 * (re, im) interleaved); w_tilde_data[M] as aj_w_tilde_data_interferometer_from.  dirty_image and
 * w_tilde_data may each be null (not computed; then data, resp. visibilities_real, may be null).  The data
 * sums are FP64 at every precision: with AJ_PREC_FP64 and the residue modes they reuse the FP64 cos/sin
 * of the W-tilde features; the single-precision modes run the stand-alone sums after the build. */
size_t aj_w_tilde_with_data_workspace_bytes(int64_t K, int64_t M, int precision);
int aj_w_tilde_with_data_interferometer_from(const double* noise_map_real, const double* uv_wavelengths,
                                             const double* grid_radians_slim, int64_t K, int64_t M, int precision,
                                             const double* data, const double* visibilities_real, double* w_tilde,
                                             double* dirty_image, double* w_tilde_data, void* workspace,
                                             size_t workspace_bytes, void* stream);

/* numba.py:293-376 mapping_matrix_from(pix_indexes_for_sub_slim_index i8[M,B],
 * pix_weights_for_sub_slim_index f8[M,B], pixels) -> f8[M,pixels] (dense, assignment). */
int aj_mapping_matrix_from(const int64_t* pix_indexes_for_sub_slim_index, const double* pix_weights_for_sub_slim_index,
                           int64_t M, int64_t B, int64_t pixels, double* mapping_matrix, void* stream);

/* numba.py:379-406 data_vector_from(mapping_matrix f8[M,S], dirty_image f8[M]) -> f8[S]. */
int aj_data_vector_from(const double* mapping_matrix, const double* dirty_image, int64_t M, int64_t S,
                        double* data_vector, void* stream);

/* numba.py:409-435 curvature_matrix_via_w_tilde_from(w_tilde f8[M,M], mapping_matrix f8[M,S]) -> f8[S,S]
 * (dense T^T W T on FP64 MFMA). */
size_t aj_curvature_dense_workspace_bytes(int64_t M, int64_t S);
int aj_curvature_matrix_via_w_tilde_from(const double* w_tilde, const double* mapping_matrix, int64_t M, int64_t S,
                                         double* curvature_matrix, void* workspace, size_t workspace_bytes,
                                         void* stream);

/* numba.py:552-587 sparse_mapping_matrix_transpose_matmul(matrix f8[M,S2], pix_indexes i8[M,B],
 * pix_weights f8[M,B], pixels) -> f8[pixels,S2]. */
size_t aj_sparse_transpose_matmul_workspace_bytes(int64_t M, int64_t B, int64_t S);
int aj_sparse_mapping_matrix_transpose_matmul(const double* matrix, const int64_t* pix_indexes_for_sub_slim_index,
                                              const double* pix_weights_for_sub_slim_index, int64_t M, int64_t S2,
                                              int64_t B, int64_t pixels, double* out, void* workspace,
                                              size_t workspace_bytes, void* stream);

/* numba.py:590-614 curvature_matrix_via_w_wilde_sparse_mapping_matrix_from(w_tilde f8[M,M],
 * pix_indexes i8[M,B], pix_weights f8[M,B], pixels) -> f8[pixels,pixels] (w_tilde symmetric). */
size_t aj_curvature_sparse_workspace_bytes(int64_t M, int64_t B, int64_t S);
int aj_curvature_matrix_via_w_wilde_sparse_mapping_matrix_from(const double* w_tilde,
                                                               const int64_t* pix_indexes_for_sub_slim_index,
                                                               const double* pix_weights_for_sub_slim_index, int64_t M,
                                                               int64_t B, int64_t pixels, double* curvature_matrix,
                                                               void* workspace, size_t workspace_bytes, void* stream);

/* numba.py:617-647 curvature_matrix_via_w_compact_sparse_mapping_matrix_from(w_compact f8[N,2N-1],
 * native_index_for_slim_index i8[M,2], pix_indexes i8[M,B], pix_weights f8[M,B], pixels) -> f8[pixels,pixels]:
 * as the sparse form above with W-tilde read from the lattice table.  Uses aj_curvature_sparse_workspace_bytes. */
int aj_curvature_matrix_via_w_compact_sparse_mapping_matrix_from(
    const double* w_compact, int64_t grid_size, const int64_t* native_index_for_slim_index,
    const int64_t* pix_indexes_for_sub_slim_index, const double* pix_weights_for_sub_slim_index, int64_t M, int64_t B,
    int64_t pixels, double* curvature_matrix, void* workspace, size_t workspace_bytes, void* stream);

/* original.py:600-700 curvature_matrix_via_w_tilde_curvature_preload_interferometer_from(curvature_preload
 * f8[Py,Px], pix_indexes i8[M,B], pix_size i8[M], pix_weights f8[M,B], native_index_for_slim_index i8[M,2],
 * pix_pixels) -> f8[pixels,pixels] = T^T W T with W[i, j] read from the preload table at the offset of the
 * later pixel in slim order from the earlier one, as original.py:640-663 reads it (the pairs ip0 < ip1, then
 * the transpose).  The mapping is in this library's format (-1 ends a row; the
 * Python layer turns pix_size into it).  Uses aj_curvature_sparse_workspace_bytes. */
int aj_curvature_matrix_via_w_tilde_curvature_preload_interferometer_from(
    const double* curvature_preload, int64_t preload_rows, int64_t preload_cols,
    const int64_t* native_index_for_slim_index, const int64_t* pix_indexes_for_sub_slim_index,
    const double* pix_weights_for_sub_slim_index, int64_t M, int64_t B, int64_t pixels, double* curvature_matrix,
    void* workspace, size_t workspace_bytes, void* stream);

/* numba.py:656-695 constant_regularization_matrix_from(coefficient, neighbors i8[S,P],
 * neighbors_sizes i8[S]) -> f8[S,S]. */
int aj_constant_regularization_matrix_from(double coefficient, const int64_t* neighbors,
                                           const int64_t* neighbors_sizes, int64_t S, int64_t P,
                                           double* regularization_matrix, void* stream);

/* numba.py:698-731 reconstruction_positive_negative_from(data_vector f8[S], curvature_reg_matrix f8[S,S])
 * -> f8[S], as its np.linalg.solve (numba.py:731): LU with partial pivoting of the whole matrix (symmetric or
 * not, definite or not), S <= 16384.  logdet (may be NULL) receives log|det A| (np.linalg.slogdet(A)[1],
 * numba.py:849); reconstruction may be NULL (then data_vector may be NULL too: log-determinant only);
 * info[0] = 0 or the 1-based column of an exact zero pivot (A singular: numpy raises LinAlgError). */
size_t aj_solve_workspace_bytes(int64_t S);
int aj_reconstruction_positive_negative_from(const double* data_vector, const double* curvature_reg_matrix, int64_t S,
                                             double* reconstruction, double* logdet, int64_t* info, void* workspace,
                                             size_t workspace_bytes, void* stream);

/* aj_reconstruction_positive_negative_from without a workspace argument and without any allocation: no
 * counterpart in the reference; for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its scratch is a file-scope device array
 * of 33,571,072 bytes (the workspace of S = 2048), one per GPU, held until the process ends; S above 2048
 * is rejected with AJ_EINVAL.  Every call on a GPU shares that GPU's scratch: calls that overlap on one GPU
 * (other streams or host threads) race and must be serialised by the caller. */
int aj_reconstruction_positive_negative_from_noalloc(const double* data_vector, const double* curvature_reg_matrix,
                                                     int64_t S, double* reconstruction, double* logdet, int64_t* info,
                                                     void* stream);

/* numba.py:734-752 noise_normalization_complex_from(noise_map c16[K]) -> f8 (device scalar). */
int aj_noise_normalization_complex_from(const double* noise_map, int64_t K, double* out, void* stream);

/* chi.chi of numba.py:846-847 and the noise normalisation of numba.py:752 in one pass:
 * out2[0] = sum over 2K of (data/noise)^2, out2[1] = noise normalisation. */
int aj_chi_squared_noise_normalization(const double* data, const double* noise_map, int64_t K, double* out2,
                                       void* stream);

/* numba.py:755-857 log_likelihood_function_via_w_tilde_from(dirty_image f8[M], data c16[K],
 * noise_map c16[K], w_tilde f8[M,M], pix_indexes i8[M,B], pix_weights f8[M,B], neighbors i8[S,P],
 * neighbors_sizes i8[S]) -> f8, coefficient 1 as at numba.py:819.  log_likelihood is a device
 * scalar; info[2] = failing pivots of A and H (0 = none).  S at most 14272 (as the batched form). */
size_t aj_log_likelihood_workspace_bytes(int64_t M, int64_t K, int64_t B, int64_t S, int64_t P);
int aj_log_likelihood_function_via_w_tilde_from(const double* dirty_image, const double* data, const double* noise_map,
                                                const double* w_tilde, const int64_t* pix_indexes_for_sub_slim_index,
                                                const double* pix_weights_for_sub_slim_index, const int64_t* neighbors,
                                                const int64_t* neighbors_sizes, int64_t M, int64_t K, int64_t B,
                                                int64_t S, int64_t P, double* log_likelihood, int64_t* info,
                                                void* workspace, size_t workspace_bytes, void* stream);

/* numba.py:860-967 log_likelihood_function_via_w_compact_from(dirty_image, data, noise_map,
 * w_compact f8[N,2N-1], native_index_for_slim_index i8[M,2], pix_indexes, pix_weights, neighbors,
 * neighbors_sizes) -> f8, coefficient 1 as at numba.py:923.  Uses aj_log_likelihood_workspace_bytes. */
int aj_log_likelihood_function_via_w_compact_from(
    const double* dirty_image, const double* data, const double* noise_map, const double* w_compact, int64_t grid_size,
    const int64_t* native_index_for_slim_index, const int64_t* pix_indexes_for_sub_slim_index,
    const double* pix_weights_for_sub_slim_index, const int64_t* neighbors, const int64_t* neighbors_sizes, int64_t M,
    int64_t K, int64_t B, int64_t S, int64_t P, double* log_likelihood, int64_t* info, void* workspace,
    size_t workspace_bytes, void* stream);

/* Batched log-evidence: numba.py:755-857 evaluated for n samples that share one w_tilde, with an
 * explicit regularisation coefficient per sample (the reference fixes it to 1, numba.py:819).
 * Sample b has S_host[b] source pixels (1 .. 14272: the solve holds a sample's solution in LDS; larger
 * samples are rejected with AJ_EINVAL); mapping blocks are [n,M,B]; neighbour rows of all samples
 * are concatenated ([sum S, P], [sum S]).  chi2_noise_norm is the device pair from
 * aj_chi_squared_noise_normalization.  deflation_beta_host (may be NULL) gives per sample the
 * rank-one shift used for log|H| (0 = plain factorisation of H).  Outputs: log_evidence[n];
 * reconstruction[sum S] (may be NULL); logdets[n,2] = (log|A|, log|H|) (may be NULL);
 * info[n,2] failing pivots (may be NULL). */
size_t aj_log_evidence_batched_workspace_bytes(int64_t M, int64_t B, int64_t n_samples, const int64_t* S_host);
int aj_log_evidence_batched(const double* w_tilde, int64_t M, const double* dirty_image, const double* chi2_noise_norm,
                            int64_t n_samples, int64_t B, int64_t P, const int64_t* S_host, const int64_t* pix_indexes,
                            const double* pix_weights, const int64_t* neighbors, const int64_t* neighbors_sizes,
                            const double* coefficients_host, const double* deflation_beta_host, double* log_evidence,
                            double* reconstruction, double* logdets, int64_t* info, void* workspace,
                            size_t workspace_bytes, void* stream);

/* How the batched log-evidence factors its matrices (process-wide tuning, no reference counterpart):
 * 0 = three launches per 32-column step (diagonal blocks, panels, trailing tiles); 1 = the same steps in one
 * cooperative launch with a grid-wide barrier between phases; 2 = 1 with lookahead (the next diagonal block
 * factored by the workgroup that updates it); 3 = one launch per step with lookahead (the next step's
 * diagonal block and panel updated, factored and solved beside the step's trailing update); 4 = left-looking,
 * one launch per 32-column block after one that factors every first diagonal block (each strip of the block
 * updated from all earlier blocks in registers and solved, the next diagonal block formed and factored by
 * the block's first task; no trailing matrix written between steps); 5 (default, also -1) = the left-looking
 * tasks in one persistent launch (a workgroup per compute unit taking tasks from its XCD's queue, each
 * waiting on its matrix's count of completed tasks of the earlier block columns, with lookahead: a task
 * updates its strip by all but the last earlier block before the previous column is done).  All give bitwise
 * the same factors. */
int aj_log_evidence_configure(int chol_mode);

/* Tuning of schedule 5 (process-wide, no reference counterpart): a block column runs as 32-row tasks once
 * its 32-row tasks on an XCD number at most small_per_cu per compute unit of the XCD (256-row tasks before;
 * -1 = default 1); a dependency wait gives up after spin_limit polls (-1 = default 2^18, ~0.25 s), after
 * which the batch's info codes read -1 (the evaluation is reported failed, never returned). */
int aj_log_evidence_flow_configure(int small_per_cu, int spin_limit);

/* Task trace of schedule 5 (process-wide, a tracing aid with no reference counterpart): while set, every
 * batch whose task list fits capacity_tasks writes per task t, at device_buffer[16 t ..], sixteen int64: the
 * wall-clock ticks (100 MHz) when it was taken, when its first wait ended and when it was counted complete,
 * its matrix (2 sample + 0 for A, 1 for H), kind (0 its first diagonal block, 1 a 256-row task, 2 a 32-row
 * task), block column, task within the column, the compute unit's hardware id; then for 256-row tasks the
 * ticks when its first blocks' loads were issued, when its second wait began and ended, when its update
 * ended and when its substitution ended (slots 8-12), and for the first 256-row task of a block column the
 * ticks when the next diagonal block was updated, factored and stored (slots 13-15); the others are left as they were. */
int aj_log_evidence_flow_trace(void* device_buffer, int64_t capacity_tasks);

/* aj_log_evidence_batched with W-tilde given by its lattice table (numba.py:860-967 per sample):
 * same arguments and workspace (aj_log_evidence_batched_workspace_bytes) with w_tilde replaced by
 * w_compact f8[N,2N-1], grid_size N and native_index_for_slim_index i8[M,2].  No reference counterpart
 * for the batch. */
int aj_log_evidence_batched_compact(const double* w_compact, int64_t grid_size,
                                    const int64_t* native_index_for_slim_index, int64_t M, const double* dirty_image,
                                    const double* chi2_noise_norm, int64_t n_samples, int64_t B, int64_t P,
                                    const int64_t* S_host, const int64_t* pix_indexes, const double* pix_weights,
                                    const int64_t* neighbors, const int64_t* neighbors_sizes,
                                    const double* coefficients_host, const double* deflation_beta_host,
                                    double* log_evidence, double* reconstruction, double* logdets, int64_t* info,
                                    void* workspace, size_t workspace_bytes, void* stream);

/* Backward of aj_log_evidence_batched (no reference counterpart: the reference differentiates its
 * JAX form, jax.py:1032-1058, by autodiff).  Uses the Cholesky factors and Omega = W T left in
 * forward_workspace by the immediately preceding aj_log_evidence_batched call with the same batch,
 * and that call's reconstruction.  coefficients_dev is the device copy of the coefficients.
 * Outputs dcoefficients[n] = dlogL_b/dc_b and dweights[n,M,B] = dlogL_b/dt_b[m,k] (0 after a -1). */
size_t aj_log_evidence_backward_workspace_bytes(int64_t M, int64_t n_samples, const int64_t* S_host);
int aj_log_evidence_batched_backward(const double* w_tilde, int64_t M, const double* dirty_image, int64_t n_samples,
                                     int64_t B, int64_t P, const int64_t* S_host, const int64_t* pix_indexes,
                                     const double* pix_weights, const int64_t* neighbors,
                                     const int64_t* neighbors_sizes, const double* coefficients_dev,
                                     const double* reconstruction, const void* forward_workspace,
                                     size_t forward_workspace_bytes, double* dcoefficients, double* dweights,
                                     void* workspace, size_t workspace_bytes, void* stream);

/* Barycentric weights of fixed points positions[M,2] (y, x) in the triangles triangles[M,3] of
 * vertices[S,2]: weights[M,3] (each image pixel's source-mesh interpolation, done upstream of the
 * reference by its mapper).  aj_barycentric_vjp maps dweights[M,3] to dvertices[S,2] with the
 * points held fixed. */
int aj_barycentric_weights(const double* vertices, const double* positions, const int64_t* triangles, int64_t M,
                           double* weights, void* stream);
size_t aj_barycentric_vjp_workspace_bytes(int64_t M, int64_t S);
int aj_barycentric_vjp(const double* vertices, const double* positions, const int64_t* triangles, int64_t M, int64_t S,
                       const double* dweights, double* dvertices, void* workspace, size_t workspace_bytes,
                       void* stream);

/* Source-mesh mapper (the reference builds it upstream, external/get_dataset.py:38-60: a Delaunay mesh
 * of the source vertices, its neighbour lists and the image pixels' barycentric mapping; no function of
 * the reference's modules).  Samples s = 0 .. n_samples-1 own vertices [vertex_offsets[s],
 * vertex_offsets[s+1]) of vertices f8[*, 2] (y, x), 3 to 3072 each (vertex_offsets on the device and
 * the same offsets on the host).
 * aj_delaunay_mesh_batched: the Delaunay triangulation of each sample's vertices (one workgroup per
 * sample): finite triangles i8[2 * total vertices, 3] (sample s from row 2 * vertex_offsets[s], local
 * vertex ids, counter-clockwise in (x, y)), n_triangles[s]; neighbors i8[total, max_neighbors] (sorted,
 * -1 padded, local ids) and neighbors_sizes[total] in the format of numba.py:672-677; status[s] = 0 ok,
 * 1 too many vertices, 2 all vertices on a line, 3 / 4 an insertion failed (capacity / geometry), 5 a
 * vertex has more than max_neighbors neighbours.  Coincident vertices after the first are left out.
 * aj_delaunay_locate_batched: each pixel position f8[n_samples, M, 2] (sample s at positions +
 * 2 * s * positions_stride; stride 0 shares one set) located in its sample's triangles:
 * pix_indexes_for_sub_slim_index i8[n_samples, M, 3] and pix_weights_for_sub_slim_index f8[n_samples,
 * M, 3] (barycentric, as aj_barycentric_weights; outside the hull the nearest vertex with weight 1, then
 * -1 / 0), the mapping format of numba.py:355-360. */
int aj_delaunay_mesh_batched(const double* vertices, const int64_t* vertex_offsets, const int64_t* vertex_offsets_host,
                             int64_t n_samples, int64_t max_neighbors, int64_t* triangles, int64_t* n_triangles,
                             int64_t* neighbors, int64_t* neighbors_sizes, int* status, void* stream);
int aj_delaunay_locate_batched(const double* vertices, const int64_t* vertex_offsets, int64_t n_samples,
                               const int64_t* triangles, const int64_t* n_triangles, const double* positions,
                               int64_t positions_stride, int64_t M, int64_t* pix_indexes_for_sub_slim_index,
                               double* pix_weights_for_sub_slim_index, void* stream);

#ifdef __cplusplus
}
#endif

#endif
