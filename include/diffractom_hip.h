/*
 * C ABI of the MI355X (gfx950) kernels of the diffractom single-phase operator and its SPDHG-Poisson solver.
 *
 * Plain device pointers, sizes and a hipStream_t; no torch types. Every function is asynchronous on the given
 * stream, allocates nothing, and returns 0 (DFM_OK) or a negative status; dfm_last_error() gives the message of
 * the last failure on the calling thread.
 *
 * The reference (diffractom, Python + OpenCL) launches its kernels through pyopencl; each entry point below replaces
 * the pyopencl launch(es) cited with it (paths relative to the reference's src/diffractom/). Layouts:
 *   coefficients  (K, Ny, Nx) float32, pixel p = x + Nx*y       (utils/arrays.py:3-8)
 *   channel-fastest images   img_k (Npix, kstride) float32, kstride a multiple of 4 (operators/parallel_radon.py:34-38)
 *   sinograms     (n_angles, My, kstride) float32, channel fastest
 *   data          (N_Omega, My, CP) float32, CP = N_eta * N_rings, segment j = c*N_rings + p
 *   geometry      geo (n_angles, 8) float32 per angle: X, Y, T0, wscale, hw, top, 0, 0 (see dfm_radon_forward)
 *   CSR           forward rows (angle, j) listing orientations k (uint16, ascending);
 *                 adjoint rows (angle, k) listing segments j (uint16, ascending); float32 values; int32 row pointers
 */
#ifndef DIFFRACTOM_HIP_H
#define DIFFRACTOM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* dfm_stream_t; /* the hipStream_t of the caller (NULL: the null stream) */

enum {
  DFM_OK = 0,
  DFM_ERR_ARG = -1,         /* bad argument (shape, size, range): the reference raises ValueError/TypeError */
  DFM_ERR_UNSUPPORTED = -2, /* unsupported combination: the reference raises NotImplementedError */
  DFM_ERR_LAUNCH = -3,      /* kernel launch failure */
  DFM_ERR_MEMORY = -4       /* out of memory: the reference raises MemoryError */
};

/* Library version (major*10000 + minor*100 + patch) and the message of the last failure on this thread. */
int dfm_version(void);
const char* dfm_last_error(void);

/* ---------------------------------------------------------------- layout (operators/parallel_radon.py:221-236,
 * radon_kernels.cl:192-244: gather_channels_k_fastest / scatter_channels_k_fastest) */
/* img_k[p, k] = coeffs[k0 + k, p] for k < kb, 0 for kb <= k < kstride. */
int dfm_gather_k(const float* coeffs, float* img_k, int64_t npix, int64_t k0, int32_t kb, int32_t kstride,
                 dfm_stream_t stream);
/* coeffs[k0 + k, p] = img_k[p, k] for k < kb. */
int dfm_scatter_k(const float* img_k, float* coeffs, int64_t npix, int64_t k0, int32_t kb, int32_t kstride,
                  dfm_stream_t stream);

/* ---------------------------------------------------------------- Radon (operators/parallel_radon.py:283-318,
 * radon_kernels.cl:44-141: radon_forward_k / radon_forward_tiled / radon_backward_k / radon_backward_tiled)
 * Per angle a: geo[8a..8a+7] = X, Y, T0, wscale, hw, top, 0, 0. Pixel (x, y) lands at t = X x + Y y + T0 (bins) and
 * adds K(t - s) to bin s, K(d) = clamp(hw - |d|, 0, top) (hw = (1 + A)/2, top = min(1, A), A = max(|X|, |Y|) for the
 * trapezoid; hw = top = 1 for the hat). ntap: bins a pixel can touch (2 when A <= 1).
 * forward:  sino[a, s, k] = scale * wscale_a * sum_pix K(t - s) img[pix, k]   (wscale_a = 1/A for the trapezoid)
 * backward: img[pix, k]   = sum_a wscale_a * sum_s K(t - s) sino[a, s, k]     (wscale_a = w_a/A: the angle weight)
 * Image channels k_off .. k_off + k_n - 1 of an (Npix, kstride) array are sinogram channels 0 .. k_n - 1 of an
 * (n_angles, ns, skstride) array (k_off, k_n, kstride, skstride multiples of 4). */
int dfm_radon_forward(const float* img_k, float* sino, const float* geo, int32_t nx, int32_t ny, int32_t ns,
                      int32_t n_angles, int32_t kstride, int32_t k_off, int32_t k_n, int32_t skstride, float scale,
                      int32_t ntap, dfm_stream_t stream);
int dfm_radon_backward(const float* sino, float* img_k, const float* geo, int32_t nx, int32_t ny, int32_t ns,
                       int32_t n_angles, int32_t kstride, int32_t k_off, int32_t k_n, int32_t skstride, int32_t ntap,
                       dfm_stream_t stream);

/* The same projector driven by precomputed geometry tables (radon_kernels.cl:44-141 with the per-pixel footprint
 * arithmetic done once per angle set on the host, in float64; ntap = 2, i.e. pixels no larger than a bin):
 * forward table: CSR over (angle, bin) rows, row_ptr (n_angles*ns + 1), entries (pixel, weight) with weight =
 *   scale * wscale_a * K(t - s);  sino[a, s, k] = sum_e weight_e img[pixel_e, k]
 * pixel table ptab (n_angles, Npix, 4) int32: bin s0, bin s1 (>= 0, < ns), float bits of w0, w1 (the backprojection
 *   weights, wscale_a folded in; 0 for taps off the detector): read by the plain tile variants; the grouped variants
 *   derive the same entries on the device from bgeo (n_angles, 6) float64: X, Y, T0, hw, top, wscale_a / A
 *   (s0 = floor(t - hw) + 1, s1 = s0 + 1, w = wscale_a / A clamp(hw - |t - s|, 0, top));
 * tile bases tbase (n_angles, ceil(nx/4) * ceil(ny/4)): the first of the 8 bins every 4 x 4 pixel tile reaches.
 * mode 0: img_k = backward(sino); mode 1: the SPDHG primal update of x, z with step T, as dfm_radon_backward_spdhg;
 * mode 2 (grouped variants: + 64 G, + 1024, + 4096 w): the EM update x = T > 0 ? x backward(sino) / (T + lam) : x with
 * T the EM sensitivity, lam >= 0, z unused (osem_poisson.py:69-73 osem_update applied to the output of
 * radon_backward_k); mode 3 (grouped variants): the PDHG primal step and extrapolation x' = T > 0 ? max(0, x - T
 * (backward(sino) + lam)) : 0, z = 2 x' - x (x_bar), x = x' (pdhg_poisson.py:69-80 primal_update applied to the
 * output of radon_backward_k); the plain tile variants return DFM_ERR_UNSUPPORTED for modes 2 and 3.
 * variant: a tuning choice with identical results (forward: table entries fetched ahead, 4 or 8, plus 256 (the
 * workgroup's table entries staged in LDS), plus 512 q for channel blocks of 64 (q 0) or 128 (q 3) channels;
 * backward: 4, a workgroup per 4 x 4 tile and 64 channels (4 angles staged per pass); or ACH + 64 G (ACH 4 or 8
 * angles per pass; G 2, 4 or 8): a workgroup per tile and G channel blocks, the tile's table rows staged once; + 1024
 * (ACH 4 or 8, G 1, 2 or 4): 8 x 8 tiles, tbase then holding each 8 x 8 tile's first of 16 bins), picked by timing
 * (the reference times its tiled / plain projectors the same way, parallel_radon.py:240-281). */
int dfm_radon_forward_table(const float* img_k, float* sino, const int32_t* row_ptr, const int32_t* ent_pix,
                            const float* ent_w, int32_t ns, int32_t n_angles, int32_t kstride, int32_t k_off,
                            int32_t k_n, int32_t skstride, int32_t variant, dfm_stream_t stream);
/* The forward projection (radon_kernels.cl:44-101 radon_forward_k) by line sweep: per angle group (grp_ang: n_grp rows
 * of na_grp angle indices, -1 for none; grp_cls: 0 = sweep image rows, for |X| >= |Y|, 1 = columns) and 16
 * orientations, `lines` image lines are staged in LDS at a time and each ray adds its pixels on them; geo: per angle
 * (X, Y, T0, hw, top, scale) in float64, the weight scale * min(hw - |X x + Y y + T0 - s|, top) where positive.
 * Needs nx, ny <= 256; lines: image lines per LDS block (within the LDS: dfm_radon_sweep_lines) + 16 w for 16 << w
 * orientations per workgroup: 16 orientations (w 0) x 2, 4 or 8 lines or 32 (w 1) x 1 or 2 lines, 1024 threads;
 * na_grp * ns <= 1024. */
int dfm_radon_forward_sweep(const float* img_k, float* sino, const double* geo, const int32_t* grp_ang,
                            const int32_t* grp_cls, int32_t n_grp, int32_t na_grp, int32_t ns, int32_t nx, int32_t ny,
                            int32_t kstride, int32_t k_off, int32_t k_n, int32_t skstride, int32_t lines,
                            dfm_stream_t stream);
/* The most image lines per LDS block of dfm_radon_forward_sweep for an nx x ny image (0: unsupported). */
int dfm_radon_sweep_lines(int32_t nx, int32_t ny);
/* Line sweep with two neighbouring rays per thread (the same result as dfm_radon_forward_sweep): groups of na_grp
 * angles with na_grp * ceil(ns / 2) <= 512; usable when (2 hw + 1) / max(|X|, |Y|) < 4 for every angle (square
 * images with the trapezoid footprint); lines: image lines staged per LDS block (1, 2, or 4 for images of at most
 * 128 pixels a side), + 8 for the form that loads each block at its start instead of holding the next block's loads
 * across the sums (lines 1, or 2 for images of at most 128 pixels a side), or + 16 for the layout holding every
 * staged pixel twice so that the LDS reads of 16 lanes never share a bank (the same lines; twice the LDS). */
int dfm_radon_forward_pairs_sweep(const float* img_k, float* sino, const double* geo, const int32_t* grp_ang,
                                  const int32_t* grp_cls, int32_t n_grp, int32_t na_grp, int32_t ns, int32_t nx,
                                  int32_t ny, int32_t kstride, int32_t k_off, int32_t k_n, int32_t skstride,
                                  int32_t lines, dfm_stream_t stream);
/* The forward projection of dfm_radon_forward_table tiled by nb bands of image rows: band_ptr[(a ns + s) nb + b]
 * is the first table entry of ray (a, s) in band b (entries ascend in pixel index), band_ptr[na ns nb] the total.
 * Per band the partial sums go to part (nb * n_angles * ns * skstride floats), then sino = their sum in band order.
 * variant: table entries fetched ahead (4 or 8). */
int dfm_radon_forward_band(const float* img_k, float* sino, float* part, const int32_t* band_ptr,
                           const int32_t* ent_pix, const float* ent_w, int32_t ns, int32_t n_angles, int32_t nb,
                           int32_t kstride, int32_t k_off, int32_t k_n, int32_t skstride, int32_t variant,
                           dfm_stream_t stream);
int dfm_radon_backward_table(const float* sino, float* img_k, const int32_t* ptab, const double* bgeo,
                             const int32_t* tbase, int32_t nx, int32_t ny, int32_t ns, int32_t n_angles,
                             int32_t kstride, int32_t k_off, int32_t k_n, int32_t skstride, float* x, float* z,
                             const float* T, float inv_p, float lam, int32_t mode, int32_t variant,
                             dfm_stream_t stream);

/* ---------------------------------------------------------------- pole-figure matrix
 * Dense PF batch (create_pfo_matrix.py:108-199 pfmatrix_eval_poles): out (R, kmax, C, P) for orientations
 * k0 .. k0+kb-1 (rows kb..kmax-1 zero). coords (R, S, C, Se, P, 3); poles (K, A, 3); axis_start (P+1);
 * axis_count (A); inv_sigma2, norm (K); qw (S*Se) used when quad_weighted (else the node mean). Directions, poles
 * and kernel widths are float64 (the reference uses float32 copies; float64 keeps the entries of narrow kernels
 * accurate to float32 rounding). */
int dfm_pf_eval_dense(const double* coords, const double* poles, const int32_t* axis_start, const float* axis_count,
                      const double* inv_sigma2, const double* norm, const float* qw, float* out, int32_t R,
                      int32_t kmax, int32_t C, int32_t P, int32_t A, int32_t S, int32_t Se, int32_t k0, int32_t kb,
                      float pf_cut, int32_t quad_weighted, dfm_stream_t stream);
/* Dense batch -> CSR (pf_kernels.cl:482-548 pf_count_rows_* / pf_fill_rows_*): counts per forward row (r, j) /
 * adjoint row (r, k); fill writes col and val from the row pointers (exclusive scan of the counts). */
int dfm_pf_count_fwd(const float* pf, int32_t* counts, int32_t R, int32_t kmax, int32_t CP, int32_t kb,
                     dfm_stream_t stream);
int dfm_pf_count_adj(const float* pf, int32_t* counts, int32_t R, int32_t kmax, int32_t CP, int32_t kb,
                     dfm_stream_t stream);
int dfm_pf_fill_fwd(const float* pf, const int32_t* row_ptr, uint16_t* col_k, float* val, int32_t R, int32_t kmax,
                    int32_t CP, int32_t kb, dfm_stream_t stream);
int dfm_pf_fill_adj(const float* pf, const int32_t* row_ptr, uint16_t* col_j, float* val, int32_t R, int32_t kmax,
                    int32_t CP, int32_t kb, dfm_stream_t stream);

/* ---------------------------------------------------------------- PF products (pf_kernels.cl:597-687
 * spmm_pf_forward_c / spmm_pf_adjoint_c; single_phase_forward_operator.py:38-123 the _sub forms)
 * Sinogram row rs (0..n_angles-1) belongs to data row omegas[rs] (omegas NULL: rs). CSR rows: rs if compact, else
 * omegas[rs].
 * forward: data[r, y, j] (+)= sum_k PF[r, k, j] sino[rs, y, k]   (accumulate 0: overwrite every j; 1: add)
 * adjoint: sino[rs, y, k] = alpha sum_j PF[r, k, j] data[r, y, j]  (k < kb; kb <= k < kstride set to 0) */
int dfm_spmm_forward(const float* sino, const int32_t* row_ptr, const uint16_t* col_k, const float* val, float* data,
                     const int32_t* omegas, int32_t n_angles, int32_t my, int32_t cp, int32_t kstride, int32_t kb,
                     int32_t compact, int32_t accumulate, dfm_stream_t stream);
/* forward with the orientations split over ksplit workgroup ranges (ksplit times the parallelism of one angle set),
 * partial sums in part (ksplit * n_angles * my * cp floats, unused when ksplit = 1), summed in a fixed order;
 * kc: orientations staged in LDS per pass (32 .. 768). Same result as dfm_spmm_forward up to summation order. */
int dfm_spmm_forward_split(const float* sino, const int32_t* row_ptr, const uint16_t* col_k, const float* val,
                           float* data, const int32_t* omegas, int32_t n_angles, int32_t my, int32_t cp,
                           int32_t kstride, int32_t kb, int32_t compact, int32_t accumulate, float* part,
                           int32_t ksplit, int32_t kc, dfm_stream_t stream);
/* The same products on transposed operands (wavefront-segmented SpMM: a wave per R CSR rows, lanes over y, column
 * chunks of the operand staged in LDS), for either CSR form:
 *   outT[s, a, row, y] = alpha sum_{e in row (rr, row), col(e) in split s} val[e] inT[a, col(e), y]
 * with rr = a if compact else omegas[a]; inT (n_angles, ncols, ys), outT (nsplit, n_angles, rows_out, ys), ys >= my a
 * multiple of 4, inT 16-byte aligned; split s covers columns [s cspan, (s + 1) cspan), cspan = ceil(ncols / nsplit);
 * rows nrows .. rows_out - 1 are written 0; my <= 512; kc: columns per LDS chunk (0: the most that fit).
 * forward: CSR (row_ptr_f, col_k), rows = cp, columns = kb, inT = the sinogram transposed;
 * adjoint: CSR (row_ptr_a, col_j), rows = kb (rows_out = kstride), columns = cp, inT = the data rows transposed. */
int dfm_spmm_t(const float* inT, const int32_t* row_ptr, const uint16_t* col, const float* val, float* outT,
               const int32_t* omegas, int32_t n_angles, int32_t my, int32_t ys, int32_t ncols, int32_t nrows,
               int32_t rows_out, int32_t compact, float alpha, int32_t nsplit, int32_t kc, dfm_stream_t stream);
/* dfm_spmm_t with either operand untransposed: in_ld > 0: the input is (n_angles, my, in_ld) (the sinogram for the
 * forward), transposed while it is staged in LDS (in_ld a multiple of 4; the column splits and chunks then start at
 * multiples of 4); in_ld == 0: transposed (n_angles, ncols, in_ys). The input is 16-byte aligned.
 * out_ld > 0 (nsplit 1): the output is written as (n_angles, my, out_ld) (the sinogram for the adjoint), through an
 * LDS transpose of each workgroup's rows; out_ld == 0: as dfm_spmm_t's outT with row stride out_ys. */
int dfm_spmm_t_direct(const float* in, int64_t in_ld, int32_t in_ys, const int32_t* row_ptr, const uint16_t* col,
                      const float* val, float* out, int64_t out_ld, int32_t out_ys, const int32_t* omegas,
                      int32_t n_angles, int32_t my, int32_t ncols, int32_t nrows, int32_t rows_out, int32_t compact,
                      float alpha, int32_t nsplit, int32_t kc, dfm_stream_t stream);
/* CSR rows per workgroup of dfm_spmm_t for a given my (0: my unsupported). */
int dfm_spmm_t_rows_per_block(int32_t my);
/* out[ao, c, r] (+)= sum_{s < nsplit} in[s split_stride + ai in_mat + r in_ld + c] for a = 0 .. n_mat-1, with
 * ai = omegas[a] if omegas and omega_out == 0, ao = omegas[a] if omegas and omega_out == 1, else a. */
int dfm_transpose_sum(const float* in, float* out, const int32_t* omegas, int32_t omega_out, int32_t n_mat,
                      int32_t nr, int32_t nc, int64_t in_ld, int64_t in_mat, int64_t out_ld, int64_t out_mat,
                      int32_t nsplit, int64_t split_stride, int32_t accumulate, dfm_stream_t stream);
int dfm_spmm_adjoint(const float* data, const int32_t* row_ptr, const uint16_t* col_j, const float* val, float* sino,
                     const int32_t* omegas, int32_t n_angles, int32_t my, int32_t cp, int32_t kstride, int32_t kb,
                     float alpha, int32_t compact, dfm_stream_t stream);
/* data rows omegas[s] (each row_len values) <- 0 (single_phase_forward_operator.py:126-132 zero_rows). */
int dfm_zero_rows(float* data, const int32_t* omegas, int32_t n_sub, int64_t row_len, dfm_stream_t stream);

/* ---------------------------------------------------------------- dense PF products on the matrix cores
 * Replace the dense mode's CLBlast gemmStridedBatched calls (single_phase_forward_operator.py:2063-2127; PF batch
 * from _eval_pf_batch 1131-1166) with BF16 MFMA (v_mfma_f32_16x16x32_bf16), FP32 accumulation.
 * PF is stored as bf16 "hi" fragments and, for the residual correction, bf16 "lo" fragments (lo = bf16(PF - hi));
 * lo NULL: hi only. Forward fragments (r, nt, kc) hold PF[r, 32 kc + 8 (l / 16) + i, 16 nt + l % 16] in lane l,
 * adjoint fragments (r, kt, jc) hold PF[r, 16 kt + l % 16, 32 jc + 8 (l / 16) + i]; element offset
 * ((r * n_tiles + tile) * n_chunks + chunk) * 512 + 8 l + i; entries beyond kb or cp are 0.
 * pack: from a dense FP32 batch pf[r, k, j] (R x km x cp, rows k < kb used). Each fragment array holds
 * R * n_tiles * n_chunks * 512 bf16 (forward: n_nt = ceil(cp / 16), n_kc = ceil(kb / 32); adjoint:
 * n_kt = ceil(kb / 16) (at least kstride / 16), n_jc = ceil(cp / 32)). */
int dfm_mfma_pack_fwd(const float* pf, void* hi, void* lo, int32_t R, int32_t km, int32_t kb, int32_t cp,
                      int32_t n_nt, int32_t n_kc, dfm_stream_t stream);
int dfm_mfma_pack_adj(const float* pf, void* hi, void* lo, int32_t R, int32_t km, int32_t kb, int32_t cp,
                      int32_t n_kt, int32_t n_jc, dfm_stream_t stream);
/* forward: data[omegas[a], y, j] (+)= sum_k PF[omegas[a], k, j] sino[a, y, k] (sino row stride sks, columns < kb
 * read). ksplit > 1 splits k over that many partial sums in part (ksplit * n_ang * my * cp floats), summed in a
 * fixed order. variant: j tiles per wave (2 or 4). The y-side operand is split into bf16 hi (+ lo when lo != NULL:
 * hi*hi + hi*lo + lo*hi). */
int dfm_mfma_forward(const float* sino, int32_t sks, int32_t kb, const void* hi, const void* lo,
                     const int32_t* omegas, int32_t n_ang, int32_t my, int32_t cp, int32_t n_nt, int32_t n_kc,
                     float* data, float* part, int32_t ksplit, int32_t accumulate, int32_t variant,
                     dfm_stream_t stream);
/* adjoint: sino[a, y, k] = alpha sum_j PF[omegas[a], k, j] data[omegas[a], y, j] for k < sks (columns >= kb are
 * 0). variant: k tiles per wave (2 or 4). */
/* The FP32 dense PF products (gemmStridedBatched, single_phase_forward_operator.py:2063-2127), per angle as of
 * the call's list (omegas, NULL: 0 .. n_ang - 1): C[y][c] (+)= alpha sum_r A[y][r] B(r, c), A row-major (lda), B(r, c)
 * = b[c ldb + r] with transpose_b (the adjoint: data x PF^T) else b[r ldb + c] (the forward: sino x PF), C
 * row-major (ldc); M x N outputs, R terms. Each operand's angle block is at its stride times the angle (a_by_angle /
 * c_by_angle: the angle index omegas[as]) or times the position as in the list; b always by angle. r_split > 1:
 * the R terms in r_split parts (multiples of 16) summed in order from partial (r_split n_ang M N floats);
 * n_ang r_split < 65536. */
int dfm_dense_f32(int32_t transpose_b, const float* a, int64_t a_stride, int32_t lda, int32_t a_by_angle,
                  const float* b, int64_t b_stride, int32_t ldb, float* c, int64_t c_stride, int32_t ldc,
                  int32_t c_by_angle, const int32_t* omegas, int32_t n_ang, int32_t m, int32_t n, int32_t r,
                  float alpha, int32_t accumulate, int32_t r_split, float* partial, dfm_stream_t stream);
int dfm_mfma_adjoint(const float* data, const void* hi, const void* lo, const int32_t* omegas, int32_t n_ang,
                     int32_t my, int32_t cp, int32_t n_kt, int32_t n_jc, float* sino, int32_t sks, float alpha,
                     int32_t variant, dfm_stream_t stream);
/* Grouped launch: several products of one direction (forward != 0: the forward) with different PF matrices, K,
 * sinogram column offsets and data rows in one launch (slices of a multi-slice batch). descs: a device array of
 * products; wg_map: a device array of n_wg int32 pairs (product, workgroup of that product); a product's workgroups
 * are n_ang * n_yg * n_tg * ksplit, decomposed (tile group fastest, then y group, k split, angle). residual: the
 * products have lo fragments. The fields (their meaning per direction as in dfm_mfma_forward / dfm_mfma_adjoint):
 *   src: forward the sinogram (first column of the product), adjoint the data; hi, lo: fragments; omegas: the data
 *   rows of the angles (NULL: 0..n_ang-1); out: forward data (direct = 1) or partial sums, adjoint the sinogram (first
 *   column); sks: sinogram row stride; kb: forward columns read; n_tiles, n_chunks: as n_nt, n_kc / n_kt, n_jc;
 *   mt_n, n_yg: 16-row y tiles per group and y groups (ceil(ceil(my/16)/8) groups); n_tg: groups of 4 * variant
 *   output tiles; ch_per: chunks per k split; ncols: adjoint columns written. */
typedef struct dfm_mfma_desc {
  const float* src;
  const void* hi;
  const void* lo;
  const int32_t* omegas;
  float* out;
  int32_t n_ang, my, cp, sks, kb;
  int32_t n_tiles, n_chunks;
  int32_t mt_n, n_yg, n_tg;
  int32_t ksplit, ch_per;
  int32_t direct, accumulate;
  float alpha;
  int32_t ncols;
} dfm_mfma_desc;
int dfm_mfma_grouped(int32_t forward, const dfm_mfma_desc* descs, const int32_t* wg_map, int64_t n_wg,
                     int32_t residual, int32_t variant, dfm_stream_t stream);

/* Block-sparse PF on the matrix cores (the sparse PF mode with 16-bit operands; replaces spmm_pf_forward_c /
 * spmm_pf_adjoint_c, pf_kernels.cl:597-687, and the subset forms single_phase_forward_operator.py:38-123). The PF of
 * an orientation batch (kb orientations) is cut into 32 x 32 blocks (angle r, orientation block of 32, segment block
 * of 32), block index (r * n_kb + kb_i) * n_jb + jb_i with n_kb = ceil(kb / 32), n_jb = ceil(cp / 32); block b lists
 * its non-zeros ptr[b] .. ptr[b + 1] - 1 (int32 pointers, R * n_kb * n_jb + 1 of them) in pos (uint16: k % 32 in
 * bits 0-4, j % 32 in bits 5-9) and val. One copy serves both directions.
 * count: counts[b] = non-zeros of block b of a dense FP32 batch pf[r, k, j] (R x km x cp, rows k < kb).
 * fill: with ptr = the exclusive scan of counts, the entries in element order (k-major within the block); fp16 = 0:
 * val uint32 = bf16(v) | bf16(v - bf16(v)) << 16; fp16 = 1: val uint16 = fp16(v * ring_scale[j % P]).
 * forward: data[omegas[a], y, j] (+)= sum_k PF[omegas[a], k, j] sino[a, y, k] (sino row stride sks, columns < kb);
 * ksplit > 1 splits the orientation blocks over that many partial sums in part (ksplit * n_ang * my * cp floats),
 * summed in a fixed order. adjoint: sino[a, y, k] = alpha sum_j PF[omegas[a], k, j] data[omegas[a], y, j] for every
 * k < sks (0 at and beyond kb). mode: 0 bf16 (val hi only), 1 bf16 with the hi/lo residual correction (hi*hi + hi*lo
 * + lo*hi), 2 fp16 (val fp16; ring_inv[p] = 1 / ring_scale[p]; the y-side operand scaled by a power of two per row
 * and chunk of 32); + 16: the blocks are indexed by the position a in the angle list (a generated PF of just these
 * angles, dfm_pfgen) instead of by omegas[a]. FP32 accumulation. */
int dfm_pfblk_count(const float* pf, int32_t* counts, int32_t R, int32_t km, int32_t kb, int32_t cp,
                    dfm_stream_t stream);
int dfm_pfblk_fill(const float* pf, const int32_t* ptr, uint16_t* pos, void* val, int32_t fp16,
                   const float* ring_scale, int32_t P, int32_t R, int32_t km, int32_t kb, int32_t cp,
                   dfm_stream_t stream);
int dfm_pfblk_forward(const float* sino, int32_t sks, int32_t kb, const int32_t* ptr, const uint16_t* pos,
                      const void* val, int32_t mode, const float* ring_inv, int32_t P, const int32_t* omegas,
                      int32_t n_ang, int32_t my, int32_t cp, float* data, float* part, int32_t ksplit,
                      int32_t accumulate, dfm_stream_t stream);
int dfm_pfblk_adjoint(const float* data, const int32_t* ptr, const uint16_t* pos, const void* val, int32_t mode,
                      const float* ring_inv, int32_t P, const int32_t* omegas, int32_t n_ang, int32_t my, int32_t cp,
                      int32_t kb, float* sino, int32_t sks, float alpha, dfm_stream_t stream);
/* Grouped launch of block-sparse products of one direction (forward != 0: the forward) over several PF matrices
 * (the slices of a multi-slice batch): descs is a device array of dfm_pfblk_desc, wg_map a device array of n_wg
 * int32 pairs (product, workgroup of that product); a product's workgroups are n_ang * n_yg * n_tg * ksplit,
 * decomposed (group of 4 column blocks fastest, then y group, k split, angle). Fields as in dfm_pfblk_forward /
 * dfm_pfblk_adjoint: src the sinogram (forward, first column of the product) or the data; out the data (direct = 1)
 * or the partial sums (forward), the sinogram (adjoint); n_kb / n_jb the stored blocks; n_cb the output column
 * blocks of 32 (forward n_jb, adjoint ceil(sks / 32)) and n_ch the contraction chunks; mt_n, n_yg the y tiling
 * (n_yg = ceil(ceil(my / 16) / 8) groups of mt_n 16-row tiles); n_tg = ceil(n_cb / 4); ch_per chunks per split. */
typedef struct dfm_pfblk_desc {
  const float* src;
  const int32_t* ptr;
  const uint16_t* pos;
  const void* val;
  const int32_t* omegas;
  const float* ring_inv;
  float* out;
  int32_t n_ang, my, cp, sks, kb, P;
  int32_t n_kb, n_jb;
  int32_t n_cb, n_ch;
  int32_t mt_n, n_yg, n_tg;
  int32_t ksplit, ch_per;
  int32_t direct, accumulate;
  float alpha;
  int32_t ncols;
  int32_t blk_local;
} dfm_pfblk_desc;
/* mode: 0 bf16, 1 bf16 residual, 2 fp16, + 256 mt_n when every descriptor has the same y-tile count mt_n (a form
 * with that count compiled in where there is one, for 5; the same results). */
int dfm_pfblk_grouped(int32_t forward, const void* descs, const int32_t* wg_map, int64_t n_wg, int32_t mode,
                      dfm_stream_t stream);
/* sizeof(dfm_pfblk_desc) as compiled into the library (for bindings that build the descriptor array by hand). */
int dfm_pfblk_desc_size(void);
/* The PF generated for the angles omegas[0 .. n_ang) of an orientation batch (k0 .. k0 + kb) in the block-sparse
 * form of dfm_pfblk_*, blocks indexed (a, kb_i, jb_i) (the generated mode, pf_generate.py:218-303 and
 * pf_generate.cl:146-454: candidates from the eta windows where a pole meets a ring's cone, |n.v| > 1 - cut sigma^2
 * - 1e-4, then the entries of dfm_pf_eval_dense evaluated for the candidates only). mode 0: counts[block] = the
 * block's non-zeros; mode 1: the entries at ptr[block] (the exclusive scan of the counts), in element order, values
 * as dfm_pfblk_fill. coords .. qw: as dfm_pf_eval_dense; frames (R, S, 9) float64: R(omega_node)^T applied to the
 * detector directions d0, d90 and the beam-normal p0 (the rows of the probed directions' cone); sin_th, cos_th: of
 * half the rings' two-theta; the eta nodes of bin c lie at eta0 + (c + x) deta, xlo <= x <= xhi (windows are
 * taken modulo a turn). */
int dfm_pfgen(int32_t mode, const double* coords, const double* poles, const int32_t* axis_start,
              const float* axis_count, const double* inv_sigma2, const double* norm, const float* qw,
              int32_t quad_weighted, const double* frames, const double* sin_th, const double* cos_th,
              const int32_t* omegas, int32_t n_ang, int32_t k0, int32_t kb, int32_t C, int32_t P, int32_t A, int32_t S,
              int32_t Se, float pf_cut, double eta0, double deta, double xlo, double xhi,
              int32_t* counts, const int32_t* ptr, uint16_t* pos, void* val, int32_t fp16, const float* ring_scale,
              dfm_stream_t stream);

/* ---------------------------------------------------------------- PF parameter gradients
 * (differentiable/gradients.py:20-189 PFGradient, pf_grad.cl:42-187 pf_param_vjp / rows_to_orientations)
 * Per (angle a of omegas, orientation k0 + k of the batch): part_k[(a kb + k) nqk + q] = the angle's share of the
 * gradient of <r, A c> with respect to the rotation increment (q 0..2) and the width (q 3) of the orientation, with
 * curv (a data-shaped curvature) also the Gauss-Newton block J^T diag(curv) J of (rotation, width) (q 4..13, upper
 * triangle row-major; nqk = 14, else 4); per workgroup (a, orientation block of 32): part_g[w (3 + P) + i] = the
 * shares of the eta offset (i 0), the axis tilt (1, 2) and the rings' 2 theta (3 + p) (a null dv_* skips its
 * group). sino (n_ang, My, sks): the projection of c for the angle list, orientation-fastest; r (R, My, C P): the
 * data-space derivative of the loss; dv_eta, dv_t0, dv_t1, dv_tth: float64 derivatives of the probed directions,
 * shaped as coords; the other arguments as dfm_pfgen. Candidates from the same eta windows as dfm_pfgen. */
int dfm_pf_grad(const double* coords, const double* poles, const int32_t* axis_start, const float* axis_count,
                const double* inv_sigma2, const double* norm, const float* qw, const double* frames,
                const double* sin_th, const double* cos_th, const double* dv_eta, const double* dv_t0,
                const double* dv_t1, const double* dv_tth, const float* sino, int32_t sks, const float* r,
                const float* curv, const int32_t* omegas, int32_t n_ang, int32_t k0, int32_t kb, int32_t My,
                int32_t C, int32_t P, int32_t A, int32_t S, int32_t Se, float pf_cut, double eta0, double deta,
                double xlo, double xhi, double* part_k, double* part_g, dfm_stream_t stream);
/* out_k[k nqk + q] = sum over a < n_ang of part_k (in order); out_g[i] = sum over the n_wg rows of part_g (a fixed
 * tree): the totals of dfm_pf_grad, deterministic. */
int dfm_pf_grad_reduce(const double* part_k, const double* part_g, int32_t n_ang, int32_t kb, int32_t nqk,
                       int64_t n_wg, int32_t n_glob, double* out_k, double* out_g, dfm_stream_t stream);
/* out[a] = sum over y < my, k < kn of u[a, y, k] v[a, y, k] in float64, for sinogram-shaped u, v (n_ang, my, sks)
 * (gradients.py:220-258 ProjGradient's shift_dot, proj_grad.cl:60-88: the PF adjoint of r against the derivative
 * projection of c). */
int dfm_angle_dots(const float* u, const float* v, int32_t n_ang, int32_t my, int32_t sks, int32_t kn, double* out,
                   dfm_stream_t stream);

/* ---------------------------------------------------------------- Fourier-slice projector (projector="fft")
 * (operators/fourier_radon.py:1-311, fourier_radon.cl fr_pad / fr_slices / fr_spread / fr_crop), every array
 * (points, channels) with the channels fastest; complex values as float pairs. pad: grid[gidx[p] kn + k] =
 * img_k[p kstride + k_off + k] deapod[p] (real; the caller zeroes the grid first). interp: spec[(ai H + j) kn + k] =
 * rowfac[r] sum_e w[e] G[col[e] kn + k] over the CSR row r = angles[ai] H + j (the polar samples' taps). spread:
 * G[t kn + k] = sum_e tw[e] colfac[r] Z[(apos[a] H + j) kn + k] over the transposed CSR row t, r = smp[e] = a H + j,
 * skipping angles with apos[a] < 0 (the transpose of interp, as a fixed-order gather). crop: out_k[p kstride + k_off
 * + k] (+)= scale Re(g[gidx[p] kn + k]) deapod[p]. */
int dfm_nufft_pad(const float* img_k, int32_t kstride, int32_t k_off, int32_t kn, const float* deapod,
                  const int32_t* gidx, int64_t npix, void* grid, dfm_stream_t stream);
int dfm_nufft_interp(const void* G, const int32_t* ptr, const int32_t* col, const float* w, const void* rowfac,
                     const int32_t* angles, int32_t n_ang, int32_t H, int32_t kn, void* spec, dfm_stream_t stream);
int dfm_nufft_spread(const void* Z, const int32_t* tptr, const int32_t* smp, const float* tw, const void* colfac,
                     const int32_t* apos, int32_t H, int32_t kn, int64_t n_grid, void* G, dfm_stream_t stream);
int dfm_nufft_crop(const void* g, const int32_t* gidx, const float* deapod, int64_t npix, int32_t kn, float scale,
                   float* out_k, int32_t kstride, int32_t k_off, int32_t accumulate, dfm_stream_t stream);
/* The SPDHG update of dfm_spdhg_update on a column block: x, z, T (rows, ld), dz (rows, ld_dz), columns < cols. */
int dfm_spdhg_update_2d(float* x, float* z, const float* T, const float* dz, int64_t rows, int32_t cols, int32_t ld,
                        int32_t ld_dz, float inv_p, float lam, dfm_stream_t stream);

/* ---------------------------------------------------------------- local basis (operators/local_forward_operator.py)
 * Patches of an orientation batch: box[p] = (y0, x0, h, w) int32 x 4, coff[p] the offset of its row-major
 * coefficients in the flat vector, mask a byte per coefficient (1: allowed); strips of W[p] bins per angle, stored
 * strips[a * sw + pw_off[p] + j] (sw = the sum of W over the batch), jp[jg] the patch of strip bin jg; ptab (R, P)
 * int32 pairs (off, tb): the strip's first detector bin at angle a and tb = t(x0, y0) - off as float bits; geo the
 * operator's geometry records (R x 8 floats). Angles: omegas[0 .. n_ang - 1].
 * strips (local_kernels.cl:38-117 local_radon_forward): strips = scale / A sum of footprint x allowed coefficients;
 *   max_area: the batch's largest patch (h w), which selects a workgroup per patch staged in LDS (up to 64 KB);
 *   n_reach: ceil(2 max_a hw_a / A_a), the most pixels a bin's footprint covers along a row (0: no bound). 
 * assemble (the sinogram rows local_kernels.cl:211-244 local_pf_forward assembles in local memory): the batch's
 *   orientation-fastest sinogram (n_ang, my, kstride), sino[as, y, k] = sum over the patches of orientation k
 *   (kp_ptr[k] .. kp_ptr[k + 1]) of their strip bin at line y, 0 elsewhere.
 * extract (the strip bins local_kernels.cl:289-319 local_pf_adjoint writes): strips = sino at (as, off + j, pk[p]),
 *   0 off the detector.
 * backproject (local_kernels.cl:120-174 local_radon_backward): for the batch's coefficients c_begin .. c_begin + n - 1
 *   (pid: the patch of each) acc = sum_a (w / A)_a sum of footprint x strip bins; mode 0: out = acc (0 where not
 *   allowed); mode 1: the SPDHG primal update with dz = acc (as dfm_spdhg_update). n_taps 2 .. 4. blk: (patch,
 *   first pixel) per 256 pixels (n_blk of them) for the form that stages a patch's strips in LDS (n_ang (32 + 4
 *   w_max) bytes up to 64 KB, w_max the batch's widest strip); pid (the patch of each coefficient) for the other. */
int dfm_local_strips(const float* coef, const uint8_t* mask, const int32_t* box, const int64_t* coff,
                     const int32_t* W, const int32_t* pw_off, int32_t sw, const int32_t* jp, const int32_t* ptab,
                     int32_t n_patches, int32_t max_area, int32_t n_reach, const float* geo, const int32_t* omegas,
                     int32_t n_ang, float scale, float* strips, dfm_stream_t stream);
int dfm_local_assemble(const float* strips, int32_t sw, const int32_t* pw_off, const int32_t* W, const int32_t* ptab,
                       int32_t n_patches, const int32_t* kp_ptr, int32_t k_batch, const int32_t* omegas,
                       int32_t n_ang, int32_t my, int32_t kstride, float* sino, dfm_stream_t stream);
int dfm_local_extract(const float* sino, int32_t n_ang, int32_t my, int32_t kstride, const int32_t* omegas,
                      const int32_t* ptab, int32_t n_patches, const int32_t* pk, const int32_t* jp,
                      const int32_t* pw_off, int32_t sw, float* strips, dfm_stream_t stream);
int dfm_local_backproject(const float* strips, int32_t sw, const int32_t* pw_off, const int32_t* W,
                          const int32_t* ptab, int32_t n_patches, int32_t w_max, const float* geo,
                          const int32_t* omegas, int32_t n_ang, const int32_t* box, const int64_t* coff,
                          const uint8_t* mask, const int32_t* pid, int64_t c_begin, int64_t n, const int32_t* blk,
                          int32_t n_blk, int32_t n_taps, int32_t mode, float* out, float* x, float* z, const float* T,
                          float inv_p, float lam, dfm_stream_t stream);

/* ---------------------------------------------------------------- SPDHG-Poisson (optimization/spdhg_poisson.py)
 * Periodic arrays w, g, bg: element i uses a[i % period]; period 0 means the constant 1, 1, 0.
 * Dual update on the rows of a subset (spdhg_poisson.py:69-91 dual_update_sub), float64 arithmetic:
 * d holds A x on entry and y' - y on exit; y <- y'. */
int dfm_dual_update_sub(float* y, float* d, const float* sig, const float* b, const float* w, int64_t pw,
                        const float* g, int64_t pg, const float* bg, int64_t pbg, float eps, const int32_t* omegas,
                        int32_t n_sub, int64_t row_len, dfm_stream_t stream);
/* Backprojection of the subset's dual change fused with the primal update (spdhg_poisson.py:95-106 spdhg_update
 * applied to the output of radon_backward_k): dz = backward(sino); z += dz; x = T > 0 ? max(0, x - T (z + inv_p dz
 * + lam)) : 0, on the channel-fastest arrays x, z, T (Npix, kstride), channels k_off .. k_off + k_n - 1. */
int dfm_radon_backward_spdhg(const float* sino, const float* geo, int32_t nx, int32_t ny, int32_t ns, int32_t n_angles,
                             int32_t kstride, int32_t k_off, int32_t k_n, int32_t skstride, int32_t ntap, float* x,
                             float* z, const float* T, float inv_p, float lam, dfm_stream_t stream);
/* The plain update (spdhg_poisson.py:95-106) on n elements: z += dz; x = T > 0 ? max(0, x - T (z + inv_p dz + lam)) : 0 */
int dfm_spdhg_update(float* x, float* z, const float* T, const float* dz, float inv_p, float lam, int64_t n,
                     dfm_stream_t stream);
/* Primal steps (spdhg_poisson.py:252-272 tinit / tmin / tfinal) on channel-fastest (Npix, kstride) arrays,
 * channels < kvalid (the others 0): tinit T = inf; tmin T = min(T, c / colsum) where supp[p] and colsum > 0
 * (supp NULL: every pixel); tfinal T = 0 where T is inf. */
int dfm_tinit(float* T, int64_t npix, int32_t kstride, dfm_stream_t stream);
int dfm_tmin(float* T, const float* colsum, const uint8_t* supp, int64_t npix, int32_t kstride, int32_t kvalid,
             float c, dfm_stream_t stream);
int dfm_tfinal(float* T, int64_t npix, int32_t kstride, int32_t kvalid, dfm_stream_t stream);
/* sig = (w != 0 && rowsum > 0) ? sigma0 / rowsum : 0 (pdhg_poisson.py:79-84 dual_steps). */
int dfm_dual_steps(float* sig, const float* rowsum, const float* w, int64_t pw, float sigma0, int64_t n,
                   dfm_stream_t stream);
/* u = (w != 0) ? 1 : 0 (pdhg_poisson.py:113-116 set_mask_ones). */
int dfm_set_mask_ones(float* u, const float* w, int64_t pw, int64_t n, dfm_stream_t stream);
/* Channel-fastest ones on the support, 0 outside and on channels >= kvalid (fista_poisson.py:219-223). */
int dfm_support_ones_k(float* x, const uint8_t* supp, int64_t npix, int32_t kstride, int32_t kvalid,
                       dfm_stream_t stream);
/* x = 0 outside the support, where x < 0 and on channels >= kvalid (fista_poisson.py:277-281 project_support). */
int dfm_project_support_k(float* x, const uint8_t* supp, int64_t npix, int32_t kstride, int32_t kvalid,
                          dfm_stream_t stream);
/* float64 partial sums, one per work-group, into partial[0 .. *n_partial) (the caller sums them):
 * deviance: sum over w != 0 of w/g (m - b + b log(b/m)), m = max(mu, 0) + bg + eps g (fista_poisson.py:72-109);
 * mean_counted: partial[i] = sum b, partial[n_partial + i] = count (pdhg_poisson.py:97-111);
 * dot: sum a*b (single_phase_forward_operator.py:2168-2189 vdot). n_partial receives the number of partials
 * (at most DFM_MAX_PARTIALS). */
#define DFM_MAX_PARTIALS 2048
int dfm_deviance_partial(const float* mu, const float* b, const float* w, int64_t pw, const float* g, int64_t pg,
                         const float* bg, int64_t pbg, float eps, int64_t n, double* partial, int32_t* n_partial,
                         dfm_stream_t stream);
int dfm_mean_counted_partial(const float* b, const float* w, int64_t pw, int64_t n, double* partial,
                             int32_t* n_partial, dfm_stream_t stream);
/* Allocation-free deviance: out[0] (a device double) = the deviance of dfm_deviance_partial, its partial sums added
 * in index order on the device. No counterpart in the reference: for callers that cannot allocate (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc). The partial sums live in a file-scope device
 * array of DFM_MAX_PARTIALS doubles (16 KiB per GPU, held until the process ends), which bounds every input size;
 * no hipMalloc, hipFree or allocating library call. Calls that overlap on one GPU share that array and race. */
int dfm_deviance_noalloc(const float* mu, const float* b, const float* w, int64_t pw, const float* g, int64_t pg,
                         const float* bg, int64_t pbg, float eps, int64_t n, double* out, dfm_stream_t stream);
/* The EM pieces of the Poisson term for OSEM / MLEM (optimization/poisson_term.py:39-47 pt_ratio, pt_wg;
 * osem_poisson.py:69-73 osem_update), over n data points or n coefficients; w, g, bg periodic as in
 * dfm_deviance_partial (a zero weight: the point is not fitted, e.g. its angle is in no subset):
 *   dfm_poisson_em_ratio:         out = w / g * b / mu, mu = max(pred, 0) + bg + eps g  (0 where w = 0)
 *   dfm_poisson_weight_over_gain: out = w / g
 *   dfm_em_update:                x = 0 where mask[i % pm] == 0 (pm 0: no mask), else x back / (sens + lam) where
 *                                 sens > 0, unchanged where sens <= 0 (lam >= 0) */
int dfm_poisson_em_ratio(const float* pred, const float* b, const float* w, int64_t pw, const float* g, int64_t pg,
                         const float* bg, int64_t pbg, float eps, int64_t n, float* out, dfm_stream_t stream);
int dfm_poisson_weight_over_gain(const float* w, int64_t pw, const float* g, int64_t pg, int64_t n, float* out,
                                 dfm_stream_t stream);
int dfm_em_update(float* x, const float* back, const float* sens, const uint8_t* mask, int64_t pm, float lam,
                  int64_t n, dfm_stream_t stream);
/* PDHG pieces (pdhg_poisson.py:69-86 primal_update, average): x' = T > 0 ? max(0, x - T (grad + lam)) : 0,
 * xb = 2 x' - x, x = x' over n coefficients; and ax = (axb + ax) / 2 over n data points (A x of the new iterate from
 * A x_bar). */
int dfm_pdhg_primal_update(float* x, float* xb, const float* grad, const float* T, float lam, int64_t n,
                           dfm_stream_t stream);
int dfm_average(float* ax, const float* axb, int64_t n, dfm_stream_t stream);
/* FISTA in the SQS metric (optimization/fista_poisson.py:61-282 kernels, the API coefficient layout; supp periodic
 * over npix pixels, npix 0: no support; float64 partial sums as dfm_deviance_partial, NV sums stored NV-major):
 *   dfm_fista_value_grad_partial (poisson_value_grad): mode 0 the deviance, mode 1 its quadratic approximation at the
 *     data; write_grad: mu <- df/dmu;
 *   dfm_fista_curvature_partial (times_curvature): u <- c u with the Poisson (mode 0) or start (mode 1) curvature;
 *     sum c u^2;
 *   dfm_fista_sums_partial: what 0, sum of x over the support (support_sum); what 1, sum w b and sum w x
 *     (weighted_sums);
 *   dfm_fista_elementwise: op 0 invert_metric (out = 1/a where a > 0 in the support), 1 support_ones, 2
 *     project_support (out in place), 3 metric_step (out = max(0, a - th D (c + lam)) where D > 0), 4 extrapolate
 *     (out = a + th (a - c));
 *   dfm_fista_step_sums_partial (step_sums): <grad, xn - y>, sum (xn - y)^2 / D over D > 0, sum xn. */
int dfm_fista_value_grad_partial(float* mu, const float* b, const float* w, int64_t pw, const float* g, int64_t pg,
                                 const float* bg, int64_t pbg, float eps, int32_t write_grad, int32_t mode,
                                 float floor_counts, int64_t n, double* partial, int32_t* n_partial,
                                 dfm_stream_t stream);
int dfm_fista_curvature_partial(float* u, const float* mu, const float* b, const float* w, int64_t pw, const float* g,
                                int64_t pg, const float* bg, int64_t pbg, float floor_counts, int32_t mode, int64_t n,
                                double* partial, int32_t* n_partial, dfm_stream_t stream);
int dfm_fista_sums_partial(const float* x, const uint8_t* supp, int64_t npix, const float* b, const float* w,
                           int64_t pw, int32_t what, int64_t n, double* partial, int32_t* n_partial,
                           dfm_stream_t stream);
int dfm_fista_elementwise(int32_t op, float* out, const float* a, const float* c, const float* D,
                          const uint8_t* supp, int64_t npix, float th, float lam, int64_t n, dfm_stream_t stream);
int dfm_fista_step_sums_partial(const float* xn, const float* y, const float* grad, const float* D, int64_t n,
                                double* partial, int32_t* n_partial, dfm_stream_t stream);
int dfm_dot_partial(const float* a, const float* b, int64_t n, double* partial, int32_t* n_partial,
                    dfm_stream_t stream);
/* Smoothed isotropic TV of coefficient images c (K, ny, nx) float32, a loss term of the differentiable refinement
 * (differentiable/data_model.py:20-64 is the intensity model it is used with): partial[b] (b < *n_partial <=
 * DFM_MAX_PARTIALS, float64) sum to sum sqrt(dx^2 + dy^2 + eps) over forward differences (0 at the last column and
 * row); grad (K, ny, nx) = its gradient with respect to c. */
int dfm_smooth_tv(const float* c, float* grad, int32_t K, int32_t ny, int32_t nx, float eps, double* partial,
                  int32_t* n_partial, dfm_stream_t stream);
/* Total-variation block of SPDHG (spdhg_poisson.py:108-193 tv_dual_dense / tv_adjoint_dense / tv_value_dense) on the
 * channel-fastest arrays (Npix, kstride), pixel p = ix + nx iy; a forward difference (right, down) is free when T > 0
 * at both ends.
 * dual:  per pixel s = ps / (4 max over its free differences of (T_j + T_j')); q = proj_{|q| <= lam}(q + s D x) on
 *        the free differences, 0 elsewhere; dq = q' - q.
 * adjoint_update: dz = D^T dq; z += dz; x = T > 0 ? max(0, x - T (z + inv_p dz + lam)) : 0.
 * value: float64 partial sums of |D x| over the free differences (n_partial <= DFM_MAX_PARTIALS). */
int dfm_tv_dual_k(const float* x, const float* T, float* qx, float* qy, float* dqx, float* dqy, int32_t nx,
                  int32_t ny, int32_t kstride, float lam, float ps, dfm_stream_t stream);
int dfm_tv_adjoint_update_k(const float* dqx, const float* dqy, float* x, float* z, const float* T, int32_t nx,
                            int32_t ny, int32_t kstride, float inv_p, float lam, dfm_stream_t stream);
int dfm_tv_value_partial(const float* x, const float* T, int32_t nx, int32_t ny, int32_t kstride, double* partial,
                         int32_t* n_partial, dfm_stream_t stream);
/* The same block on a flat coefficient vector whose neighbours are index arrays (right, down, left, up; -1: none):
 * the local basis's patches (spdhg_poisson.py:195-236 tv_dual_local / tv_adjoint_local / tv_value_local; the
 * adjoint fused with the primal update as above). n elements. */
int dfm_tv_dual_nbr(const float* x, const float* T, float* qx, float* qy, float* dqx, float* dqy, const int32_t* right,
                    const int32_t* down, float lam, float ps, int64_t n, dfm_stream_t stream);
int dfm_tv_adjoint_update_nbr(const float* dqx, const float* dqy, const int32_t* left, const int32_t* up, float* x,
                              float* z, const float* T, float inv_p, float lam, int64_t n, dfm_stream_t stream);
int dfm_tv_value_nbr_partial(const float* x, const float* T, const int32_t* right, const int32_t* down, int64_t n,
                             double* partial, int32_t* n_partial, dfm_stream_t stream);
/* a *= s (single_phase_forward_operator.py:2169, 2192 x *= 1/|x|). */
int dfm_scale(float* a, float s, int64_t n, dfm_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* DIFFRACTOM_HIP_H */
