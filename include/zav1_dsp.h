/*
 * zav1_dsp.h — C-ABI of the MI355X (gfx950) backend for zenav1-svt's
 * data-parallel intra-encode DSP kernels.
 *
 * Two forms share one set of HIP kernels:
 *
 *  1. The PLAIN form: a versioned function table (zav1_dsp_v1) whose entries
 *     mirror, argument for argument, the reference's per-kernel entry points
 *     that its encoder calls through archmage tier dispatch.  Pointers are
 *     HOST pointers (caller-owned slices, never retained); each call copies
 *     its inputs to HBM, runs one launch on the calling thread's stream and
 *     copies the result back.  This is what a drop-in FFI crate binds.
 *
 *  2. The BATCHED form: zav1_*_batch() entry points taking DEVICE pointers to
 *     many jobs at once, ordered on a caller-supplied HIP stream.  The loop
 *     filter, candidate-RD and serving drivers build on these.
 *
 * Status codes (map onto the reference's EncodeError, typ/error.rs:22-40):
 *   ZAV1_OK 0, ZAV1_EINVAL bad argument, ZAV1_ENOTSUP unsupported combination
 *   (e.g. ADST on a 32/64 dimension — the reference's `false` return),
 *   ZAV1_ENOMEM device allocation / HBM budget, ZAV1_ECANCELED, ZAV1_EIO HIP
 *   runtime failure.  No exception crosses the ABI.  Every entry point is
 *   thread-safe; the plain form uses the per-thread default stream.
 *
 * Enumerations follow the reference (C SVT-AV1) order:
 *   tx_size: rust/crates/svtav1-types/src/transform.rs:7-29 (19 sizes)
 *   tx_type: rust/crates/svtav1-types/src/transform.rs:84-101 (16 types)
 */
#ifndef ZAV1_DSP_H
#define ZAV1_DSP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZAV1_ABI_VERSION 3u

#define ZAV1_OK 0
#define ZAV1_EINVAL (-22)
#define ZAV1_ENOTSUP (-95)
#define ZAV1_ENOMEM (-12)
#define ZAV1_ECANCELED (-125)
#define ZAV1_EIO (-5)

typedef void *zav1_stream_t; /* a hipStream_t; NULL = the per-thread default stream */

/* ------------------------------------------------------------------------ */
/* Runtime plumbing (lets hosts without their own GPU runtime drive HBM).    */
/* ------------------------------------------------------------------------ */
int zav1_init(int device);                   /* select + initialise a device */
int zav1_device_count(void);
int zav1_malloc(void **dptr, size_t bytes);  /* device memory */
int zav1_free(void *dptr);
int zav1_host_alloc(void **hptr, size_t bytes); /* pinned host memory */
int zav1_host_free(void *hptr);
int zav1_memcpy_h2d(void *dst, const void *src, size_t bytes, zav1_stream_t s);
int zav1_memcpy_d2h(void *dst, const void *src, size_t bytes, zav1_stream_t s);
int zav1_memcpy_d2d(void *dst, const void *src, size_t bytes, zav1_stream_t s);
int zav1_memset(void *dst, int value, size_t bytes, zav1_stream_t s);
int zav1_stream_create(zav1_stream_t *s);
int zav1_stream_destroy(zav1_stream_t s);
int zav1_stream_sync(zav1_stream_t s);
int zav1_device_sync(void);
/* Event timing on a stream: record start, run work, record stop, elapsed. */
int zav1_event_create(void **ev);
int zav1_event_destroy(void *ev);
int zav1_event_record(void *ev, zav1_stream_t s);
int zav1_event_elapsed_ms(void *start, void *stop, float *ms);
/* Work enqueued on s after this call waits for ev's recorded work (joins two
 * streams on the device). */
int zav1_stream_wait_event(zav1_stream_t s, void *ev);
/* 1 when ev's recorded work has completed, 0 while it is pending. */
int zav1_event_query(void *ev);
const char *zav1_last_error(void);

/* ------------------------------------------------------------------------ */
/* Quantizer rows (host side, no device work).                               */
/* ------------------------------------------------------------------------ */
typedef struct zav1_quant_row {
  int32_t zbin[2], round[2], quant[2], quant_shift[2];
  int32_t round_fp[2], quant_fp[2], dequant[2];
} zav1_quant_row;

/* Replaces build_quant_table_bd (rust/crates/svtav1-encoder/src/quant.rs:180),
 * sharpness 0.  bd in {8, 10}.  [0] = DC, [1] = AC. */
int zav1_build_quant_row(int qindex, int bd, zav1_quant_row *row);
/* The same row with the quantizer sharpness adjustment (-7..7; 0 = the row
 * above).  Replaces build_quant_table_bd_sharp
 * (rust/crates/svtav1-encoder/src/quant.rs:195, apply_quant_sharpness_factors
 * :105). */
int zav1_build_quant_row_sharp(int qindex, int bd, int sharpness, zav1_quant_row *row);

/* ------------------------------------------------------------------------ */
/* PLAIN form: function table, host pointers, one launch per call.           */
/* ------------------------------------------------------------------------ */
typedef struct zav1_dsp_v1 {
  uint32_t abi_version; /* ZAV1_ABI_VERSION */
  uint32_t backend;     /* 1 = HIP gfx950 */

  /* fwd_txfm2d_dispatch (rust/crates/svtav1-dsp/src/txfm_dispatch.rs:18).
   * in: w x h residual with row stride `stride`; out: w x h raster.
   * Returns 1 on success, 0 when no kernel exists (the reference's false),
   * negative on runtime failure. */
  int (*fwd_txfm2d)(const int32_t *in, int32_t *out, size_t stride, uint8_t tx_size,
                    uint8_t tx_type);

  /* inv_txfm2d_dispatch_bd (rust/crates/svtav1-dsp/src/txfm_dispatch.rs:55).
   * in: coefficients with row stride `stride` (top-left 32x32 read for 64-dim);
   * out: residual with row stride `stride`. bd in {8, 10}. 1 / 0 / negative. */
  int (*inv_txfm2d_bd)(const int32_t *in, int32_t *out, size_t stride, uint8_t tx_size,
                       uint8_t tx_type, uint8_t bd);

  /* quantize_fp_raster (rust/crates/svtav1-dsp/src/quant_coding.rs:58), bd8:
   * `rounding` is the pre-shifted _fp round row. */
  int (*quantize_fp_raster)(const int32_t *coeffs, int32_t *qcoeff, int32_t *dqcoeff, size_t n,
                            const int32_t rounding[2], const int32_t quant_fp[2],
                            const int32_t dequant[2], int32_t log_scale);

  /* quantize_b_raster (rust/crates/svtav1-dsp/src/quant_coding.rs:336), bd8:
   * zbins/round pre-shifted by log_scale as the caller (quant.rs:255) does. */
  int (*quantize_b_raster)(const int32_t *coeffs, int32_t *qcoeff, int32_t *dqcoeff, size_t n,
                           const int32_t zbins[2], const int32_t round[2],
                           const int32_t quant[2], const int32_t quant_shift[2],
                           const int32_t dequant[2], int32_t log_scale);

  /* sse_i32 / sq_sum_i32 (rust/crates/svtav1-dsp/src/residual.rs:245/:364):
   * wrapping u64 result through *out. b == NULL gives sq_sum. */
  int (*sse_i32)(const int32_t *a, const int32_t *b, size_t n, uint64_t *out);

  /* aom_hadamard_{4x4,8x8,16x16,32x32} + aom_satd
   * (rust/crates/svtav1-dsp/src/hadamard.rs:691,719,885,913,942).
   * src_diff: n x n i16 residual, row stride `stride`; coeff (n*n, may be
   * NULL) receives the reference's coefficient order; *satd the SATD. */
  int (*hadamard_satd)(const int16_t *src_diff, size_t stride, int n, int32_t *coeff,
                       int32_t *satd);

  /* ---- ABI 2 (appended: an ABI 1 caller reads the prefix unchanged) ---- */
  /* In the entries below 8-bit samples are uint8_t and 10-bit uint16_t; the
   * void * sample buffers follow bd (8: uint8_t, 10: uint16_t).  Outputs are
   * written only where the reference writes. */

  /* residual_i32 (rust/crates/svtav1-dsp/src/residual.rs:26) and its 16-bit
   * twin: out[r*w + c] = src - pred, w x h (1..128). */
  int (*residual_i32)(const uint8_t *src, size_t src_stride, const uint8_t *pred, size_t pred_stride,
                      size_t w, size_t h, int32_t *out);
  int (*residual_i32_u16)(const uint16_t *src, size_t src_stride, const uint16_t *pred, size_t pred_stride,
                          size_t w, size_t h, int32_t *out);
  /* recon_add_clamp (residual.rs:140): out[r*w + c] = clip(pred + inv, 0, 255);
   * the 16-bit twin clips to 2^bd - 1 (highbd_clip_pixel_add, hbd.rs:145). */
  int (*recon_add_clamp)(const uint8_t *pred, size_t pred_stride, const int32_t *inv, size_t w, size_t h,
                         uint8_t *out);
  int (*recon_add_clamp_u16)(const uint16_t *pred, size_t pred_stride, const int32_t *inv, size_t w, size_t h,
                             uint16_t *out, int bd);
  /* variance::sse (variance.rs:60) / full_distortion_kernel16_bits
   * (hbd.rs:936): sum of squared differences over w x h into *out. */
  int (*sse_u8)(const uint8_t *a, size_t a_stride, const uint8_t *b, size_t b_stride, size_t w, size_t h,
                uint64_t *out);
  int (*sse_u16)(const uint16_t *a, size_t a_stride, const uint16_t *b, size_t b_stride, size_t w, size_t h,
                 uint64_t *out);
  /* predict_dc / _v / _h / _smooth / _smooth_v / _smooth_h / _paeth
   * (intra_pred.rs:18-154; hbd.rs:169-361): mode 0 DC (has_above / has_left
   * as the reference's flags), 1 V, 2 H, 9 SMOOTH, 10 SMOOTH_V, 11 SMOOTH_H,
   * 12 PAETH (top_left used).  above[0..w), left[0..h); w, h in 4..64. */
  int (*intra_pred)(int mode, void *dst, size_t dst_stride, const void *above, const void *left, int top_left,
                    int w, int h, int has_above, int has_left, int bd);
  /* dr_predictor_edged (intra_pred.rs:810; hbd.rs:436-700): final (filtered,
   * upsampled) edge buffers of edge_len samples with sample 0 of the row /
   * column at `origin` and the corner at origin - 1; angle = a directional
   * mode's nominal angle (45, 67, 90, 113, 135, 157, 180, 203) + 3 * delta,
   * |delta| <= 3; upsampling only for w + h <= 16. */
  int (*dr_predictor_edged)(void *dst, size_t dst_stride, const void *above, const void *left, size_t edge_len,
                            size_t origin, int upsample_above, int upsample_left, int w, int h, int angle, int bd);
  /* predict_filter_intra (intra_pred.rs:988; hbd.rs:784): above[0] the
   * top-left sample, above[1..=w] the row above; mode 0..4; w, h <= 32. */
  int (*predict_filter_intra)(void *dst, size_t dst_stride, const void *above, const void *left, int w, int h,
                              int mode, int bd);
  /* cfl_predict_lbd (intra_pred.rs:1137; hbd.rs:889): ac_q3 with row stride
   * 32 (CFL_BUF_LINE), pred the DC prediction; dst = clip(pred + alpha*ac). */
  int (*cfl_predict)(const int16_t *ac_q3, const void *pred, size_t pred_stride, void *dst, size_t dst_stride,
                     int alpha_q3, int w, int h, int bd);
  /* lpf_{horizontal,vertical}_{4,6,8,14} (loop_filter.rs:307-405) and the
   * _hbd twins (hbd.rs:1335-1433): filters the 4 lines of the call at sample
   * offset `off` of buf (buf_len samples) in place; thresholds as
   * lf_thresholds (loop_filter.rs:50) gives them at 8 bits. */
  int (*lpf)(void *buf, size_t buf_len, size_t off, size_t pitch, int vertical, int taps, uint8_t mblim,
             uint8_t lim, uint8_t hev_thr, int bd);
  /* cdef_find_dir (cdef.rs:143): returns the direction 0..7 of the 8x8 u16
   * block (negative on error), *var its variance. */
  int (*cdef_find_dir)(const uint16_t *img, size_t stride, int coeff_shift, int32_t *var);
  /* cdef_filter_block (cdef.rs:236) / cdef_filter_block_hbd (hbd.rs:1533):
   * `in` the padded u16 source (row stride 144 = CDEF_BSTRIDE, in_len
   * samples, 3-row / 8-column border around the block at in + ioff); bsize
   * 0 4x4, 1 4x8, 2 8x4, 3 8x8; rows stepped by subsampling_factor. */
  int (*cdef_filter_block)(void *dst, size_t dst_stride, const uint16_t *in, size_t in_len, size_t ioff,
                           int pri_strength, int sec_strength, int dir, int pri_damping, int sec_damping,
                           int bsize, int coeff_shift, size_t subsampling_factor, int bd);
  /* compute_stats (restoration.rs:232) / compute_stats_hbd (:1738): Wiener
   * M (win^2) and H (win^4) over [h_start, h_end) x [v_start, v_end), dgd
   * with win/2 samples of margin; window sides <= 384. */
  int (*compute_stats)(int wiener_win, const void *dgd, size_t dgd_len, size_t dgd_origin, size_t dgd_stride,
                       const void *src, size_t src_len, size_t src_origin, size_t src_stride, int h_start,
                       int h_end, int v_start, int v_end, int64_t *M, int64_t *H, int bd);

  /* ---- ABI 3 (appended) ---- */
  /* quantize_fp_hbd / quantize_b_hbd (rust/crates/svtav1-encoder/src/
   * quant.rs:335 / :391): the bd > 8 quantizers on n (<= 1024) coefficients
   * in raster order with their scan (a permutation of 0..n-1), the row as
   * zav1_build_quant_row(_sharp) builds it (log_scale applied inside), no
   * INT16 clamp.  Returns the eob (>= 0) or a negative status. */
  int (*quantize_fp_hbd)(const int32_t *coeffs, const uint16_t *scan, size_t n, const zav1_quant_row *t,
                         int32_t log_scale, int32_t *qcoeff, int32_t *dqcoeff);
  int (*quantize_b_hbd)(const int32_t *coeffs, const uint16_t *scan, size_t n, const zav1_quant_row *t,
                        int32_t log_scale, int32_t *qcoeff, int32_t *dqcoeff);
  /* compute_cdef_dist_8bit / _16bit (rust/crates/svtav1-dsp/src/cdef.rs:1192
   * / :1160): SSE of the filtered blocks `packed` (back to back, bsize 0 4x4,
   * 1 4x8, 2 8x4, 3 8x8) against the source plane (plane_len samples, block
   * b at plane_off + (by << log2 bh) * pstride + (bx << log2 bw) for dlist's
   * (by, bx) byte pairs), rows stepped by subsampling_factor, >>
   * 2 * coeff_shift, into *out.  Samples are u8 at bd 8, u16 at bd 10. */
  int (*cdef_dist)(const void *plane, size_t plane_len, size_t plane_off, size_t pstride, const void *packed,
                   size_t packed_len, const uint8_t *dlist, size_t n_dlist, int bsize, int coeff_shift,
                   size_t subsampling_factor, int bd, uint64_t *out);
} zav1_dsp_v1;

/* Returns the table for `backend` (1 = HIP), or NULL if no GPU is usable. */
const zav1_dsp_v1 *zav1_dsp_get(int backend);

/* ------------------------------------------------------------------------ */
/* BATCHED form: device pointers, stream-ordered, many jobs per launch.      */
/* Jobs are packed back to back (job j at offset j * w * h).                 */
/* ------------------------------------------------------------------------ */

/* n_jobs forward 2D transforms of one tx_size; types[j] per job (device). */
int zav1_fwd_txfm2d_batch(const int32_t *d_in, int32_t *d_out, const uint8_t *d_types,
                          int tx_size, int n_jobs, zav1_stream_t s);
/* n_jobs inverse 2D transforms (residual out, HIGHBD_WRAPLOW at bd). */
int zav1_inv_txfm2d_batch(const int32_t *d_in, int32_t *d_out, const uint8_t *d_types,
                          int tx_size, int n_jobs, int bd, zav1_stream_t s);
/* Quantize n_jobs blocks of n_per_job coefficients; d_qidx[j] selects the
 * quantizer row of qindex d_qidx[j] at `bd`.  kind 0 = fp, 1 = b. */
int zav1_quantize_batch(const int32_t *d_coeff, int32_t *d_qcoeff, int32_t *d_dqcoeff,
                        const uint8_t *d_qidx, int n_per_job, int n_jobs, int log_scale, int bd,
                        int kind, zav1_stream_t s);
/* d_out[j] = sum over job j of (a - b)^2 (wrapping u64); d_b NULL = sq_sum. */
int zav1_sse_i32_batch(const int32_t *d_a, const int32_t *d_b, uint64_t *d_out, int n_per_job,
                       int n_jobs, zav1_stream_t s);
/* n x n Hadamard + SATD of packed i16 tiles; d_coeff may be NULL. */
int zav1_satd_batch(const int16_t *d_res, int32_t *d_coeff, int32_t *d_satd, int n, int n_jobs,
                    zav1_stream_t s);

/* Coefficient rate tables (1/512 bit) the RDOQ trellis prices with, in
 * the reference's layout: one (txs_ctx, plane) of TxbCosts
 * (rust/crates/svtav1-encoder/src/quant.rs:459, C LvMapCoeffCost), one
 * (eob_multi_size, plane) of EobCosts (:469, LvMapEobCost), and a frame's
 * complete set, CoeffCostTables (:475): txb[txs_ctx * 2 + plane],
 * eob[eob_multi_size][plane].  Built by the caller from its CDFs
 * (build_coeff_cost_tables, :514). */
typedef struct {
  int32_t txb_skip_cost[13][2];
  int32_t base_eob_cost[4][3];
  int32_t base_cost[42][8];
  int32_t eob_extra_cost[9][2];
  int32_t dc_sign_cost[3][2];
  int32_t lps_cost[21][26];
} zav1_txb_costs;
typedef struct {
  int32_t eob_cost[2][11];
} zav1_eob_costs;
typedef struct {
  zav1_txb_costs txb[10];
  zav1_eob_costs eob[7][2];
} zav1_coeff_costs;

/* One block of an RDOQ batch: the per-call fields of OptimizeCtx
 * (quant.rs:887). */
typedef struct {
  int64_t rdmult;       /* RDCOST multiplier (rdoq_rdmult_full, quant.rs:923) */
  uint32_t coeff;       /* element offset of the block's packed coefficients */
  uint8_t tx_class;     /* 0 2D, 1 horizontal, 2 vertical (TX_TYPE_TO_CLASS) */
  uint8_t plane_type;   /* 0 luma, 1 chroma: selects the cost tables */
  uint8_t txb_skip_ctx; /* < 13 */
  uint8_t dc_sign_ctx;  /* < 3 */
  uint8_t sharpness;    /* sharpness_flag: no eob shortening, no skip */
  uint8_t pad[7];
} zav1_rdoq_job;

/* RDOQ: the trellis requantization of n_jobs blocks of tx_size after fp
 * quantization at (qindex, bd), in place on d_qcoeff / d_dqcoeff (packed,
 * coded size <= 32x32, raster) with d_eob[j] in (scan order) and out.
 * Replaces optimize_b (rust/crates/svtav1-encoder/src/quant.rs:1312) on the
 * path without quantization matrices.  d_costs is the frame's
 * zav1_coeff_costs; d_scans holds tx_size's three scans (zav1_scan_order
 * classes 0, 1, 2, each of the coded block's length).  Blocks with eob 0 are
 * left as they are.  cut_off_num / cut_off_denum: rdoq_cutoffs (:1551),
 * 0 / 0 for the full trellis.  A job with a field out of range is left as it
 * is.  d_dqcoeff must hold the fp quantizer's dequantized values,
 * sign((|q| * dequant) >> log_scale), which the trellis recomputes from the
 * levels.  d_eob[j] must be the quantizer's eob: positions at or past it
 * hold level 0 (a block whose level at scan[eob - 1] is 0 is left as it
 * is).  No allocation: one wave per block, in on-chip memory. */
int zav1_rdoq_batch(int tx_size, int qindex, int bd, const int32_t *d_tcoeff, int32_t *d_qcoeff,
                    int32_t *d_dqcoeff, uint16_t *d_eob, const zav1_rdoq_job *d_jobs, int n_jobs,
                    const zav1_coeff_costs *d_costs, const uint16_t *d_scans, uint32_t cut_off_num,
                    uint32_t cut_off_denum, zav1_stream_t s);

/* The same trellis with a caller workspace (d_work, work_bytes; both from
 * the caller, nothing allocated): a batch of at least the tx size's
 * threshold of jobs (zav1_rdoq_lane_min) whose workspace holds
 * zav1_rdoq_workspace(tx_size, n_jobs) bytes runs in the lane form (one
 * lane per block over scan-ordered pairs gathered into the workspace: the
 * throughput form for large batches); otherwise in the wave form of
 * zav1_rdoq_batch.  Both forms give identical results. */
int zav1_rdoq_batch_ws(int tx_size, int qindex, int bd, const int32_t *d_tcoeff, int32_t *d_qcoeff,
                       int32_t *d_dqcoeff, uint16_t *d_eob, const zav1_rdoq_job *d_jobs, int n_jobs,
                       const zav1_coeff_costs *d_costs, const uint16_t *d_scans, uint32_t cut_off_num,
                       uint32_t cut_off_denum, void *d_work, size_t work_bytes, zav1_stream_t s);
/* Workspace bytes of the lane form: n_jobs x (8.5 bytes per coded
 * coefficient of tx_size). */
size_t zav1_rdoq_workspace(int tx_size, int n_jobs);
/* The lane form's minimum batch for tx_size: set to n_jobs when n_jobs >= 0
 * (0: the lane form whenever a workspace is given) and returns the previous
 * value, or ZAV1_EINVAL.  Process-wide.  No counterpart in the reference
 * (a scheduling choice of this library). */
int zav1_rdoq_lane_min(int tx_size, int n_jobs);

/* One block of a coefficient-rate batch (cost_coeffs_txb's arguments,
 * rust/crates/svtav1-encoder/src/leaf_funnel/coeff_rate.rs:60). */
typedef struct {
  uint32_t coeff;       /* element offset of the block's packed levels */
  int32_t txt_rate;     /* the tx type's rate (MdRates::txt_rate), added for luma */
  uint8_t tx_class;     /* 0 2D, 1 horizontal, 2 vertical */
  uint8_t plane_type;   /* 0 luma, 1 chroma */
  uint8_t txb_skip_ctx; /* < 13 */
  uint8_t dc_sign_ctx;  /* < 3 */
} zav1_rate_job;

/* Coefficient rate (1/512 bit) of n_jobs quantized blocks of tx_size
 * (packed, coded size, raster; d_eob[j] in scan order): d_rate[j] =
 * cost_coeffs_txb (coeff_rate.rs:60, the full loop) priced with the frame's
 * zav1_coeff_costs and tx_size's three scans (as for zav1_rdoq_batch).
 * Blocks with eob 0, or a job field out of range, get 0. */
int zav1_coeff_rate_batch(int tx_size, const int32_t *d_qcoeff, const uint16_t *d_eob,
                          const zav1_rate_job *d_jobs, int n_jobs, const zav1_coeff_costs *d_costs,
                          const uint16_t *d_scans, int32_t *d_rate, zav1_stream_t s);

/* AV1 coefficient scan of tx_size's packed (<= 32x32) block for a scan
 * class (0 default diagonal, 1 row, 2 column; the reference's
 * TX_TYPE_TO_SCAN_INDEX classes).  Replaces scan()
 * (rust/crates/svtav1-entropy/src/scan_tables.rs:1640).  Host only; returns
 * the scan length or a negative status. */
int zav1_scan_order(int tx_size, int scan_class, uint16_t *scan);

/* The frame's coefficient rate tables from the AV1 default CDFs of
 * base_qindex's q context.  Replaces build_coeff_cost_tables
 * (rust/crates/svtav1-encoder/src/quant.rs:514, over
 * rust/crates/svtav1-entropy/src/coeff_c.rs:651 default_for_qindex).  Host
 * only. */
int zav1_build_coeff_costs(int base_qindex, zav1_coeff_costs *out);

/* Rate (1/512 bit) of signalling tx_type for an intra luma transform of
 * tx_size whose (filter-intra mapped) prediction direction is intra_dir
 * (0..12), from the default intra ext-tx CDFs; 0 where the size's intra set
 * holds one type or tx_type is not in it.  Replaces MdRates::txt_rate
 * (rust/crates/svtav1-encoder/src/leaf_funnel/rate_tables.rs:291, intra).
 * Host only; negative on a bad argument. */
int zav1_intra_txt_rate(int tx_size, int intra_dir, int tx_type);

/* Luma rate (1/512 bit) of an intra candidate (mode 0..12, angle delta
 * -3..3, filter-intra mode 0..4 or 5 = none) of a w x h key-frame block under
 * the above / left mode contexts (0..4, intra_mode_context of the neighbour
 * modes, 0 where unavailable), from the default CDFs: the kf y-mode symbol,
 * the angle delta of a directional mode on blocks of at least 8x8 and, for
 * DC_PRED up to 32x32, the filter-intra flag and mode.  Replaces the luma
 * terms of build_md_rates
 * (rust/crates/svtav1-encoder/src/leaf_funnel/rate_tables.rs:146) as the
 * fast loop sums them.  Host only; negative on a bad argument. */
int zav1_intra_mode_rate(int w, int h, int above_ctx, int left_ctx, int mode, int delta, int fi_mode);

/* The key-frame full mode-decision lambda of an allintra still at qindex
 * (CLI qp cli_qp, bd 8 or 10).  Replaces kf_full_lambda_8bit / _bd10
 * (rust/crates/svtav1-encoder/src/pd0.rs:223 / :294).  Host only; 0 on a bad
 * argument. */
uint64_t zav1_kf_full_lambda(int qindex, int cli_qp, int bd);

/* Fused transform-unit evaluation of the tx-type search (the per-candidate
 * body of tx_unit_inner, rust/crates/svtav1-encoder/src/leaf_funnel/
 * tx_pipeline.rs:268, without RDOQ / rate): for every block (d_xy[2j],
 * d_xy[2j+1]) of tx_size on the 16-bit plane (8-bit samples widened), DC
 * prediction from the source neighbours, residual, forward transform,
 * 64-point fold, fp quantization at qindex, frequency distortion.  For
 * each block j and host-side candidate type types[k] (n_types <= 16) writes
 * d_dist[j*n_types+k] (u64), d_eob[...] (scan order), d_lsum[...] (sum of
 * |level|). Samples must lie in [0, 2^bd - 1] (the fused transforms size their
 * arithmetic for residuals of at most 10-bit samples). */
int zav1_txsearch_batch(const uint16_t *d_plane, int stride, int bd, int tx_size, const int32_t *d_xy,
                        int n_blocks, const uint8_t *types, int n_types, int qindex, uint64_t *d_dist,
                        int32_t *d_eob, uint32_t *d_lsum, zav1_stream_t s);
/* Origins of the full grid of bw x bh blocks covering a plane_w x plane_h
 * plane in raster order (x, y pairs, the layout d_xy of the search takes):
 * writes (plane_w / bw) * (plane_h / bh) pairs and returns that count. */
int zav1_grid_xy(int32_t *d_xy, int plane_w, int plane_h, int bw, int bh, zav1_stream_t s);
/* The same search at n_q quantizers (qindices[q], 1 <= n_q <= 8): each
 * block's transforms are computed once and quantized at every qindex.
 * Outputs per (q, block j, type k) at q*n_blocks*n_types + j*n_types + k;
 * identical to n_q calls of zav1_txsearch_batch (the evaluation of one
 * candidate set at several quantizers, as a quantizer or rate-control
 * search asks for). */
int zav1_txsearch_batch_q(const uint16_t *d_plane, int stride, int bd, int tx_size, const int32_t *d_xy,
                          int n_blocks, const uint8_t *types, int n_types, const int *qindices, int n_q,
                          uint64_t *d_dist, int32_t *d_eob, uint32_t *d_lsum, zav1_stream_t s);

/* Fused transform units with a given prediction: for every job (one unit of
 * tx_size at (x, y) of d_src, its w*h u16 prediction at d_pred + pred_off,
 * its own tx type) residual -> forward 2D -> 64-point fold -> quantize
 * (quantize_fp, or quantize_b with ZAV1_TXU_QUANT_B) -> eob (scan order),
 * cul level byte, and distortion: frequency-domain (sse_i32 / sq_sum_i32 +
 * folded energy, >> (1 - log_scale) * 2) or, with ZAV1_TXU_SPATIAL, the SSE
 * of the reconstruction against the source over crop_w x crop_h, << 4.
 * With ZAV1_TXU_RECON (implied by SPATIAL) the unit is reconstructed:
 * inverse 2D of the dequantized levels + prediction, clipped to bd.
 * Replaces tx_unit_inner / tx_unit_hbd
 * (rust/crates/svtav1-encoder/src/leaf_funnel/tx_pipeline.rs:268 / :854)
 * without the RDOQ trellis, quantization matrices and coefficient rate
 * (the caller prices d_qcoeff), with compute_cul_level (:130).
 * Per job: d_dist (u64), d_eob (i32; -1 = no kernel for the type at this
 * size, the reference's `false`; -22 = unit or prediction out of bounds),
 * d_cul (u8); optional d_qcoeff (packed <= 32x32 levels, job-major) and
 * d_recon (w*h u16, job-major).  d_src is u8 at bd 8, u16 at bd 10. Source and prediction samples must lie in [0, 2^bd - 1]. */
typedef struct zav1_txunit_job {
  int32_t x, y;            /* unit origin in d_src (samples) */
  uint32_t pred_off;       /* d_pred offset (samples) of the w*h prediction */
  uint8_t tx_type;         /* 0..15 */
  uint8_t crop_w, crop_h;  /* spatial-distortion extent, <= w, h */
  uint8_t pad;
} zav1_txunit_job;
#define ZAV1_TXU_QUANT_B 1u
#define ZAV1_TXU_SPATIAL 2u
#define ZAV1_TXU_RECON 4u
int zav1_txunit_batch(const void *d_src, int stride, int plane_w, int plane_h, int bd, int tx_size,
                      const uint16_t *d_pred, size_t pred_len, const zav1_txunit_job *d_jobs, int n_jobs, int qindex,
                      unsigned flags, uint64_t *d_dist, int32_t *d_eob, uint8_t *d_cul, int32_t *d_qcoeff,
                      uint16_t *d_recon, zav1_stream_t s);

/* The full-RD transform unit of the mode decision's full loop: the body of
 * tx_unit_inner with RDOQ and the exact coefficient rate
 * (rust/crates/svtav1-encoder/src/leaf_funnel/tx_pipeline.rs:268-640,
 * RateMode::Exact; tx_unit_hbd :854 at 10 bits), for n_jobs units of one
 * tx_size.  Per job: residual -> forward 2D -> 64-point fold -> quantize_b
 * (ZAV1_TXU_QUANT_B) or quantize_fp, then with ZAV1_TXU_RDOQ the trellis
 * (optimize_b, rust/crates/svtav1-encoder/src/quant.rs:1312, cut-offs
 * rdoq_cutoffs :1551) -> eob -> coefficient rate cost_coeffs_txb
 * (leaf_funnel/coeff_rate.rs:60) including the job's tx-type rate, or the
 * skip cost txb_skip_cost[ctx][1] when the unit ends with eob 0 -> cul level
 * -> distortion: frequency-domain (sse of the final levels + folded energy,
 * >> (1 - log_scale) * 2) or, with ZAV1_TXU_SPATIAL, SSE << 4 of the
 * reconstruction (inverse of the final levels + prediction) over the crop.
 * d_costs: the frame's tables (zav1_build_coeff_costs).  d_work: a device
 * workspace of at least zav1_txunit_rd_workspace(tx_size, n_jobs) bytes,
 * the caller's (the call allocates nothing); calls sharing a workspace must
 * be ordered on one stream.  Outputs per job: d_dist, d_eob (-1 no kernel
 * for the type, -22 unit out of bounds: then rate 0), d_rate (1/512 bit),
 * d_cul; optional d_qcoeff (final packed levels) and d_recon (w*h u16).
 * ZAV1_TXU_RATE_EST2: the luma rate of a unit whose eob is below
 * (w*h) >> 6 is the fast approximation 6000 + 1000 * eob instead of the
 * exact one (coeff_rate_est_lvl 2 of presets 7 and 8, tx_pipeline.rs:605). */
#define ZAV1_TXU_RDOQ 8u
#define ZAV1_TXU_RATE_EST2 16u
typedef struct zav1_txunit_rd_job {
  int32_t x, y;           /* unit origin in d_src (samples) */
  uint32_t pred_off;      /* d_pred offset (samples) of the w*h prediction */
  uint8_t tx_type;        /* 0..15 */
  uint8_t crop_w, crop_h; /* spatial-distortion extent, <= w, h */
  uint8_t plane_type;     /* 0 luma, 1 chroma: selects the cost tables */
  uint8_t txb_skip_ctx;   /* < 13 */
  uint8_t dc_sign_ctx;    /* < 3 */
  uint8_t sharpness;      /* sharpness_flag of the trellis */
  uint8_t pad;
  int32_t txt_rate;       /* tx-type rate added to the coefficient rate (luma) */
  int64_t rdmult;         /* trellis RD multiplier (rdoq_rdmult_full, quant.rs:923) */
} zav1_txunit_rd_job;
int zav1_txunit_rd_workspace(int tx_size, int n_jobs, size_t *bytes);
int zav1_txunit_rd_batch(const void *d_src, int stride, int plane_w, int plane_h, int bd, int tx_size,
                         const uint16_t *d_pred, size_t pred_len, const zav1_txunit_rd_job *d_jobs, int n_jobs,
                         int qindex, unsigned flags, uint32_t cut_off_num, uint32_t cut_off_denum,
                         const zav1_coeff_costs *d_costs, void *d_work, size_t work_bytes, uint64_t *d_dist,
                         int32_t *d_eob, int32_t *d_rate, uint8_t *d_cul, int32_t *d_qcoeff, uint16_t *d_recon,
                         zav1_stream_t s);

/* zav1_txunit_rd_batch without a workspace argument and without any device
 * allocation: no counterpart in the reference.  For callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The temporaries live in the library's fixed
 * per-GPU scratch of zav1_noalloc_scratch_bytes() (32 MiB, held by each GPU
 * the library runs on until the process ends); zav1_init(device) must have
 * run on the device first.  A call whose n_jobs need more than the scratch
 * (zav1_txunit_rd_workspace) returns ZAV1_EINVAL with the sizes in
 * zav1_last_error().  Every call on one GPU uses the same scratch: calls
 * that overlap on one GPU (different streams or host threads without
 * ordering) race.  Nothing in the scratch needs to start at zero. */
int zav1_txunit_rd_batch_noalloc(const void *d_src, int stride, int plane_w, int plane_h, int bd, int tx_size,
                                 const uint16_t *d_pred, size_t pred_len, const zav1_txunit_rd_job *d_jobs,
                                 int n_jobs, int qindex, unsigned flags, uint32_t cut_off_num,
                                 uint32_t cut_off_denum, const zav1_coeff_costs *d_costs, uint64_t *d_dist,
                                 int32_t *d_eob, int32_t *d_rate, uint8_t *d_cul, int32_t *d_qcoeff,
                                 uint16_t *d_recon, zav1_stream_t s);
size_t zav1_noalloc_scratch_bytes(void);

/* ------------------------------------------------------------------------ */
/* Mode-decision funnel, MDS0 (batched).                                     */
/* ------------------------------------------------------------------------ */
/* MDS0 SATD of (unit, prediction) jobs of one w x h unit size (w, h powers
 * of two in 4..64): residual d_src - prediction over square tiles of
 * min(w, h, 32), raster order, aom Hadamard + SATD per tile, summed.
 * Replaces hadamard_satd / hadamard_satd_hbd
 * (rust/crates/svtav1-encoder/src/leaf_funnel/predict.rs:143 / :188) over
 * every candidate of every unit of the fast loop (mod.rs:2084).  Jobs use
 * zav1_txunit_job's x, y, pred_off (the other fields are ignored); d_satd
 * (u32 per job) is 0xffffffff for a job outside the plane or d_pred.
 * d_src is u8 at bd 8, u16 at bd 10; d_pred u16 (pred_len samples). */
int zav1_pred_satd_batch(const void *d_src, int stride, int plane_w, int plane_h, int bd, int w, int h,
                         const uint16_t *d_pred, size_t pred_len, const zav1_txunit_job *d_jobs, int n_jobs,
                         uint32_t *d_satd, zav1_stream_t s);

/* MDS0 survivor selection on the device.  d_satd / d_rate (nullable: rate
 * 0) hold n_cands entries per unit, unit-major, in candidate generation
 * order; the fast cost is rdcost(lambda, rate, satd << 4) =
 * ((rate * lambda + 256) >> 9) + (satd << 11) (mod.rs:2181,
 * rate_tables.rs:1033), an invalid SATD costing UINT64_MAX.  The first k
 * positions of the reference's exchange sort over those costs
 * (c_exchange_sort_by, rate_tables.rs:108, as sort_lane uses it at
 * mod.rs:2994: not stable on ties) go to d_sel[u*k + r] (candidate index)
 * and d_sel_cost[u*k + r] (nullable), and the transform-unit jobs of the
 * survivors, one per host-side type types[t] (n_types <= 16), to
 * d_jobs[(u*k + r)*n_types + t]: x, y = d_unit_xy[u*xy_stride + 0 / 1]
 * (xy_stride 2 for (x, y) pairs, 5 for zav1_md_unit records), crop w x h,
 * pred_off = (u*n_cands + cand)*w*h, or (u*k + r)*w*h when survivors_only
 * (the layout of zav1_md_predict_survivors), ready for zav1_txunit_batch.
 * 1 <= k <= n_cands <= 64. */
int zav1_md_fast_select(const uint32_t *d_satd, const uint32_t *d_rate, uint64_t lambda, int n_units, int n_cands,
                        int k, const int32_t *d_unit_xy, int xy_stride, int w, int h, const uint8_t *types,
                        int n_types, int survivors_only, uint8_t *d_sel, uint64_t *d_sel_cost,
                        zav1_txunit_job *d_jobs, zav1_stream_t s);

/* Fused MDS0 fast loop.  A unit (x, y, availability counts as for
 * zav1_intra_job, edge-filter flags) and a candidate list shared by every
 * unit of the call (host array, n_cands <= 64, generation order):
 * mode 0..12, angle delta -3..3 (directional modes), filter-intra mode
 * 0..4 with mode 0 and w, h <= 32, or 5 = none. */
typedef struct zav1_md_unit {
  int32_t x, y;
  int16_t n_top_px, n_topright_px, n_left_px, n_bottomleft_px;
  uint8_t edge_filter, filt_type, pad0, pad1;
} zav1_md_unit;
typedef struct zav1_md_cand {
  uint8_t mode;
  int8_t delta;
  uint8_t fi_mode, pad;
} zav1_md_cand;
/* For every unit of one w x h size (an AV1 block size: powers of two 4..64,
 * aspect at most 4:1) and every
 * candidate: the prediction from d_recon (predict_unit, lf/predict.rs:49)
 * and its Hadamard SATD against d_src (hadamard_satd, lf/predict.rs:143),
 * the fast loop of mod.rs:2060-2185, into d_satd[u*n_cands + m].  The
 * predictions stay on chip; the result equals zav1_intra_pred_batch +
 * zav1_pred_satd_batch.  A unit whose reference reads (availability
 * counts) would leave the plane gets 0xffffffff for every candidate.  d_src
 * and d_recon share stride and sample type (u8 at bd 8, u16 at bd 10). */
int zav1_md_fast_satd(const void *d_src, const void *d_recon, int stride, int plane_w, int plane_h, int bd, int w,
                      int h, const zav1_md_unit *d_units, int n_units, const zav1_md_cand *cands, int n_cands,
                      uint32_t *d_satd, zav1_stream_t s);
/* Predictions of the survivors for the next stage: candidate d_sel[u*k + r]
 * of unit u into d_pred + (u*k + r)*w*h (u16; zeros for a unit that
 * zav1_md_fast_satd reports invalid). */
int zav1_md_predict_survivors(const void *d_recon, int stride, int plane_w, int plane_h, int bd, int w, int h,
                              const zav1_md_unit *d_units, int n_units, const zav1_md_cand *cands, int n_cands,
                              const uint8_t *d_sel, int k, uint16_t *d_pred, zav1_stream_t s);

/* ------------------------------------------------------------------------ */
/* Superblock-wavefront mode decision (batched form of the leaf funnel).     */
/* ------------------------------------------------------------------------ */
/* Every leaf of a partition (transform = block) -- uniform leaf x leaf
 * blocks, or the leaves of a partition map such as the partition pre-pass
 * decides (the fixed tree the reference codes at presets 6..8,
 * pipeline.rs:8000-8060 encode_fixed_tree) -- of n_images equal-size luma
 * planes, stacked vertically in one device plane, decided with the
 * reference's funnel (evaluate_leaf,
 * rust/crates/svtav1-encoder/src/leaf_funnel/mod.rs:1007): MDS0 over the
 * candidate list predicted from the reconstruction with AV1 neighbour
 * availability (enc/intra_edge.rs:355 / :436 within the tile), fast cost
 * with the default-CDF mode rates under the neighbours' mode contexts, the
 * exchange sort's first k1 into MDS1 (DCT_DCT, quantize_b, frequency
 * distortion, exact rate), the post-MDS1 exchange sort's first k3 into MDS3
 * (the intra set's tx types in TX_TYPE_GROUPS order, quantize_fp + RDOQ at
 * rdoq_level or quantize_b at level 0, spatial distortion; the tx-type rate
 * gate txt_rate_th, the coefficient-SATD early exit txt_satd_th
 * (qp-scaled), no signalable eob-0 non-DCT), in MDS1 and MDS3 the luma rate
 * at coeff_rate_est_lvl 2 approximated for small eobs, the least full cost (mode rate
 * + coefficient rate, distortion) first in order; its reconstruction is
 * committed before the leaves that read it.  A leaf runs one step after the
 * latest leaf whose reconstruction or mode it reads (the coding order's
 * dependencies within its tile), every tile and image of the batch together,
 * one chain of launches per (step, leaf size); 64x64 leaves take no filter
 * intra (the configuration's filter-intra candidates, which must come last,
 * are dropped for them).  Rate tables are the
 * frame's default-CDF tables (no per-superblock CDF adaptation); luma only. */
typedef struct zav1_wf_cfg {
  int32_t width, height;          /* per image, multiples of leaf */
  int32_t n_images;               /* images of the batch, stacked in one plane */
  int32_t bd;                     /* 8 or 10 */
  int32_t leaf;                   /* 8, 16, 32 or 64 (uniform leaves) */
  int32_t sb;                     /* superblock 64 or 128 */
  int32_t tile_cols, tile_rows;   /* uniform tiles (in superblocks) */
  int32_t qindex;
  int32_t k1, k3;                 /* candidates into MDS1 / MDS3 (k3 <= k1 <= n_cands) */
  int32_t rdoq_level;             /* 0 quantize_b; 1, 2, 3 quantize_fp + trellis (rdoq_cutoffs) */
  int32_t txt_groups;             /* TX_TYPE_GROUPS searched (1..6) */
  int32_t txt_rate_th;            /* txt_rate_cost_th */
  int32_t txt_satd_th;            /* qp-scaled SATD early-exit threshold (0 = off) */
  int32_t coeff_rate_est_lvl;     /* 0, 1 exact luma rate; 2 the fast approximation (ZAV1_TXU_RATE_EST2) */
  uint64_t lambda;                /* full lambda */
  int32_t n_cands;                /* <= 64 */
  int32_t pad;
  zav1_md_cand cands[64];
  int32_t txt_groups_lt16;        /* with a partition map: tx-type groups of leaves below 16x16 */
  int32_t keep_levels;            /* 1: keep every winner's quantized levels (for the entropy coder) */
  const uint8_t *leaf_map;        /* NULL: uniform leaves of `leaf`; else a partition map (host memory, read by
                                     zav1_wf_create / zav1_wf_footprint only): per image, per 64x64 unit in
                                     raster order, 256 log2 leaf sizes of its 4x4 cells (zav1_pd0_run's
                                     d_leaf_map layout); leaves 8..64, sides multiples of 64 */
} zav1_wf_cfg;
typedef struct zav1_wf_decision {
  uint8_t cand, tx_type, mode, fi_mode; /* candidate index, its tx type, mode and filter-intra mode */
  int32_t eob, rate, mode_rate;         /* coefficient rate incl. tx-type rate; the mode rate */
  uint64_t dist, cost;                  /* spatial distortion (SSE << 4), full cost */
} zav1_wf_decision;
typedef struct zav1_wf_info_t {
  void *d_src;                    /* stacked source plane (u8 / u16, stride width): the caller fills it */
  void *d_recon;                  /* stacked reconstruction plane, written by a run */
  zav1_wf_decision *d_decisions;  /* per leaf, in leaf order */
  const zav1_md_unit *d_units;    /* per leaf: origin in the stacked plane, availability */
  int32_t n_leaves, n_steps, max_leaves_per_step, n_types, stacked_height;
  uint8_t types[16];              /* MDS3 tx types in search order */
  size_t bytes;                   /* device memory of the object */
  int32_t *d_levels;              /* keep_levels: the winners' levels, packed min(leaf, 32)^2 raster per leaf */
  const uint64_t *d_level_off;    /* keep_levels: per leaf (storage order) its offset in d_levels */
  size_t n_levels;
  const uint64_t *d_trellis_runs; /* MDS3 jobs handed to the RDOQ trellis with a nonzero eob, summed over every
                                   * run of the object (measurement: read it before and after a run) */
} zav1_wf_info_t;
typedef struct zav1_wf zav1_wf;
/* Device bytes zav1_wf_create would take for cfg (admission by footprint). */
int zav1_wf_footprint(const zav1_wf_cfg *cfg, size_t *bytes);
/* Plans the steps and takes the device memory (one allocation); zav1_init
 * must have run on the current device. */
int zav1_wf_create(const zav1_wf_cfg *cfg, zav1_wf **out);
int zav1_wf_info(const zav1_wf *wf, zav1_wf_info_t *info);
/* Decides every leaf: use_graph 0 enqueues the steps' launches on s; 1
 * captures them once into a HIP graph (s must be a created stream) and
 * replays it. */
int zav1_wf_run(zav1_wf *wf, int use_graph, zav1_stream_t s);
/* Steps [first_step, first_step + n_steps) of the eager chain (earlier steps
 * must have been enqueued before): the unit a serving loop advances each
 * request by, so that requests on shared hardware queues interleave. */
int zav1_wf_run_steps(zav1_wf *wf, int first_step, int n_steps, zav1_stream_t s);
/* The eager run with HIP events at the stage boundaries of every step, the
 * stream synchronized after each step; stage_ms[k] (n_stages >=
 * ZAV1_WF_STAGES) receives stage k's total device time over the run:
 *   0 candidate rates and contexts    1 MDS0 prediction + SATD
 *   2 MDS0 selection (fast cost, exchange sort)
 *   3 survivor prediction             4 MDS1 job build
 *   5 MDS1 transform units (quantize_b, rate, distortion)
 *   6 MDS3 job build (tx-type gates)  7 MDS3 forward transform + quantize
 *   8 MDS3 RDOQ trellis               9 MDS3 rate, distortion, reconstruction
 *  10 winner and reconstruction commit
 * The decisions equal zav1_wf_run's.  No counterpart in the reference
 * (measurement). */
#define ZAV1_WF_STAGES 11
/* The plan's autotuner: times the RDOQ trellis of steps [first_step,
 * first_step + n_steps) under each of its forms -- 0 the persistent form fed
 * by a device work queue, 1 the direct group / wave form, 2 the lane form --
 * and keeps the fastest per leaf size (8, 16, 32, 64 -> forms[0..3]; -1
 * where the size has no step in the window or no trellis runs; ms[k * 3 +
 * f], when not NULL, the measured milliseconds).  Every form gives the same
 * levels, so the decisions do not change; the window's leaves are decided
 * again (run the plan once first so their neighbours are decided).  A graph
 * captured before is dropped and captured again by the next graph run.  No
 * counterpart in the reference (tuning). */
int zav1_wf_autotune(zav1_wf *wf, int first_step, int n_steps, zav1_stream_t s, int32_t *forms, double *ms);
int zav1_wf_run_timed(zav1_wf *wf, zav1_stream_t s, double *stage_ms, int n_stages);
/* The plain per-call form of the same decisions for leaves
 * [first_leaf, first_leaf + n_leaves) in leaf order (earlier leaves must be
 * decided): one launch per candidate prediction + SATD, per survivor
 * prediction and transform unit, per MDS3 tx type, with the selections on
 * the host from the results read back after each stage; synchronous. */
int zav1_wf_run_plain(zav1_wf *wf, int first_leaf, int n_leaves, zav1_stream_t s);
int zav1_wf_destroy(zav1_wf *wf);
const char *zav1_wf_last_error(void);

/* ------------------------------------------------------------------------ */
/* Contexts and MD waves: an HBM budget, image upload and a queue of        */
/* mode-decision waves, collected when the host needs the results.           */
/* ------------------------------------------------------------------------ */
/* A context owns one stream and one device arena of hbm_budget bytes
 * (allocated once: ZAV1_ENOMEM maps to EncodeError::AllocFailed,
 * rust/crates/svtav1-types/src/error.rs:22-40).  Allocations from the arena
 * are 256-byte aligned and live until zav1_ctx_reset or zav1_ctx_destroy. */
typedef struct zav1_ctx zav1_ctx;
int zav1_ctx_create(int device, size_t hbm_budget, zav1_ctx **ctx);
int zav1_ctx_destroy(zav1_ctx *ctx);
int zav1_ctx_alloc(zav1_ctx *ctx, size_t bytes, void **d_ptr);
/* Bytes in use / free in the arena. */
int zav1_ctx_usage(const zav1_ctx *ctx, size_t *used, size_t *capacity);
/* Drops every arena allocation (call after zav1_wave_collect). */
int zav1_ctx_reset(zav1_ctx *ctx);
zav1_stream_t zav1_ctx_stream(zav1_ctx *ctx);
/* Copies a w x h plane (host rows `host_stride` samples apart; u8 at bd 8,
 * u16 at bd 10) into a tight plane allocated from the arena. */
int zav1_image_upload(zav1_ctx *ctx, const void *host, size_t host_stride, int w, int h, int bd, void **d_plane);

/* One MD wave: the units of one size whose neighbours are reconstructed
 * (an SB-diagonal wavefront's leaves), through MDS0 and MDS1 — the fused fast
 * loop (zav1_md_fast_satd), survivor selection (zav1_md_fast_select),
 * survivor prediction (zav1_md_predict_survivors) and the survivors'
 * transform units of each type in `types` (zav1_txunit_batch) — with every
 * intermediate allocated from the context arena at submit time.  Outputs
 * (device, caller-owned) per (unit, rank[, type]): sel, sel_cost
 * (nullable), dist / eob / cul (unit-major, then rank, then type) and the
 * packed levels (nullable, <= 32x32 per job). */
typedef struct zav1_md_wave {
  const void *d_src, *d_recon;
  int stride, plane_w, plane_h, bd, w, h;
  const zav1_md_unit *d_units;
  int n_units;
  const zav1_md_cand *cands; /* host, n_cands <= 64 */
  int n_cands;
  const uint32_t *d_rate;    /* nullable */
  uint64_t lambda;
  int k;
  const uint8_t *types;      /* host, n_types <= 16 */
  int n_types;
  int qindex;
  unsigned txu_flags;        /* ZAV1_TXU_* (RECON / SPATIAL need d_recon_out) */
  uint8_t *d_sel;
  uint64_t *d_sel_cost;
  uint64_t *d_dist;
  int32_t *d_eob;
  uint8_t *d_cul;
  int32_t *d_qcoeff;
  uint16_t *d_recon_out;
} zav1_md_wave;
/* The tx-type search over a whole plane: every tx size's full block grid
 * (zav1_grid_xy), each with the legal types of its size (all 16 up to 16
 * points per side; DCT_DCT, IDTX, V_DCT, H_DCT at 32; DCT_DCT at 64), at
 * n_q quantizers (zav1_txsearch_batch_q), on the context stream; grids and
 * outputs are allocated from the context arena (ZAV1_ENOMEM, nothing
 * enqueued, when they do not fit).  d_plane: u16 plane of w x h (tight
 * rows).  out receives per size the block and type counts, the types and
 * the device outputs ([q][block][type]; NULL for sizes larger than the
 * plane).  The search body of a frame's MD in one call. */
typedef struct zav1_frame_search {
  int32_t n_blocks[19], n_types[19];
  uint8_t types[19][16];
  uint64_t *dist[19];
  int32_t *eob[19];
  uint32_t *lsum[19];
} zav1_frame_search;
int zav1_txsearch_frame(zav1_ctx *ctx, const uint16_t *d_plane, int w, int h, int bd, const int *qindices, int n_q,
                        zav1_frame_search *out);

/* Enqueues one wave on the context stream; fails with ZAV1_ENOMEM (nothing
 * enqueued) when the arena cannot hold its intermediates. */
int zav1_md_wave_submit(zav1_ctx *ctx, const zav1_md_wave *wave);
/* Waits for every submitted wave; returns the first error any of them met. */
int zav1_wave_collect(zav1_ctx *ctx);

/* ------------------------------------------------------------------------ */
/* Intra prediction (batched).                                               */
/* ------------------------------------------------------------------------ */
/* One transform unit.  Replaces the per-unit predictor routing of
 * predict_unit (rust/crates/svtav1-encoder/src/leaf_funnel/predict.rs:49) and
 * its u16 twin predict_unit_hbd (tx_pipeline.rs:667): the neighbour fill of
 * extract_neighbors_tiled (partition.rs:250), the directional edge assembly,
 * corner / edge filters and upsampling of dr_predict (intra_edge.rs:910) and
 * the predictors of rust/crates/svtav1-dsp/src/intra_pred.rs:18-1060 /
 * hbd.rs:169-850.  The caller supplies the availability counts (C
 * n_top_px / n_topright_px / n_left_px / n_bottomleft_px, i.e. real reference
 * samples after the frame-extent and coding-order rules of
 * intra_edge.rs:355/436); the device reads every sample from the plane. */
typedef struct zav1_intra_job {
  int32_t x, y;        /* unit origin in the plane (samples) */
  uint8_t w, h;        /* 4..64 */
  uint8_t mode;        /* 0 DC, 1 V, 2 H, 3 D45, 4 D135, 5 D113, 6 D157, 7 D203,
                          8 D67, 9 SMOOTH, 10 SMOOTH_V, 11 SMOOTH_H, 12 PAETH */
  int8_t delta;        /* angle delta -3..3 (3-degree steps) */
  uint8_t fi_mode;     /* filter intra 0..4 (w, h <= 32), 5 = none */
  uint8_t edge_filter; /* sequence enable_intra_edge_filter */
  uint8_t filt_type;   /* 1 if the above or left block uses a smooth mode */
  uint8_t pad0;
  int16_t n_top_px, n_topright_px, n_left_px, n_bottomleft_px;
  uint32_t out_off;    /* d_out offset (samples) of this unit's w*h raster */
  uint32_t pad1;
} zav1_intra_job;

/* d_plane: 8-bit (bd 8) or 16-bit (bd 10) reconstruction, `stride` samples;
 * d_out: u16 predictions. */
int zav1_intra_pred_batch(const void *d_plane, int stride, int bd, const zav1_intra_job *d_jobs, int n_jobs,
                          uint16_t *d_out, zav1_stream_t s);

/* Chroma-from-luma: cfl_luma_subsampling_420 + cfl_subtract_average +
 * cfl_predict (rust/crates/svtav1-dsp/src/intra_pred.rs:1088,1110,1137;
 * hbd.rs:855,889).  Per job: the 2cw x 2ch co-located luma at (lx, ly) of
 * d_luma, the cw x ch chroma DC prediction at d_dc + dc_off, alpha in Q3. */
typedef struct zav1_cfl_job {
  int32_t lx, ly, cw, ch, alpha_q3;
  uint32_t dc_off, out_off, pad;
} zav1_cfl_job;
int zav1_cfl_batch(const void *d_luma, int stride, int bd, const zav1_cfl_job *d_jobs, int n_jobs,
                   const uint16_t *d_dc, uint16_t *d_out, zav1_stream_t s);

/* ------------------------------------------------------------------------ */
/* Loop-filter stage.  Planes are tight (stride = plane width) u8 at bd 8,   */
/* u16 at bd 10; chroma is 4:2:0.                                            */
/* ------------------------------------------------------------------------ */
/* Per-4x4 (mi) luma-grid geometry, all pointers device memory, row-major
 * mi_rows x mi_cols (the reference's DeblockGeom,
 * rust/crates/svtav1-encoder/src/deblock.rs:532).  Transforms and blocks
 * must be aligned to their own size (as AV1 partitions are). */
typedef struct zav1_deblock_geom {
  int32_t mi_cols, mi_rows;   /* luma mi grid */
  int32_t true_w, true_h;     /* coded (unpadded) luma frame size */
  const uint32_t *block_id;   /* covering block id per mi */
  const uint8_t *bw, *bh;     /* covering block dims (px) */
  const uint8_t *tw, *th;     /* luma transform dims (px) */
  const uint8_t *skip;        /* inter skip per mi; NULL = all-intra */
} zav1_deblock_geom;

/* In-place deblocking of one plane (plane 0 luma, 1/2 chroma with ss = 1):
 * every vertical edge, then every horizontal edge, at the given levels.
 * Replaces filter_plane / filter_plane_hbd
 * (rust/crates/svtav1-encoder/src/deblock.rs:811 / :875) with lpf_params
 * (:698) and the lpf_{vertical,horizontal}_{4,6,8,14} filters
 * (rust/crates/svtav1-dsp/src/loop_filter.rs:50-405, hbd.rs:1040-1433).
 * geom is a host struct holding device pointers.  Luma plane dimensions must
 * be multiples of 8 and chroma multiples of 4 (the padded frame), so every
 * filter tap stays inside the plane; else -EINVAL. */
int zav1_deblock_plane(void *d_plane, int stride, int plane_w, int plane_h, int bd, int plane, int ss, int level_v,
                       int level_h, int sharpness, const zav1_deblock_geom *geom, zav1_stream_t s);

/* Deblocking level trials: for each of n_levels (<= 64) candidate levels
 * (host array, 0..63; the level applies to both edge directions, as on the
 * still-picture path) a tight copy of d_recon in d_work (n_levels planes;
 * work_bytes checked) is filtered and its SSE against d_src written to
 * d_sse[i].  Replaces try_filter_plane + plane_sse
 * (rust/crates/svtav1-encoder/src/deblock.rs:228 / :208).  Planes tight. */
int zav1_deblock_level_sse(const void *d_recon, const void *d_src, int plane_w, int plane_h, int bd, int plane, int ss,
                           int sharpness, const zav1_deblock_geom *geom, const uint8_t *levels, int n_levels,
                           void *d_work, size_t work_bytes, uint64_t *d_sse, zav1_stream_t s);

/* Deblocking level search of one plane: search_filter_level
 * (rust/crates/svtav1-encoder/src/deblock.rs:337) on the still path (start
 * level 0; luma both directions at one level, chroma per plane), with the
 * trials of each round run on the GPU (at most two levels per round; d_work
 * >= 2 planes, d_sse >= 2 entries).  conv_th is the early-exit convergence
 * count (0 = full search).  Synchronizes `s` once per round.  Outputs the
 * picked level and the SSEs at level 0 and at the pick. */
int zav1_deblock_search_level(const void *d_recon, const void *d_src, int plane_w, int plane_h, int bd, int plane,
                              int ss, int sharpness, const zav1_deblock_geom *geom, int conv_th, void *d_work,
                              size_t work_bytes, uint64_t *d_sse, int *best_level, int64_t *zero_sse,
                              int64_t *best_sse, zav1_stream_t s);

/* CDEF strength search of a whole frame: for every 64x64 filter block
 * (raster, ceil(w/64) x ceil(h/64)) and every strength index g = 4 * pri +
 * sec (pri 0..15, sec 0..3 -> 0,1,2,4), the SSE of the filtered luma
 * (d_mse_y[fb*64 + g]) and of U + V (d_mse_uv, may be NULL) against the
 * source, with the reference's per-plane shift and subsampling scale.
 * d_dirvar[(fb*64 + by*8 + bx)*2 + {0,1}] receives each 8x8's direction and
 * variance (required when d_mse_uv is given, else may be NULL).  d_skip
 * (may be NULL) is the per-mi signaled skip flag, mi_rows x mi_cols (the
 * reference's DeblockGeom::skip): an 8x8 whose mi are all skip is left out
 * of the search (is_8x8_all_skip, rust/crates/svtav1-encoder/src/
 * deblock.rs:652), and a block with none left has zero rows.  Damping is
 * 3 + (qindex >> 6); rows are sampled every `subsampling` (luma, at most 4)
 * / min(subsampling, 1) (chroma).  Replaces cdef_search_still /
 * cdef_search_still_hbd (rust/crates/svtav1-encoder/src/cdef.rs:1166 /
 * :1374) over every candidate, with cdef_find_dir and cdef_filter_block
 * (rust/crates/svtav1-dsp/src/cdef.rs:143 / :931). */
int zav1_cdef_search(const void *d_rec_y, const void *d_rec_u, const void *d_rec_v, const void *d_src_y,
                     const void *d_src_u, const void *d_src_v, int width, int height, int bd, int qindex,
                     int subsampling, const uint8_t *d_skip, int mi_cols, int mi_rows, int32_t *d_dirvar,
                     uint64_t *d_mse_y, uint64_t *d_mse_uv, zav1_stream_t s);

/* The CDEF frame pass, decoder-exact: apply_cdef_frame / apply_cdef_frame_hbd
 * (rust/crates/svtav1-encoder/src/cdef.rs:343 / :556), out of place.  The
 * d_in_* planes are the post-deblock reconstruction (read only: every read
 * is of the pre-CDEF value, as the reference's snapshot gives), d_out_*
 * receive the filtered planes; every output pixel is written (the input
 * where nothing filters), so outputs must not alias inputs.  Per 64x64
 * filter block (raster, ceil(w/64) x ceil(h/64)) d_fb_strength[2 fb] and
 * [2 fb + 1] are the luma and chroma strength indices the block applies
 * (CdefPick::fb_strengths: primary idx / 4, secondary idx % 4 with 3 ->
 * 4); damping is the frame's cdef damping (3..6; + bd - 8 inside).  d_skip
 * (may be NULL = no skip) is the per-mi signaled skip map, mi_rows x
 * mi_cols: 8x8s whose mi are all skip are not filtered.  Chroma (d_in_u /
 * d_in_v, both or neither) is 4:2:0 with the luma directions.  width and
 * height multiples of 8.  d_stats (may be NULL): CdefStats, [0] pixels of
 * every filter invocation, [1] of those the pixels that changed (at both
 * bit depths).  The frame-level gate (CdefPick::any) is the caller's: with
 * all strengths 0 the outputs equal the inputs. */
int zav1_cdef_apply(const void *d_in_y, const void *d_in_u, const void *d_in_v, void *d_out_y, void *d_out_u,
                    void *d_out_v, int width, int height, int bd, int damping, const uint8_t *d_fb_strength,
                    const uint8_t *d_skip, int mi_cols, int mi_rows, uint64_t *d_stats, zav1_stream_t s);

/* Restoration-unit grid of a plane (count_units_in_tile,
 * rust/crates/svtav1-dsp/src/restoration.rs:1388).  Host only. */
int zav1_rest_unit_grid(int unit_size, int plane_w, int plane_h, int *n_h, int *n_v);

/* Wiener statistics of every restoration unit (raster order of the grid
 * above; unit limits of foreach_rest_unit_in_tile, restoration.rs:1395,
 * with the 8 >> ss_y row offset), d_M[u*win^2 + k], d_H[u*win^4 + k*win^2
 * + l]; d_dgd is read with edge replication beyond the plane.  win 7
 * (luma) or 5 (chroma); unit_size <= 256.  Replaces compute_stats /
 * compute_stats_hbd (restoration.rs:232 / :1738) with find_average(_hbd)
 * (:198 / :1707). */
int zav1_wiener_stats(const void *d_dgd, const void *d_src, int plane_w, int plane_h, int bd, int win, int unit_size,
                      int ss_y, int64_t *d_M, int64_t *d_H, zav1_stream_t s);

/* SSE of two planes over every restoration unit (d_out[u]); at 10 bits the
 * sum of highbd_variance's parts, each truncated to 32 bits (the right
 * strip of width % 16 columns, the bottom strip left of it, the 16x16
 * blocks).  Replaces sse_region / sse_region_hbd (restoration.rs:1666 /
 * :1991). */
int zav1_unit_sse(const void *d_a, const void *d_b, int plane_w, int plane_h, int bd, int unit_size, int ss_y,
                  uint64_t *d_out, zav1_stream_t s);

/* One restoration unit's Wiener filter (WienerInfo, restoration.rs:95):
 * the 8-tap rows as signaled (taps 0..2 mirrored, centre -2 (t0 + t1 + t2),
 * tap 7 = 0) and the unit's type (0 RESTORE_NONE, 1 RESTORE_WIENER). */
typedef struct {
  int16_t hfilter[8];
  int16_t vfilter[8];
  uint8_t type;
  uint8_t pad[7];
} zav1_wiener_unit;

/* The Wiener unit filter over every restoration unit of a plane (grid and
 * limits as zav1_rest_unit_grid / zav1_wiener_stats; d_units[u] in raster
 * order): loop_restoration_filter_unit without stripe boundaries
 * (rust/crates/svtav1-dsp/src/restoration.rs:1288 with need_boundaries = 0,
 * its 10-bit twin loop_restoration_filter_unit_search_hbd :1921, over
 * wiener_convolve_add_src(_hbd) :138 / :1820), reading d_dgd edge-extended
 * by 3 (extend_frame) and writing every pixel of d_out (a RESTORE_NONE unit
 * is copied).  unit_size 32, 64, 128 or 256; ss_y the plane's vertical
 * subsampling. */
int zav1_wiener_apply(const void *d_dgd, void *d_out, int plane_w, int plane_h, int bd, int unit_size, int ss_y,
                      const zav1_wiener_unit *d_units, zav1_stream_t s);

/* finer_tile_search_wiener (rust/crates/svtav1-encoder/src/restoration.rs:
 * 513) for every unit of a plane at once: from each unit's taps in d_units
 * (the solver's filter), the SSE of the unit filtered (as
 * zav1_wiener_apply) against d_src, then, with use_refinement, the +-step
 * hill climb over taps (7 - win) / 2 .. 2 of hfilter, then vfilter (steps
 * 4, 2, 1; 4 only with max_one_step).  Writes the refined taps back into
 * d_units and the final SSE into d_err[u].  One workgroup per unit. */
int zav1_wiener_refine(const void *d_dgd, const void *d_src, int plane_w, int plane_h, int bd, int win, int unit_size,
                       int ss_y, int use_refinement, int max_one_step, zav1_wiener_unit *d_units, int64_t *d_err,
                       zav1_stream_t s);

/* Self-guided projection statistics of every restoration unit and each of
 * the 16 AV1 parameter sets: d_out[(u*16 + set)*5 + {0..4}] = H00, H01,
 * H11, C0, C1 of the projection normal equations, each summed over the unit
 * and divided by its area.  No counterpart in the reference (its encoder
 * never searches self-guided restoration); offered for an encoder that
 * adds that search. */
int zav1_sgr_stats(const void *d_dgd, const void *d_src, int plane_w, int plane_h, int bd, int unit_size, int ss_y,
                   int64_t *d_out, zav1_stream_t s);


/* ------------------------------------------------------------------------ */
/* Partition pre-pass (light PD0 of the allintra presets).                   */
/* ------------------------------------------------------------------------ */
/* The partition tree of every 64x64 unit of a batch of stills, as the
 * reference's PD0 pass decides it (rust/crates/svtav1-encoder/src/pd0.rs):
 * the 85-entry variance map (compute_b64_variance :81), the pass level per
 * unit, each tested square block's cost -- DC prediction from the source
 * neighbours, DCT_DCT, quantize_b, frequency distortion, coefficient rate --
 * and the parent-first walk with its early exits (Pd0Ctx::pick :1702).
 * Replaces pd0_pick_sb_partition (:1890, eff-M9: level 5),
 * pd0_pick_sb_partition_lvl0 (:1990, level 0) and
 * pd0_pick_sb_partition_m6_eval (:2144, presets 2..8: level 1) for every
 * unit of the batch in one call.  The plane is stacked like the wavefront's
 * (n_images x height rows, stride width); samples are 8-bit, a 10-bit plane
 * enters MSB-truncated (v >> 2) as the reference's pass reads it.  Image
 * sides must be multiples of 64. */
typedef struct zav1_pd0_cfg {
  int32_t width, height;          /* per image, multiples of 64 */
  int32_t n_images;               /* images stacked vertically */
  int32_t bd;                     /* samples of the plane: 8 (u8) or 10 (u16) */
  int32_t level;                  /* 0 LVL_0, 1 LVL_1, 5 eff-M9 (LVL_5 or LVL_6 per unit by the detector) */
  int32_t qp, qindex;             /* CLI qp (threshold scaling), frame qindex */
  int32_t ires;                   /* input_resolution_factor (LVL_0 / LVL_5 rate, pd0.rs:371) */
  int32_t cap_max_block;          /* LVL_1: the 64x64 variance cap (preset >= 8) */
  int32_t max_tx;                 /* 64, or 32 (max_tx_size) */
  int32_t min_sq;                 /* 8, or 4 where 4x4 blocks are allowed (presets <= 3) */
  int32_t coeff_rate_est_lvl;     /* LVL_1: 1 exact rate, 2 the approximation for small eobs */
  int32_t tile_w, tile_h;         /* LVL_1 tile size in pixels for neighbour availability; 0 = one tile */
  int32_t split_bits[4];          /* PARTITION_SPLIT rate at 8, 16, 32, 64 (sub-context 0) */
  int32_t none_bits, skip0_bits;  /* PARTITION_NONE at context 0, skip = 0 at context 0 */
  int32_t txt_dc[3];              /* DCT_DCT intra tx-type rate at DC_PRED: 4x4, 8x8, 16x16 */
  int32_t pad;
  uint64_t lambda;                /* kf_full_lambda_8bit(qindex, qp) (pd0.rs:224) */
} zav1_pd0_cfg;
typedef struct zav1_pd0_sb {
  uint16_t var[85];               /* 64x64, 32x32 (2x2), 16x16 (4x4), 8x8 (8x8), raster per level */
  uint8_t mode;                   /* level used: 0, 1, 5, 6 */
  uint8_t max_sq;                 /* largest tested square */
  uint8_t subres;                 /* LVL_5: 1 when the 64x64 passed the odd/even check */
  uint8_t pad[3];
} zav1_pd0_sb;
/* A node of the walk (the record the depth refinement reads): index 0 the
 * 64x64, 1..4 the 32x32s, 5..20 the 16x16s, 21..84 the 8x8s, 85..340 the
 * 4x4s (min_sq 4 only), raster within a level. */
#define ZAV1_PD0_VISITED 1
#define ZAV1_PD0_TESTED 2
#define ZAV1_PD0_SPLIT 8
typedef struct zav1_pd0_node {
  uint64_t cost;                  /* the PART_N cost when tested, else 0 */
  uint8_t flags, pad[7];
} zav1_pd0_node;
/* Fills the rate fields, lambda and ires of cfg from the AV1 default CDFs
 * (build_m6_pd0_tables, pd0.rs:1295; partition_symbol_cost, skip0_bits
 * :1460-1489); width, height, qp and qindex must be set.  Host only. */
int zav1_pd0_rates(zav1_pd0_cfg *cfg);
/* Nodes per unit: 85, or 341 with min_sq 4. */
int zav1_pd0_nodes_per_unit(const zav1_pd0_cfg *cfg);
/* Runs the pass: d_sb[u], d_nodes[u * nodes_per_unit + i] and d_leaf_map
 * [u * 256 + r * 16 + c] (log2 of the leaf size covering 4x4 cell (r, c))
 * for every unit u in raster order of the stacked plane.  d_costs: the
 * frame's coefficient cost tables (zav1_build_coeff_costs at qindex; level 1
 * only, may be null otherwise).  Needs zav1_init on the device; allocates
 * nothing. */
int zav1_pd0_run(const zav1_pd0_cfg *cfg, const void *d_plane, const zav1_coeff_costs *d_costs, zav1_pd0_sb *d_sb,
                 zav1_pd0_node *d_nodes, uint8_t *d_leaf_map, zav1_stream_t s);


/* ------------------------------------------------------------------------ */
/* AV1 bitstream of a still image's luma decisions (host).                   */
/* ------------------------------------------------------------------------ */
/* One leaf of a still: a square block whose transform is the block, its
 * intra mode info and its quantized levels (packed min(size, 32)^2 raster,
 * the coded top-left 32x32 of a 64x64). */
typedef struct zav1_still_leaf {
  int32_t x, y;                   /* luma position, pixels */
  uint8_t size;                   /* 8, 16, 32 or 64 */
  uint8_t mode;                   /* intra mode 0..12 */
  int8_t delta;                   /* angle delta -3..3 (directional modes) */
  uint8_t fi_mode;                /* filter-intra mode 0..4, 5 = none */
  uint8_t tx_type;                /* in the size's intra ext-tx set */
  uint8_t pad[3];
  int32_t eob;                    /* scan positions coded (the last nonzero); 0 = skip */
  int32_t pad2;
  uint64_t level_off;             /* its levels at levels + level_off */
} zav1_still_leaf;
typedef struct zav1_still_desc {
  int32_t width, height;          /* multiples of the superblock size */
  int32_t bd;                     /* 8 or 10 */
  int32_t sb;                     /* 64 or 128 */
  int32_t tile_cols, tile_rows;   /* uniform tiles, powers of two within AV1's tile limits */
  int32_t qindex;                 /* 1..255 */
  int32_t n_leaves;
  int32_t threads;                /* host threads coding tiles in parallel; 0 = all cores */
  int32_t pad;
  const zav1_still_leaf *leaves;  /* any order; together they cover the frame once */
  const int32_t *levels;
} zav1_still_desc;
/* The AV1 stream (temporal delimiter, reduced still-picture sequence header,
 * one frame OBU) of a monochrome key frame coding these leaves: partition
 * (NONE / SPLIT quadtrees), skip, key-frame y mode, angle delta, filter
 * intra, tx type and coefficients with adaptive CDFs, each tile with fresh
 * CDFs and contexts, TX_MODE_LARGEST, no loop filter / CDEF / restoration.
 * Replaces the entropy walk and OBU packing of the reference
 * (rust/crates/svtav1-encoder/src/pipeline.rs:2542-2811 over
 * rust/crates/svtav1-entropy/src/range_coder.rs, coeff_c.rs:924, obu.rs) for
 * the luma plane.  Writes *size bytes into out; ZAV1_ENOMEM (with *size the
 * stream's length) when cap is smaller.  Host only. */
int zav1_write_still(const zav1_still_desc *d, uint8_t *out, size_t cap, size_t *size);
/* The streams of images [first, first + n) of a wavefront batch run with
 * keep_levels, back to back in out; sizes[i] each.  The batch's tiles must
 * be powers of two and its sides multiples of the superblock.  Host-side
 * coding after the device results are read back on stream s (synchronous):
 * whole images on `threads` host threads (0 = all cores), an image's tiles
 * split among the threads left over when n < threads.  A cap short of the
 * streams' total returns ZAV1_ENOMEM with every sizes[i] set (nothing
 * written; call again with a larger buffer). */
int zav1_wf_write_stills(zav1_wf *wf, int first, int n, int threads, uint8_t *out, size_t cap, size_t *sizes,
                         zav1_stream_t s);

#ifdef __cplusplus
}
#endif
#endif
