/*
 * zav1 oracle — CPU restatement of zenav1-svt's intra-encode DSP kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, its
 * C-ABI, the host drivers) may include, link or call this code.  Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark use it,
 * and only as the checker.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the zenav1-svt checkout):
 *   dsp/ = rust/crates/svtav1-dsp/src/
 *   enc/ = rust/crates/svtav1-encoder/src/
 *   lf/  = rust/crates/svtav1-encoder/src/leaf_funnel/
 *
 * Integer semantics: the reference computes butterflies with i64 products
 * and i32 sums; this file is compiled with -fwrapv so every i32 wrap is
 * defined and equals the reference's two's-complement result.
 *
 * Parity pinning (how each group is checked, see tests/):
 *   1D transforms       golden vectors extracted from C SVT-AV1 by the
 *                       reference (svtav1/tests/golden_parity.rs:36-720),
 *                       cos/sin table closed form, float-DCT closeness.
 *   2D transforms       structure pinned by the 1D goldens; separability/flip
 *                       invariants; the 16x16 / 32x32 / 64x64 DCT_DCT (with
 *                       the 64-fold) end to end by the C encoder's captured
 *                       full-loop records (tests/test_oracle_captures.py).
 *   quantizers          per-coefficient brute force of the quoted formulas,
 *                       dequant bounds; table rows by closed form and a row
 *                       captured from the reference's tests.
 *   hadamard / SATD     brute-force Walsh matrix on small-range input.
 *   scans / eob         sha256 digests of the reference's C-extracted scan
 *                       tables (all sizes x classes).
 *   intra predictors    the reference's own known-answer cases, closed forms
 *                       of the directional projections, constant-edge and
 *                       availability fallbacks, spec edge-filter tables,
 *                       CfL identities; tables generated from the reference's
 *                       data files.
 *   deblocking          thresholds closed form; flat lines unchanged, steps
 *                       smoothed; p/q mirror symmetry of every filter.
 *   CDEF                the reference's constrain / adjust_strength known
 *                       values, find_dir spec semantics, direction table vs
 *                       the spec, flat-block identity, skip-list additivity;
 *                       single impulses isolating every primary / secondary
 *                       tap weight and offset of every direction.
 *   Wiener / unit SSE   numpy covariance of the windowed samples; unit SSE
 *                       and unit limits against the reference's C captures
 *                       (tests/golden/lr).
 *   MDS0 selection      the exchange sort's documented behaviour
 *                       (lf/rate_tables.rs:97-104): a permutation, costs
 *                       non-decreasing, equal to a stable sort on distinct
 *                       keys, and the stated tie displacement on a worked
 *                       case; rdcost by hand-computed values.
 *   self-guided stats   no reference twin (nothing in the reference pins it):
 *                       the specification's Sgr_Params and lookup tables,
 *                       and a whole-array numpy brute force of box sums,
 *                       A/B, both passes and the projection sums.
 *   RDOQ trellis /      the reference's own invariant test (dequant mirror,
 *   coefficient rate    nonzero last coefficient, enc/quant.rs:1852); brute
 *                       force of the single-coefficient block; and end to
 *                       end by the C encoder's captured MDS1 / MDS3 records
 *                       at the frame origin (eob, rate, distortion of
 *                       32x32 / 64x64 / 16x16 blocks at rdoq levels 0, 2, 3
 *                       with the default-CDF tables): every record exact.
 *   rate tables         built from the AV1 default CDFs (fixture from the
 *                       reference's default_cdfs.rs); pinned by the captures
 *                       above and av1_cost_symbol's closed form.
 */
#ifndef ZAV1_ORACLE_H
#define ZAV1_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- enums in C order (typ/transform.rs:7-29, :84-101) ---------------- */
enum {
  OR_TX_4X4 = 0, OR_TX_8X8, OR_TX_16X16, OR_TX_32X32, OR_TX_64X64,
  OR_TX_4X8, OR_TX_8X4, OR_TX_8X16, OR_TX_16X8, OR_TX_16X32, OR_TX_32X16,
  OR_TX_32X64, OR_TX_64X32, OR_TX_4X16, OR_TX_16X4, OR_TX_8X32, OR_TX_32X8,
  OR_TX_16X64, OR_TX_64X16, OR_TX_SIZES_ALL
};

/* ---- transforms ------------------------------------------------------- */
void or_tx_size_dims(int tx_size, int *w, int *h);
/* 1D kernels, exposed for the golden-vector tests.  kind: 0 DCT, 1 ADST,
 * 3 identity.  inverse uses INV_COS_BIT 12 and `range` clamps. */
int or_fwd_txfm1d(int kind, int n, const int32_t *in, int32_t *out, int cos_bit);
int or_inv_txfm1d(int kind, int n, const int32_t *in, int32_t *out, int range);
/* dsp/txfm_dispatch.rs:18 fwd_txfm2d_dispatch: returns 1 ok, 0 unsupported. */
int or_fwd_txfm2d(const int32_t *in, int32_t *out, int stride, int tx_size, int tx_type);
/* dsp/txfm_dispatch.rs:55 inv_txfm2d_dispatch_bd (residual out, stride). */
int or_inv_txfm2d(const int32_t *in, int32_t *out, int stride, int tx_size, int tx_type, int bd);
const int32_t *or_cospi_row(int cos_bit);
const int32_t *or_sinpi_row(int cos_bit);

/* ---- quantizers ------------------------------------------------------- */
typedef struct {
  int32_t zbin[2], round[2], quant[2], quant_shift[2];
  int32_t round_fp[2], quant_fp[2], dequant[2];
} or_quant_row;
/* enc/quant.rs:132-236 build_quant_table(_bd) */
void or_build_quant_row(int qindex, int bd, or_quant_row *row);
/* enc/quant.rs:195 build_quant_table_bd_sharp (sharpness -7..7) */
void or_build_quant_row_sharp(int qindex, int bd, int sharpness, or_quant_row *t);
/* dsp/quant_coding.rs:80 fp_one / enc/quant.rs:286,335 (hbd drops clamp) */
void or_quantize_fp(const int32_t *coeff, int n, const or_quant_row *t, int log_scale, int bd,
                    int32_t *qcoeff, int32_t *dqcoeff);
/* dsp/quant_coding.rs:365 b_one / enc/quant.rs:247,391 */
void or_quantize_b(const int32_t *coeff, int n, const or_quant_row *t, int log_scale, int bd,
                   int32_t *qcoeff, int32_t *dqcoeff);
int or_tx_scale(int tx_size); /* enc/quant.rs:238 TX_SCALE_TAB */

/* ---- residual / distortion ------------------------------------------- */
uint64_t or_sse_i32(const int32_t *a, const int32_t *b, int n);      /* dsp/residual.rs:245 */
uint64_t or_sq_sum_i32(const int32_t *a, int n);                     /* dsp/residual.rs:364 */
uint64_t or_sse_pix(const void *src, int src_stride, const void *ref, int ref_stride,
                    int w, int h, int hbd);                          /* dsp/variance.rs:124 */
void or_hadamard(int n, const int16_t *src, int stride, int32_t *coeff); /* dsp/hadamard.rs:691-913 */
int32_t or_satd(const int32_t *coeff, int n);                           /* dsp/hadamard.rs:942 */
/* lf/predict.rs:139/185 hadamard_satd(_hbd): block SATD over square tiles. */
uint64_t or_block_satd(const void *src, int src_stride, const void *pred, int pred_stride,
                       int w, int h, int hbd);

/* ---- scans / eob / fused transform unit (txsearch.c) ------------------- */
int or_scan_class(int tx_type);
int or_scan(int tx_size, int scan_class, uint16_t *scan);
int or_eob(const int32_t *qcoeff, const uint16_t *scan, int n);
/* tx_pipeline.rs:268 tx_unit_inner (freq-domain distortion, no RDOQ):
 * residual w x h -> dist (after the log_scale shift), eob, sum |level|. */
int or_tx_unit(const int32_t *residual, int tx_size, int tx_type, int qindex, int bd, int use_b,
               int32_t *qcoeff_out, uint64_t *dist, int *eob, uint32_t *level_sum);
/* lf/tx_pipeline.rs:268 tx_unit_inner / :854 tx_unit_hbd for one unit with a
 * given u16 prediction (no RDOQ / QM / rate): see txsearch.c. */
int or_tx_unit_full(const void *src, int src_stride, int hbd, const uint16_t *pred, int tx_size, int tx_type,
                    int qindex, int bd, int use_b, int need_recon, int spatial, int crop_w, int crop_h,
                    int32_t *qcoeff_out, uint16_t *recon_out, uint64_t *dist, int *eob, uint8_t *cul);
int32_t or_dc_value(const int32_t *above, const int32_t *left, int w, int h, int has_above,
                    int has_left, int bd);
void or_txsearch_plane(const uint16_t *plane, int stride, int width, int height, int bd, int tx_size,
                       const int32_t *xy, int n_blocks, const uint8_t *types, int n_types, int qindex,
                       uint64_t *dist, int32_t *eob, uint32_t *lsum);

/* ---- intra prediction (intra.c) ----------------------------------------- */
/* One transform unit to predict.  The host computes the availability counts
 * (C n_top_px / n_topright_px / n_left_px / n_bottomleft_px: real reference
 * samples, capped by the aligned frame / tile extent and the coding-order
 * top-right / bottom-left rules); everything sample-dependent is here. */
typedef struct {
  int32_t x, y;          /* tx-unit origin in the plane (samples) */
  uint8_t w, h;          /* tx dims 4..64 */
  uint8_t mode;          /* 0 DC,1 V,2 H,3 D45,4 D135,5 D113,6 D157,7 D203,8 D67,
                            9 SMOOTH,10 SMOOTH_V,11 SMOOTH_H,12 PAETH */
  int8_t delta;          /* angle delta, -3..3 (x 3 degrees) */
  uint8_t fi_mode;       /* filter-intra mode 0..4, 5 = none */
  uint8_t edge_filter;   /* sequence enable_intra_edge_filter */
  uint8_t filt_type;     /* 1 when an above/left neighbour block is smooth */
  uint8_t pad0;
  int16_t n_top_px, n_topright_px, n_left_px, n_bottomleft_px;
  uint32_t out_off;      /* output offset (samples) */
  uint32_t pad1;
} or_intra_job;
int or_edge_filter_strength(int bs0, int bs1, int delta, int filt_type);
int or_use_upsample(int bs0, int bs1, int delta, int filt_type);
/* lf/predict.rs:49 predict_unit (+ hbd twins): w*h samples into dst (raster). */
int or_intra_predict(const void *plane, int stride, int bd, const or_intra_job *job, int32_t *dst);
/* dsp/intra_pred.rs:810 dr_predictor_edged on final edges (160 samples,
 * origin 16). */
void or_dr_predict_edges(const int32_t *above, const int32_t *left, int w, int h, int angle, int up_a, int up_l,
                         int bd, int32_t *dst);
void or_cfl_predict(const void *luma, int luma_stride, int bd, int cw, int ch, const int32_t *dc, int alpha_q3,
                    int32_t *dst);

/* ---- deblocking / CDEF (lf.c) ------------------------------------------ */
/* Per-4x4 (mi) frame geometry (enc/deblock.rs:532 DeblockGeom). */
typedef struct {
  int mi_cols, mi_rows;     /* luma mi grid */
  int true_w, true_h;       /* coded (unpadded) luma frame size */
  const uint32_t *block_id; /* covering block id per mi */
  const uint8_t *bw, *bh;   /* covering block dims (px) */
  const uint8_t *tw, *th;   /* luma tx dims (px) */
  const uint8_t *skip;      /* inter skip per mi, may be NULL (all-intra) */
} or_deblock_geom;
void or_lf_thresholds(int level, int sharpness, int *mblim, int *lim, int *hev);
void or_lpf(void *buf, int hbd, int bd, long off, int pitch, int vertical, int len, int level, int sharpness);
void or_deblock_plane(void *buf, int stride, int plane_w, int plane_h, int bd, int plane, int ss, int level_v,
                      int level_h, int sharpness, const or_deblock_geom *g);
/* enc/deblock.rs:337 search_filter_level (still path, real trials). */
void or_deblock_search_level(const void *recon, const void *src, int pw, int ph, int bd, int plane, int ss,
                             int sharpness, const or_deblock_geom *g, int conv_th, int *best, int64_t *err0,
                             int64_t *best_err);
int or_cdef_constrain(int diff, int threshold, int damping);
int or_cdef_adjust_strength(int strength, int var);
int or_cdef_find_dir(const uint16_t *img, int stride, int coeff_shift, int *var_out);
/* dsp/cdef.rs:1160 / :1192 compute_cdef_dist_16bit / _8bit */
uint64_t or_cdef_dist(const void *plane, int hbd, long plane_off, int pstride, const void *packed,
                      const uint8_t *dlist, int n, int bsize, int coeff_shift, int sub);
void or_cdef_filter_block(int32_t *dst, int dstride, const uint16_t *in, int pri, int sec, int dir, int pdamp,
                          int sdamp, int bw, int bh, int coeff_shift, int sub);
/* enc/cdef.rs:343 apply_cdef_frame / :556 apply_cdef_frame_hbd (out of place) */
void or_cdef_apply_frame(const void *in_y, const void *in_u, const void *in_v, void *out_y, void *out_u,
                         void *out_v, int width, int height, int bd, int damping, const uint8_t *fb_str,
                         const uint8_t *skip, int mi_cols, int mi_rows, uint64_t *stats);
void or_cdef_search_fb(const void *rec_y, const void *rec_u, const void *rec_v, const void *src_y,
                       const void *src_u, const void *src_v, int width, int height, int bd, int qindex,
                       int subsampling, int fbr, int fbc, const int *fs, int n_fs, uint64_t *mse_y,
                       uint64_t *mse_uv, const uint8_t *skip, int mi_cols, int mi_rows);

/* ---- loop restoration statistics (restoration.c) ------------------------ */
int or_count_units(int unit_size, int size);
void or_unit_limits(int unit_size, int pw, int ph, int ss_y, int ux, int uy, int *hs, int *he, int *vs, int *ve);
void or_wiener_stats(int win, const void *dgd, const void *src, int pw, int ph, int bd, int hs, int he, int vs,
                     int ve, int64_t *M, int64_t *H);
uint64_t or_sse_region(const void *a, const void *b, int pw, int bd, int hs, int he, int vs, int ve);
/* dsp/restoration.rs:1288 / :1921 the Wiener unit filter (no stripe
 * boundaries) over every unit of a plane; enc/restoration.rs:513 the finer
 * tap search of one unit */
void or_wiener_filter_plane(const void *dgd, void *out, int pw, int ph, int bd, int unit_size, int ss_y,
                            const uint8_t *types, const int16_t *filters);
int64_t or_wiener_refine_unit(const void *dgd, const void *src, int pw, int ph, int bd, int win, int hs, int he,
                              int vs, int ve, int use_refinement, int max_one_step, int16_t *hf, int16_t *vf);
int or_sgr_param(int set, int k);
int or_sgr_table(int which, int i);
void or_sgr_unit_stats(const void *dgd, const void *src, int pw, int ph, int bd, int hs, int he, int vs, int ve,
                       int set, int64_t *out);

/* ---- mode-decision funnel, MDS0 (funnel.c) ------------------------------ */
/* lf/rate_tables.rs:1033 rdcost. */
uint64_t or_rdcost(uint64_t lambda, uint64_t rate, uint64_t dist);
/* lf/rate_tables.rs:108 c_exchange_sort_by over idx = 0..n-1. */
void or_exchange_sort(uint64_t *cost_of, int n, int *idx);
/* Per unit: fast costs rdcost(lambda, rate, satd << 4) (0xffffffff SATD ->
 * UINT64_MAX), exchange-sorted; first k indices / costs out. n_cands <= 64. */
void or_md_fast_select(const uint32_t *satd, const uint32_t *rate, uint64_t lambda, int n_units, int n_cands, int k,
                       uint8_t *sel, uint64_t *sel_cost);
/* MDS0 -> MDS1 over units: jobs unit-major (n_cands each, same w x h),
 * predictions from recon, SATD vs src, fast select, survivors' transform
 * units (tx_type, quantize_b when use_b, frequency distortion). */
void or_md_funnel(const void *src, const void *recon, int stride, int bd, const or_intra_job *jobs, int n_units,
                  int n_cands, const uint32_t *rate, uint64_t lambda, int k, int tx_size, int tx_type, int qindex,
                  int use_b, uint8_t *sel, uint64_t *sel_cost, uint64_t *dist, int32_t *eob, uint8_t *cul);

/* Coefficient rate tables (1/512 bit): one (txs_ctx, plane) of C
 * LvMapCoeffCost (enc/quant.rs:459 TxbCosts) and one (eob_multi_size,
 * plane) of LvMapEobCost (:469 EobCosts). */
typedef struct {
  int32_t txb_skip_cost[13][2];
  int32_t base_eob_cost[4][3];
  int32_t base_cost[42][8];
  int32_t eob_extra_cost[9][2];
  int32_t dc_sign_cost[3][2];
  int32_t lps_cost[21][26];
} or_txb_costs;
typedef struct {
  int32_t eob_cost[2][11];
} or_eob_costs;
int or_adjusted_tx_size(int tx_size);
int or_txs_ctx(int tx_size);
int or_eob_multi_size(int tx_size);
int or_rdoq_eob_cost(int eob, const or_eob_costs *e, const or_txb_costs *c, int tx_class);
/* enc/quant.rs:1312 optimize_b without quantization matrices on one packed
 * block (raster, adjusted size); eob > 0; returns the new eob. */
int or_optimize_b(const int32_t *tcoeff, int32_t *qcoeff, int32_t *dqcoeff, int eob, const uint16_t *scan,
                  const int32_t dequant[2], const or_txb_costs *c, const or_eob_costs *e, int64_t rdmult,
                  int tx_size, int tx_class, int txb_skip_ctx, int dc_sign_ctx, int sharpness_flag,
                  int cut_off_num, int cut_off_denum);
/* lf/coeff_rate.rs:60 cost_coeffs_txb (full loop) of one block, eob > 0. */
int or_cost_coeffs_txb(const int32_t *qcoeff, int eob, const uint16_t *scan, const or_txb_costs *c,
                       const or_eob_costs *e, int tx_size, int tx_class, int plane_type, int txb_skip_ctx,
                       int dc_sign_ctx, int txt_rate);

/* Rate tables from the AV1 default CDFs (rate_tables.c):
 * enc/quant.rs:514 build_coeff_cost_tables(base_qindex) into txb[txs_ctx*2 +
 * plane] / eob[eob_multi_size][plane]; context.rs:388 av1_cost_symbol;
 * leaf_funnel/rate_tables.rs:291 txt_rate for an intra tx type. */
int or_coef_q_ctx(int base_qindex);
int or_cost_symbol(unsigned p15);
void or_build_coeff_costs(int base_qindex, or_txb_costs txb[10], or_eob_costs eob[7][2]);
int or_intra_ext_tx_set_type(int tx_size);
int or_intra_tx_type_used(int tx_size, int tx_type);
int or_intra_txt_rate(int tx_size, int intra_dir, int tx_type);
int or_bsize_index(int w, int h);
/* key-frame intra luma candidate rate (rate_tables.rs:146 build_md_rates) */
int or_intra_mode_rate(int w, int h, int actx, int lctx, int mode, int delta, int fi);

/* ---- superblock-wavefront leaf funnel (wavefront.c) --------------------- */
typedef struct {
  uint8_t mode;
  int8_t delta;
  uint8_t fi_mode, pad;
} or_wf_cand;
typedef struct { /* same layout as zav1_wf_cfg */
  int32_t width, height, n_images, bd, leaf, sb, tile_cols, tile_rows, qindex, k1, k3, rdoq_level, txt_groups,
      txt_rate_th, txt_satd_th, coeff_rate_est_lvl;
  uint64_t lambda;
  int32_t n_cands, pad;
  or_wf_cand cands[64];
  int32_t txt_groups_lt16, keep_levels; /* tx-type groups of leaves below 16x16 (partition map only) */
  const uint8_t *leaf_map;       /* NULL: uniform leaves; else per image, per 64x64 unit (raster), 256 log2
                                    leaf sizes (16 x 16 cells of 4x4): the partition pre-pass's map */
} or_wf_cfg;
typedef struct {
  uint8_t cand, tx_type, mode, fi_mode;
  int32_t eob, rate, mode_rate;
  uint64_t dist, cost;
} or_wf_decision;
/* Every leaf of n_images stacked planes (src, recon: n_images * height rows);
 * uniform leaves: dec[img][leaf row][leaf col]; with a partition map:
 * dec[img][y / 8][x / 8] at each leaf's top-left 8x8 cell. */
void or_wf_frame(const or_wf_cfg *cfg, const void *src, void *recon, or_wf_decision *dec);
/* or_wf_frame that also returns the winners' quantized levels: per image a
 * width x height int32 plane, each leaf's min(L, 32)^2 levels (raster) at
 * its top-left corner (the leaf's own area, so leaves never overlap). */
void or_wf_frame_levels(const or_wf_cfg *cfg, const void *src, void *recon, or_wf_decision *dec,
                        int32_t *levels);


/* Coefficient rate with C's mds_fast_coeff_est_level (1 = full loop). */
int or_cost_coeffs_txb_fast(const int32_t *qcoeff, int eob, const uint16_t *scan, const or_txb_costs *c,
                            const or_eob_costs *e, int tx_size, int tx_class, int plane_type, int txb_skip_ctx,
                            int dc_sign_ctx, int txt_rate, int fast_level);
/* enc/quant.rs:494 syntax_rate_from_cdf over one inverse CDF. */
void or_syntax_rate(int32_t *cost, const uint16_t *icdf, int max_syms);

/* ---- partition pre-pass (pd0.c; enc/pd0.rs) ----------------------------- */
typedef struct {
  int level;              /* 0 LVL_0, 1 LVL_1, 5 eff-M9 (detector: LVL_5 / LVL_6 per superblock) */
  int qp, qindex;         /* CLI qp (threshold scaling, lambda weight), frame qindex */
  uint64_t lambda;        /* kf_full_lambda_8bit(qindex, qp) */
  int ires;               /* input_resolution_factor (LVL_0 / LVL_5 closed-form rate) */
  int cap_max_block;      /* LVL_1: the 64x64 variance cap (preset >= 8) */
  int max_tx;             /* 64, or 32 (max_tx_size) */
  int min_sq;             /* 8, or 4 when 4x4 blocks are allowed */
  int coeff_rate_est_lvl; /* LVL_1: 1 exact, 2 approximation for small eobs */
  int tile_w, tile_h;     /* LVL_1 tile size in pixels (0 = one tile) */
  int32_t split_bits[4];  /* PARTITION_SPLIT at 8, 16, 32, 64 */
  int32_t none_bits, skip0_bits;
} or_pd0_cfg;
enum { OR_PD0_VISITED = 1, OR_PD0_TESTED = 2, OR_PD0_SPLIT = 8 };
typedef struct {
  uint64_t cost;          /* the node's PART_N cost when tested */
  uint8_t flags, pad[7];
} or_pd0_node;
void or_b64_variance(const uint8_t *src, int stride, int ox, int oy, uint16_t v[85]);
void or_qp_th_scaling(int qp, uint32_t *qw, uint32_t *qwd);
uint32_t or_kf_full_lambda_8bit(int qindex, int cli_qp);
int or_max_block_size_allintra(int var64, int qp);
int or_input_resolution_factor(long pixels);
int or_pd0_detector_demotes(const uint16_t v[85], int qp);
uint64_t or_lvl6_cost(const uint16_t v[85], int sq, int ox, int oy, int qp);
void or_pd0_default_rates(int32_t split_bits[4], int32_t *none_bits, int32_t *skip0_bits);
/* One block's cost at a given level (mode 0, 1, 5, 6) of the superblock at
 * (sb_x, sb_y); *subres_state carries the LVL_5 subres decision (255
 * undetermined) between calls like the reference's Pd0Ctx. */
uint64_t or_pd0_block_cost_sb(const uint8_t *src, int stride, int sb_x, int sb_y, const or_pd0_cfg *cfg, int mode,
                              int sq, int ox, int oy, int *subres_state);
/* One 64x64 unit: node records (index 0 the 64x64, 1..4 the 32x32s,
 * 5..20 the 16x16s, 21..84 the 8x8s, 85..340 the 4x4s, raster within each
 * level), leaf map (log2 of the leaf size per 4x4 cell, raster 16 x 16) and
 * the level used (0, 1, 5, 6). */
int or_pd0_sb(const uint8_t *src, int stride, int sb_x, int sb_y, const or_pd0_cfg *cfg, or_pd0_node nodes[341],
              uint8_t leaf_map[256], int *mode_out);
/* Every 64x64 unit of a frame (sides multiples of 64), raster order. */
int or_pd0_frame(const uint8_t *src, int stride, int width, int height, const or_pd0_cfg *cfg, or_pd0_node *nodes,
                 uint8_t *leaf_map, uint8_t *modes);


/* ---- AV1 still decoder (av1dec.c; the specification's decoding process) --- */
typedef struct {
  int32_t x, y;                       /* luma position */
  uint8_t size, mode, fi_mode, skip;  /* square block, y mode, filter intra (5 none), skip */
  int8_t delta;                       /* angle delta */
  uint8_t tx_type, pad[2];
  int32_t eob, pad2;
  uint64_t level_off;                 /* its levels (min(size, 32)^2, raster) in the level array */
} or_dec_block;
typedef struct {
  int32_t width, height, bd, sb, qindex, tile_cols, tile_rows, n_blocks;
  int64_t n_levels;
} or_dec_info;
/* Decodes a monochrome reduced-still-picture stream of the library's subset
 * into frame (u8 at 8 bits, u16 at 10, width x height); blocks in decoding
 * order.  0 ok, -1 malformed / too small, -2 outside the subset. */
int or_av1_decode_still(const uint8_t *data, size_t size, void *frame, int max_w, int max_h, or_dec_info *info,
                        or_dec_block *blocks, int max_blocks, int32_t *levels, size_t max_levels);

#ifdef __cplusplus
}
#endif
#endif
