/* Oracle (TEST INFRASTRUCTURE ONLY — see zav1_oracle.h): the leaf funnel
 * over a uniform partition, sequential in the reference's coding order.
 *
 *   coding order          enc/pipeline.rs:7588-7598: tiles in order,
 *                         superblocks in raster order within a tile, leaves
 *                         in Z-order within a superblock
 *   availability          enc/intra_edge.rs:355 has_top_right / :436
 *                         has_bottom_left: coded earlier, inside the tile
 *   MDS0                  lf/mod.rs:2060-2185: every candidate predicted
 *                         from the reconstruction, Hadamard SATD, fast cost
 *                         rdcost(lambda, mode rate, satd << 4); mode rates
 *                         lf/rate_tables.rs:146 under the above / left mode
 *                         contexts; exchange sort lf/rate_tables.rs:108
 *   MDS1                  lf/mod.rs:3067-3110: DCT_DCT, quantize_b,
 *                         frequency distortion, exact coefficient rate;
 *                         post-MDS1 exchange sort lf/mod.rs:3236
 *   MDS3 tx-type search   lf/mod.rs:7140-7390: TX_TYPE_GROUPS order within the
 *                         intra ext-tx set, rate gate, coefficient-SATD early
 *                         exit, no eob-0 non-DCT, first strictly smaller cost
 *   winner                least full cost (mode rate + coefficient rate,
 *                         spatial distortion), first in order
 * (lf/ = rust/crates/svtav1-encoder/src/leaf_funnel/, enc/ = .../src/).
 * Every piece is an oracle function pinned elsewhere (intra prediction,
 * SATD, transforms, quantizers, RDOQ, coefficient rate, rates). */
#include <stdlib.h>
#include <string.h>

#include "zav1_oracle.h"

static const int TX_OF_LEAF[65] = {[8] = 1, [16] = 2, [32] = 3, [64] = 4};
static const int FI_DIR[5] = {0, 1, 2, 6, 0};
static const int MODE_CTX[13] = {0, 1, 2, 3, 4, 4, 4, 4, 3, 0, 1, 2, 0};

static int ld(const void *p, int hbd, long i) { return hbd ? ((const uint16_t *)p)[i] : ((const uint8_t *)p)[i]; }
static void st(void *p, int hbd, long i, int v) {
  if (hbd)
    ((uint16_t *)p)[i] = (uint16_t)v;
  else
    ((uint8_t *)p)[i] = (uint8_t)v;
}

/* Z-order index of leaf (i, j) of a superblock: row bits odd, column bits even */
static int zidx(int i, int j) {
  int z = 0;
  for (int b = 0; b < 8; b++) z |= ((j >> b) & 1) << (2 * b) | ((i >> b) & 1) << (2 * b + 1);
  return z;
}

/* The types of the tx search of an L x L intra leaf in TX_TYPE_GROUPS order. */
static int search_types(int tx, int groups, int *out) {
  static const int G[6][8] = {{0}, {10, 11}, {3}, {1, 2}, {6, 9}, {4, 5, 7, 8, 12, 13, 14, 15}};
  static const int GN[6] = {1, 2, 1, 2, 2, 8};
  int n = 0;
  for (int g = 0; g < groups && g < 6; g++)
    for (int k = 0; k < GN[g]; k++)
      if (or_intra_tx_type_used(tx, G[g][k])) out[n++] = G[g][k];
  return n;
}

typedef struct {
  uint64_t dist, csatd;
  int eob, rate;
  int32_t rec[64 * 64];
  int32_t lev[32 * 32]; /* quantized levels, min(L, 32)^2 raster */
} rd_out;

/* One transform unit of the funnel (tx_unit_inner with do_rdoq and the exact
 * rate): quantize_b (rdoq level 0) or quantize_fp + optimize_b, rate with the
 * tx-type rate (skip cost at eob 0), distortion frequency-domain or spatial
 * with the reconstruction.  est2: coeff_rate_est_lvl 2, the luma rate of an
 * eob below L*L/64 is 6000 + 1000 * eob
 * (rust/crates/svtav1-encoder/src/leaf_funnel/tx_pipeline.rs:605-616).  A
 * 64x64 unit keeps its top-left 32x32 coefficients (quantized, priced, and
 * the inverse's input) and adds the rest's energy to the frequency
 * distortion (tx_pipeline.rs:335-365). */
static void rd_unit(const void *src, int stride, int hbd, int bd, long off, const int32_t *pred, int L, int tx,
                    int type, int qidx, const or_txb_costs *txb, const or_eob_costs *eobc, int use_rdoq,
                    int cut_num, int cut_den, int64_t rdmult, int txt_rate, int spatial, int est2, rd_out *o) {
  const int n = L * L, ls = or_tx_scale(tx), P = L < 32 ? L : 32, np = P * P;
  static __thread int32_t res[64 * 64], coef[64 * 64], inv[64 * 64], dqf[64 * 64];
  int32_t pk[1024], q[1024], dq[1024];
  for (int r = 0; r < L; r++)
    for (int c = 0; c < L; c++) res[r * L + c] = ld(src, hbd, off + (long)r * stride + c) - pred[r * L + c];
  or_fwd_txfm2d(res, coef, L, tx, type);
  uint64_t cs = 0, fold = 0;
  for (int i = 0; i < n; i++) cs += (uint64_t)(coef[i] < 0 ? -(int64_t)coef[i] : coef[i]);
  o->csatd = cs;
  for (int r = 0; r < L; r++)
    for (int c = 0; c < L; c++) {
      if (r < P && c < P) pk[r * P + c] = coef[r * L + c];
      else fold += (uint64_t)((int64_t)coef[r * L + c] * coef[r * L + c]);
    }
  or_quant_row row;
  or_build_quant_row(qidx, bd, &row);
  uint16_t scan[1024];
  or_scan(tx, or_scan_class(type), scan);
  if (use_rdoq)
    or_quantize_fp(pk, np, &row, ls, bd, q, dq);
  else
    or_quantize_b(pk, np, &row, ls, bd, q, dq);
  int eob = or_eob(q, scan, np);
  const int txs = or_txs_ctx(tx), ems = or_eob_multi_size(tx);
  const int cls = type < 10 ? 0 : ((type & 1) ? 1 : 2);
  if (use_rdoq && eob)
    eob = or_optimize_b(pk, q, dq, eob, scan, row.dequant, &txb[txs * 2], &eobc[ems * 2], rdmult, tx, cls, 0, 0,
                        0, cut_num, cut_den);
  o->eob = eob;
  memcpy(o->lev, q, sizeof(int32_t) * np);
  o->rate = eob ? or_cost_coeffs_txb(q, eob, scan, &txb[txs * 2], &eobc[ems * 2], tx, cls, 0, 0, 0, txt_rate)
                : txb[txs * 2].txb_skip_cost[0][1];
  if (est2 && eob < (n >> 6)) o->rate = 6000 + 1000 * eob;
  if (eob) {
    memset(dqf, 0, sizeof(int32_t) * n);
    for (int r = 0; r < P; r++)
      for (int c = 0; c < P; c++) dqf[r * L + c] = dq[r * P + c];
    or_inv_txfm2d(dqf, inv, L, tx, type, bd);
  } else {
    memset(inv, 0, sizeof(int32_t) * n);
  }
  const int mx = (1 << bd) - 1;
  uint64_t sse = 0, fd = fold;
  for (int i = 0; i < n; i++) {
    int v = pred[i] + inv[i];
    v = v < 0 ? 0 : (v > mx ? mx : v);
    o->rec[i] = v;
    const int64_t e = (int64_t)ld(src, hbd, off + (long)(i / L) * stride + i % L) - v;
    sse += (uint64_t)(e * e);
  }
  for (int i = 0; i < np; i++) {
    const int64_t f = (int64_t)pk[i] - dq[i];
    fd += (uint64_t)(f * f);
  }
  const int sh = (1 - ls) * 2;
  o->dist = spatial ? sse << 4 : (sh >= 0 ? fd >> sh : fd << -sh);
}

typedef struct {
  int x, y, L, X0, X1, Y0, Y1;
} wf_leaf;

/* Leaves of one 64x64 unit in Z-order from its 16 x 16 map of log2 sizes. */
static int unit_leaves(const uint8_t *map, int sq, int ox, int oy, int ux, int uy, wf_leaf *out, int n) {
  if ((1 << map[(oy / 4) * 16 + ox / 4]) == sq) {
    out[n].x = ux + ox, out[n].y = uy + oy, out[n].L = sq;
    return n + 1;
  }
  const int h = sq / 2;
  for (int i = 0; i < 4; i++) n = unit_leaves(map, h, ox + (i & 1) * h, oy + (i >> 1) * h, ux, uy, out, n);
  return n;
}

void or_wf_frame(const or_wf_cfg *cfg, const void *src, void *recon, or_wf_decision *dec) {
  or_wf_frame_levels(cfg, src, recon, dec, NULL);
}

void or_wf_frame_levels(const or_wf_cfg *cfg, const void *src, void *recon, or_wf_decision *dec,
                        int32_t *levels) {
  const int W = cfg->width, H = cfg->height, SB = cfg->sb, bd = cfg->bd, hbd = bd > 8;
  const int tree = cfg->leaf_map != NULL;
  const int gw8 = W / 8, gh8 = H / 8, uw = W / 64;
  or_txb_costs txb[10];
  or_eob_costs eobc[7][2];
  or_build_coeff_costs(cfg->qindex, txb, eobc);
  const int64_t rdmult = ((int64_t)cfg->lambda * 17 + 2) >> 2;
  const int use_rdoq = cfg->rdoq_level > 0;
  const int cut_num = cfg->rdoq_level >= 2 ? (cfg->rdoq_level == 2 ? 80 : 60) : 0;
  const int cut_den = cfg->rdoq_level >= 2 ? 100 : 0;
  const int sbc = (W + SB - 1) / SB, sbr = (H + SB - 1) / SB;
  const int tcols = cfg->tile_cols, trows = cfg->tile_rows;
  const int tcs = (sbc + tcols - 1) / tcols, trs = (sbr + trows - 1) / trows;
  uint8_t *modes = (uint8_t *)malloc((size_t)gw8 * gh8);
  int32_t *pred = (int32_t *)malloc(sizeof(int32_t) * 64 * 64 * 64);
  rd_out *o3 = (rd_out *)malloc(sizeof(rd_out) * 16 * 16);
  uint8_t *pp = (uint8_t *)malloc((size_t)2 * 64 * 64);
  wf_leaf *lv = (wf_leaf *)malloc(sizeof(wf_leaf) * (size_t)(SB / 8) * (SB / 8));
  for (int img = 0; img < cfg->n_images; img++) {
    const void *s = (const char *)src + (size_t)img * W * H * (hbd ? 2 : 1);
    void *rc = (char *)recon + (size_t)img * W * H * (hbd ? 2 : 1);
    const int L0 = cfg->leaf;
    or_wf_decision *dimg = dec + (size_t)img * (tree ? (size_t)gw8 * gh8 : (size_t)(W / L0) * (H / L0));
    const uint8_t *imap = tree ? cfg->leaf_map + (size_t)img * (W / 64) * (H / 64) * 256 : NULL;
    for (int ty = 0; ty < trows; ty++)
      for (int tx0 = 0; tx0 < tcols; tx0++) {
        const int cx0 = tx0 * tcs < sbc ? tx0 * tcs : sbc, cx1 = (tx0 + 1) * tcs < sbc ? (tx0 + 1) * tcs : sbc;
        const int ry0 = ty * trs < sbr ? ty * trs : sbr, ry1 = (ty + 1) * trs < sbr ? (ty + 1) * trs : sbr;
        const int X0 = cx0 * SB, X1 = cx1 * SB < W ? cx1 * SB : W, Y0 = ry0 * SB, Y1 = ry1 * SB < H ? ry1 * SB : H;
        for (int r = ry0; r < ry1; r++)
          for (int q = cx0; q < cx1; q++) {
            /* the superblock's leaves in coding order */
            int nlv = 0;
            if (tree) {
              for (int u = 0; u < (SB / 64) * (SB / 64); u++) {
                const int ux = q * SB + (u & 1) * 64, uy = r * SB + (u >> 1) * 64;
                if (ux >= W || uy >= H) continue;
                nlv = unit_leaves(imap + (size_t)((uy / 64) * uw + ux / 64) * 256, 64, 0, 0, ux, uy, lv, nlv);
              }
            } else {
              const int nl = SB / L0;
              for (int z = 0; z < nl * nl; z++) {
                int i = 0, j = 0;
                for (int b = 0; b < 8; b++) {
                  j |= ((z >> (2 * b)) & 1) << b;
                  i |= ((z >> (2 * b + 1)) & 1) << b;
                }
                const int x = q * SB + j * L0, y = r * SB + i * L0;
                if (x >= W || y >= H) continue;
                lv[nlv].x = x, lv[nlv].y = y, lv[nlv].L = L0;
                nlv++;
              }
            }
            for (int li = 0; li < nlv; li++) {
              const int x = lv[li].x, y = lv[li].y, L = lv[li].L, tx = TX_OF_LEAF[L], nl = SB / L;
              const int i = (y % SB) / L, j = (x % SB) / L;
              /* the leaf size's funnel: no filter intra on 64x64 (the
               * candidates carrying it come last), tx types of its groups */
              int nc = cfg->n_cands;
              while (L > 32 && nc > 0 && cfg->cands[nc - 1].fi_mode < 5) nc--;
              const int k1 = cfg->k1 < nc ? cfg->k1 : nc, k3 = cfg->k3 < k1 ? cfg->k3 : k1;
              int types[16];
              const int nt = search_types(tx, (tree && L < 16) ? cfg->txt_groups_lt16 : cfg->txt_groups, types);
              const int64_t satd_th = nt > 1 ? cfg->txt_satd_th : 0;
              const int up = y > Y0, lf = x > X0, zi = zidx(i, j);
              int tr = 0, bl = 0;
              if (up) tr = i == 0 ? 1 : (j + 1 < nl && zidx(i - 1, j + 1) < zi);
              if (lf) bl = j == 0 ? (i + 1 < nl) : (i + 1 < nl && zidx(i + 1, j - 1) < zi);
              int ntr = tr ? X1 - (x + L) : 0, nbl = bl ? Y1 - (y + L) : 0;
              ntr = ntr < 0 ? 0 : (ntr > L ? L : ntr);
              nbl = nbl < 0 ? 0 : (nbl > L ? L : nbl);
              const int am = up ? modes[((y - 1) / 8) * gw8 + x / 8] : 0;
              const int lm = lf ? modes[(y / 8) * gw8 + (x - 1) / 8] : 0;
              const int filt = (up && am >= 9 && am <= 11) || (lf && lm >= 9 && lm <= 11);
              /* MDS0 */
              uint64_t fc[64], c1[64];
              int mrate[64], ord[64], ord3[64];
              for (int m = 0; m < nc; m++) {
                const or_wf_cand *cd = &cfg->cands[m];
                or_intra_job jb;
                memset(&jb, 0, sizeof(jb));
                jb.x = x, jb.y = y, jb.w = (uint8_t)L, jb.h = (uint8_t)L, jb.mode = cd->mode, jb.delta = cd->delta;
                jb.fi_mode = cd->fi_mode, jb.edge_filter = 1, jb.filt_type = (uint8_t)filt;
                jb.n_top_px = (int16_t)(up ? L : 0), jb.n_topright_px = (int16_t)ntr;
                jb.n_left_px = (int16_t)(lf ? L : 0), jb.n_bottomleft_px = (int16_t)nbl;
                int32_t *pm = pred + (size_t)m * L * L;
                or_intra_predict(rc, W, bd, &jb, pm);
                for (int k = 0; k < L * L; k++) st(pp, hbd, k, pm[k]);
                const uint64_t satd = or_block_satd((const char *)s + ((size_t)y * W + x) * (hbd ? 2 : 1), W, pp, L,
                                                    L, L, hbd);
                mrate[m] = or_intra_mode_rate(L, L, MODE_CTX[am], MODE_CTX[lm], cd->mode, cd->delta, cd->fi_mode);
                fc[m] = or_rdcost(cfg->lambda, (uint64_t)mrate[m], satd << 4);
              }
              or_exchange_sort(fc, nc, ord);
              /* MDS1 */
              const long off = (long)y * W + x;
              for (int rr = 0; rr < k1; rr++) {
                const int m = ord[rr];
                const or_wf_cand *cd = &cfg->cands[m];
                const int dir = cd->fi_mode < 5 ? FI_DIR[cd->fi_mode] : cd->mode;
                rd_unit(s, W, hbd, bd, off, pred + (size_t)m * L * L, L, tx, 0, cfg->qindex, txb, &eobc[0][0], 0, 0,
                        0, rdmult, or_intra_txt_rate(tx, dir, 0), 0, cfg->coeff_rate_est_lvl == 2, &o3[0]);
                c1[rr] = or_rdcost(cfg->lambda, (uint64_t)mrate[m] + (uint32_t)o3[0].rate, o3[0].dist);
              }
              or_exchange_sort(c1, k1, ord3);
              /* MDS3 + winner */
              uint64_t best_full = ~0ull;
              int win_i = 0, win_t = 0;
              for (int ii = 0; ii < k3; ii++) {
                const int m = ord[ord3[ii]];
                const or_wf_cand *cd = &cfg->cands[m];
                const int dir = cd->fi_mode < 5 ? FI_DIR[cd->fi_mode] : cd->mode;
                uint64_t best = ~0ull, dct_cost = ~0ull;
                int64_t best_satd = INT64_MAX;
                int pick = 0;
                for (int t = 0; t < nt; t++) {
                  rd_out *o = &o3[ii * 16 + t];
                  const int txr = or_intra_txt_rate(tx, dir, types[t]);
                  rd_unit(s, W, hbd, bd, off, pred + (size_t)m * L * L, L, tx, types[t], cfg->qindex, txb,
                          &eobc[0][0], use_rdoq, cut_num, cut_den, rdmult, txr, 1, cfg->coeff_rate_est_lvl == 2, o);
                  if (types[t] != 0 && dct_cost != ~0ull &&
                      or_rdcost(cfg->lambda, (uint64_t)txr, 0) * 1000 > dct_cost * (uint64_t)cfg->txt_rate_th)
                    continue;
                  if (satd_th > 0) {
                    const int64_t sd = (int64_t)o->csatd;
                    if (sd < best_satd)
                      best_satd = sd;
                    else if ((sd - best_satd) * 100 > best_satd * satd_th)
                      continue;
                  }
                  if (o->eob == 0 && types[t] != 0) continue;
                  const uint64_t cst = or_rdcost(cfg->lambda, (uint64_t)(uint32_t)o->rate, o->dist);
                  if (cst < best) {
                    best = cst;
                    pick = t;
                    if (types[t] == 0) dct_cost = cst;
                  } else if (types[t] == 0) {
                    dct_cost = cst;
                  }
                }
                const rd_out *ob = &o3[ii * 16 + pick];
                const uint64_t full = or_rdcost(cfg->lambda, (uint64_t)mrate[m] + (uint32_t)ob->rate, ob->dist);
                if (full < best_full) {
                  best_full = full;
                  win_i = ii;
                  win_t = pick;
                }
              }
              const int m = ord[ord3[win_i]];
              const rd_out *ow = &o3[win_i * 16 + win_t];
              for (int rr = 0; rr < L; rr++)
                for (int cc = 0; cc < L; cc++) st(rc, hbd, (long)(y + rr) * W + x + cc, ow->rec[rr * L + cc]);
              or_wf_decision *d = tree ? &dimg[(y / 8) * gw8 + x / 8] : &dimg[(y / L) * (W / L) + x / L];
              d->cand = (uint8_t)m, d->tx_type = (uint8_t)types[win_t], d->mode = cfg->cands[m].mode;
              d->fi_mode = cfg->cands[m].fi_mode, d->eob = ow->eob, d->rate = ow->rate, d->mode_rate = mrate[m];
              d->dist = ow->dist, d->cost = best_full;
              if (levels) {
                const int P = L < 32 ? L : 32;
                int32_t *lp = levels + (size_t)img * W * H;
                for (int rr = 0; rr < P; rr++)
                  for (int cc = 0; cc < P; cc++) lp[(size_t)(y + rr) * W + x + cc] = ow->lev[rr * P + cc];
              }
              for (int rr = y / 8; rr < (y + L) / 8; rr++)
                for (int cc = x / 8; cc < (x + L) / 8; cc++) modes[rr * gw8 + cc] = cfg->cands[m].mode;
            }
          }
      }
  }
  free(modes);
  free(pred);
  free(o3);
  free(pp);
  free(lv);
}
