/*
 * Oracle: MDS0 fast cost and survivor ordering of the mode-decision funnel.
 * TEST INFRASTRUCTURE ONLY (see zav1_oracle.h).
 *
 *   fast cost   lf/mod.rs:2181-2185: rdcost(lambda, rate, satd << 4);
 *               rdcost lf/rate_tables.rs:1033 = ((rate*lambda + 256) >> 9)
 *               + (dist << 7), u64 arithmetic.
 *   ordering    lf/rate_tables.rs:108 c_exchange_sort_by (C
 *               sort_fast_cost_based_candidates): for i, for j > i, swap
 *               idx[i], idx[j] when cost[idx[j]] < cost[idx[i]]; applied to
 *               the candidates in generation order (lf/mod.rs:2994).
 */
#include <stdlib.h>

#include "zav1_oracle.h"

uint64_t or_rdcost(uint64_t lambda, uint64_t rate, uint64_t dist) { return ((rate * lambda + 256) >> 9) + (dist << 7); }

void or_exchange_sort(uint64_t *cost_of, int n, int *idx) {
  for (int i = 0; i < n; i++) idx[i] = i;
  for (int i = 0; i + 1 < n; i++)
    for (int j = i + 1; j < n; j++)
      if (cost_of[idx[j]] < cost_of[idx[i]]) {
        int t = idx[i];
        idx[i] = idx[j];
        idx[j] = t;
      }
}

void or_md_fast_select(const uint32_t *satd, const uint32_t *rate, uint64_t lambda, int n_units, int n_cands, int k,
                       uint8_t *sel, uint64_t *sel_cost) {
  uint64_t cost[64];
  int idx[64];
  for (int u = 0; u < n_units; u++) {
    for (int m = 0; m < n_cands; m++) {
      const uint32_t s = satd[(size_t)u * n_cands + m];
      cost[m] = s == 0xffffffffu ? ~0ull : or_rdcost(lambda, rate ? rate[(size_t)u * n_cands + m] : 0, (uint64_t)s << 4);
    }
    or_exchange_sort(cost, n_cands, idx);
    for (int r = 0; r < k; r++) {
      sel[(size_t)u * k + r] = (uint8_t)idx[r];
      sel_cost[(size_t)u * k + r] = cost[idx[r]];
    }
  }
}

/* The MDS0 -> MDS1 chain over units (unit-major jobs, n_cands per unit,
 * each job's out_off ignored): predict every candidate from `recon`
 * (or_intra_predict), SATD against `src` (or_block_satd), fast select,
 * then the transform unit of each survivor with tx_type (quantize_b when
 * use_b, frequency distortion).  Outputs per (unit, rank): sel, sel_cost,
 * dist, eob, cul.  OpenMP over units. */
void or_md_funnel(const void *src, const void *recon, int stride, int bd, const or_intra_job *jobs, int n_units,
                  int n_cands, const uint32_t *rate, uint64_t lambda, int k, int tx_size, int tx_type, int qindex,
                  int use_b, uint8_t *sel, uint64_t *sel_cost, uint64_t *dist, int32_t *eob, uint8_t *cul) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int u = 0; u < n_units; u++) {
    const or_intra_job *uj = jobs + (size_t)u * n_cands;
    const int w = uj[0].w, h = uj[0].h, hbd = bd > 8;
    int32_t p32[64 * 64];
    uint16_t *pred = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)n_cands * w * h);
    uint8_t p8[64 * 64];
    uint32_t satd[64];
    const size_t so = (size_t)uj[0].y * stride + uj[0].x;
    const void *sp = hbd ? (const void *)((const uint16_t *)src + so) : (const void *)((const uint8_t *)src + so);
    for (int m = 0; m < n_cands; m++) {
      or_intra_predict(recon, stride, bd, uj + m, p32);
      uint16_t *pm = pred + (size_t)m * w * h;
      for (int i = 0; i < w * h; i++) {
        pm[i] = (uint16_t)p32[i];
        p8[i] = (uint8_t)p32[i];
      }
      satd[m] = (uint32_t)or_block_satd(sp, stride, hbd ? (const void *)pm : (const void *)p8, w, w, h, hbd);
    }
    or_md_fast_select(satd, rate ? rate + (size_t)u * n_cands : NULL, lambda, 1, n_cands, k, sel + (size_t)u * k,
                      sel_cost + (size_t)u * k);
    for (int r = 0; r < k; r++) {
      const size_t o = (size_t)u * k + r;
      int e = 0;
      or_tx_unit_full(sp, stride, hbd, pred + (size_t)sel[o] * w * h, tx_size, tx_type, qindex, bd, use_b, 0, 0, w,
                      h, NULL, NULL, dist + o, &e, cul + o);
      eob[o] = e;
    }
    free(pred);
  }
}
