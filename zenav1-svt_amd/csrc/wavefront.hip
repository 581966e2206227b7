// Superblock-wavefront intra mode decision for gfx950: every leaf of a
// uniform partition of one or more still images (and their tiles) decided
// with the reference's funnel -- MDS0 over every intra candidate predicted
// from the real reconstruction, the exchange-sorted survivors into MDS1
// (DCT_DCT, quantize_b, frequency distortion, exact coefficient rate), the
// best of those into MDS3 (the tx-type search with RDOQ and spatial
// distortion, in TX_TYPE_GROUPS order with the rate and SATD gates), the
// winner by full cost, committed to the reconstruction the next leaves read.
//
// Reference: leaf_funnel/mod.rs:1007 evaluate_leaf (MDS0 :2060-2185, the
// exchange sort :2994, MDS1 :3067-3110, the post-MDS1 sort :3236, MDS3's
// tx-type search :7140-7330), rates rate_tables.rs:146 / :291, the coding
// order enc/pipeline.rs:7588-7598 (superblocks in raster order, tiles
// independent) and the neighbour availability of enc/intra_edge.rs:355 /
// :436 (has_top_right / has_bottom_left) -- all paths relative to
// rust/crates/svtav1-encoder/src.
//
// Order: a superblock needs its left, top and top-right neighbours, so the
// superblocks of one anti-diagonal (column + 2 * row within a tile) are
// independent; within a superblock the leaves follow Z-order.  A step is
// (anti-diagonal, leaf index): every superblock of that anti-diagonal, in
// every tile and every image of the batch, decides that leaf together, one
// launch per stage over all of them.  Each stage writes the next stage's
// job list on the device (survivor lists, transform-unit jobs), so a step is
// a fixed chain of launches over device-resident queues; the whole frame's
// chain is captured once into a HIP graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "../../include/zav1_dsp.h"
#include "zav1_dev.h"
#include "zav1_kernels.h"

namespace zav1 {
// the host AV1 still writer (bitstream.cpp): one still into `stream`
int write_still_vec(const zav1_still_desc* d, std::vector<uint8_t>& stream, std::string* why = nullptr);
}  // namespace zav1

namespace zav1 {
namespace {

struct WfCand {
  uint8_t mode;
  int8_t delta;
  uint8_t fi;   // 0..4 filter intra, 5 none
  uint8_t dir;  // intra direction for the tx-type rate (filter intra maps to DC, V, H, D157, DC)
};

struct WfDecision {  // layout of zav1_wf_decision
  uint8_t cand, tx_type, mode, fi;
  int32_t eob, rate, mode_rate;
  uint64_t dist, cost;
};
static_assert(sizeof(WfDecision) == 32, "decision layout");

__host__ __device__ inline uint64_t rdc(uint64_t lambda, uint64_t rate, uint64_t dist) {
  return ((rate * lambda + 256) >> 9) + (dist << 7);
}

// intra_mode_context (AV1): the kf y-mode context of a neighbour's mode
__host__ __device__ inline int mode_ctx(int m) {
  constexpr int8_t c[13] = {0, 1, 2, 3, 4, 4, 4, 4, 3, 0, 1, 2, 0};
  return c[m];
}
__host__ __device__ inline bool is_smooth(int m) { return m >= 9 && m <= 11; }

// The first k positions of the reference's exchange sort (swap whenever a
// later candidate costs strictly less), over n costs in generation order.
__host__ __device__ inline void exchange_first_k(const uint64_t* cost, int n, int k, uint8_t* ord) {
  for (int i = 0; i < n; i++) ord[i] = (uint8_t)i;
  for (int i = 0; i < k; i++) {
    int mi = ord[i];
    uint64_t ci = cost[mi];
    for (int j = i + 1; j < n; j++) {
      const int mj = ord[j];
      if (cost[mj] < ci) {
        ord[j] = (uint8_t)mi;
        mi = mj;
        ci = cost[mj];
      }
    }
    ord[i] = (uint8_t)mi;
  }
}

// The MDS3 tx-type search over one candidate's evaluated types (in search
// order): the rate gate against the DCT cost, the SATD early exit, the
// unsignalable eob-0 rule, the first strictly smaller cost wins
// (leaf_funnel/mod.rs:7228-7390).  Returns the type index and its cost.
__host__ __device__ inline int txt_pick(int nt, const uint8_t* types, const int32_t* txt_row, const uint64_t* dist,
                                        const int32_t* rate, const int32_t* eob, const uint64_t* csatd,
                                        uint64_t lambda, uint64_t rate_th, int64_t satd_th, uint64_t* best_cost) {
  uint64_t best = ~0ull, dct_cost = ~0ull;
  int64_t best_satd = INT64_MAX;
  int pick = 0;
  for (int t = 0; t < nt; t++) {
    const int tt = types[t];
    if (tt != 0 && dct_cost != ~0ull && rdc(lambda, (uint64_t)txt_row[tt], 0) * 1000 > dct_cost * rate_th) continue;
    if (satd_th > 0) {
      const int64_t sd = (int64_t)csatd[t];
      if (sd < best_satd) best_satd = sd;
      else if ((sd - best_satd) * 100 > best_satd * satd_th) continue;
    }
    if (eob[t] == 0 && tt != 0) continue;
    const uint64_t c = rdc(lambda, (uint64_t)(uint32_t)rate[t], dist[t]);
    if (c < best) {
      best = c;
      pick = t;
      if (tt == 0) dct_cost = c;
    } else if (tt == 0) {
      dct_cost = c;
    }
  }
  *best_cost = best;
  return pick;
}

// ---- per-step kernels ------------------------------------------------------

// Mode contexts and the edge-filter type of every unit from its coded
// neighbours, and each candidate's luma rate under those contexts.
__global__ __launch_bounds__(256) void k_wf_prep(MdUnit* __restrict__ units, const int32_t* __restrict__ nbr, int n,
                                                 const uint8_t* __restrict__ mode_map,
                                                 const uint32_t* __restrict__ rate_tab, int nc,
                                                 uint32_t* __restrict__ rate_out) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long)n * nc) return;
  const int u = (int)(i / nc), c = (int)(i - (long)u * nc);
  const int a = nbr[2 * u], l = nbr[2 * u + 1];
  const int am = a >= 0 ? mode_map[a] : 0, lm = l >= 0 ? mode_map[l] : 0;
  if (c == 0) units[u].filt_type = (uint8_t)((a >= 0 && is_smooth(am)) || (l >= 0 && is_smooth(lm)));
  rate_out[i] = rate_tab[(mode_ctx(am) * 5 + mode_ctx(lm)) * nc + c];
}

// MDS1 jobs: every survivor's DCT_DCT unit (quantize_b, frequency distortion).
__global__ __launch_bounds__(256) void k_wf_mds1_jobs(const MdUnit* __restrict__ units, int n,
                                                      const uint8_t* __restrict__ sel, int k1,
                                                      const WfCand* __restrict__ cands,
                                                      const int32_t* __restrict__ txt, int L, int64_t rdmult,
                                                      TxUnitRdJob* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n * k1) return;
  const int u = i / k1;
  const WfCand c = cands[sel[i]];
  TxUnitRdJob j{};
  j.x = units[u].x, j.y = units[u].y, j.pred_off = (uint32_t)i * (uint32_t)(L * L);
  j.tx_type = 0, j.crop_w = (uint8_t)L, j.crop_h = (uint8_t)L, j.plane_type = 0;
  j.txt_rate = txt[c.dir * 16 + 0];
  j.rdmult = rdmult;
  out[i] = j;
}

// After MDS1: full costs of the survivors, the exchange sort's first k3 into
// MDS3 (lane-parallel, one wave per unit, lane = survivor), and their
// tx-type search jobs (every type of the intra set, in search order).
__global__ __launch_bounds__(64) void k_wf_mds3_jobs(const MdUnit* __restrict__ units, int n,
                                                     const uint8_t* __restrict__ sel, int k1, int k3,
                                                     const WfCand* __restrict__ cands,
                                                     const uint32_t* __restrict__ mode_rate, int nc,
                                                     const uint64_t* __restrict__ dist1,
                                                     const int32_t* __restrict__ rate1, uint64_t lambda,
                                                     const uint8_t* __restrict__ types, int nt,
                                                     const int32_t* __restrict__ txt, int L, int64_t rdmult,
                                                     uint8_t* __restrict__ m3, TxUnitRdJob* __restrict__ out) {
  __shared__ uint8_t ord[64];
  const int u = blockIdx.x, lane = threadIdx.x;
  if (u >= n) return;
  int idx = lane;
  uint64_t cost = ~0ull;
  if (lane < k1) {
    const int c = sel[u * k1 + lane];
    cost = rdc(lambda, (uint64_t)mode_rate[u * nc + c] + (uint32_t)rate1[u * k1 + lane], dist1[u * k1 + lane]);
  }
  wave_exchange_first_k(idx, cost, k1, k3, lane, m3 + u * k3, nullptr);
  ord[lane] = (uint8_t)idx;
  wave_sync();
  const int32_t x = units[u].x, y = units[u].y;
  for (int q = lane; q < k3 * nt; q += 64) {
    const int i = q / nt, t = q - i * nt;
    const int r = ord[i];
    const WfCand c = cands[sel[u * k1 + r]];
    TxUnitRdJob j{};
    j.x = x, j.y = y, j.pred_off = (uint32_t)(u * k1 + r) * (uint32_t)(L * L);
    j.tx_type = types[t], j.crop_w = (uint8_t)L, j.crop_h = (uint8_t)L, j.plane_type = 0;
    j.txt_rate = txt[c.dir * 16 + types[t]];
    j.rdmult = rdmult;
    out[(u * k3 + i) * nt + t] = j;
  }
}

// The leaf's winner: per MDS3 candidate the tx-type search, then the least
// full cost (mode rate + coefficient rate, spatial distortion), first in
// order; its reconstruction is written into the plane, its decision and mode
// recorded.  One workgroup per unit; `pick_in` (host-chosen winner job) is
// used instead of the search when not null.
template <typename Pix>
__global__ __launch_bounds__(256) void k_wf_commit(const MdUnit* __restrict__ units, int n,
                                                   const int32_t* __restrict__ unit_id,
                                                   const uint8_t* __restrict__ m3, const uint8_t* __restrict__ sel,
                                                   int k1, int k3, const WfCand* __restrict__ cands,
                                                   const uint32_t* __restrict__ mode_rate, int nc,
                                                   const uint8_t* __restrict__ types, int nt,
                                                   const int32_t* __restrict__ txt,
                                                   const uint64_t* __restrict__ dist3,
                                                   const int32_t* __restrict__ rate3,
                                                   const int32_t* __restrict__ eob3,
                                                   const uint64_t* __restrict__ csatd3,
                                                   const uint16_t* __restrict__ rec3, uint64_t lambda,
                                                   uint64_t rate_th, int64_t satd_th, Pix* __restrict__ recon,
                                                   int stride, int L, WfDecision* __restrict__ dec,
                                                   uint8_t* __restrict__ mode_map, const int32_t* __restrict__ q3,
                                                   int32_t* __restrict__ levels,
                                                   const uint64_t* __restrict__ level_off) {
  __shared__ int win_job;
  __shared__ uint64_t s_full[64];
  __shared__ int s_t[64];
  const int u = blockIdx.x;
  if (u >= n) return;
  // each survivor's tx-type pick on its own lane (k3 <= 64), then the
  // first-in-order least full cost
  if (threadIdx.x < k3) {
    const int i = threadIdx.x;
    const int r = m3[u * k3 + i];
    const int c = sel[u * k1 + r];
    const long j0 = (long)(u * k3 + i) * nt;
    uint64_t tc;
    const int t = txt_pick(nt, types, txt + cands[c].dir * 16, dist3 + j0, rate3 + j0, eob3 + j0, csatd3 + j0,
                           lambda, rate_th, satd_th, &tc);
    s_full[i] = rdc(lambda, (uint64_t)mode_rate[u * nc + c] + (uint32_t)rate3[j0 + t], dist3[j0 + t]);
    s_t[i] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t best_full = ~0ull;
    int bi = 0, bt = 0;
    uint64_t bcost = 0;
    for (int i = 0; i < k3; i++) {
      const uint64_t full = s_full[i];
      if (full < best_full) {
        best_full = full;
        bi = i;
        bt = s_t[i];
        bcost = full;
      }
    }
    const long jw = (long)(u * k3 + bi) * nt + bt;
    win_job = (int)jw;
    const int r = m3[u * k3 + bi];
    const int c = sel[u * k1 + r];
    WfDecision d;
    d.cand = (uint8_t)c, d.tx_type = types[bt], d.mode = cands[c].mode, d.fi = cands[c].fi;
    d.eob = eob3[jw], d.rate = rate3[jw], d.mode_rate = (int32_t)mode_rate[u * nc + c];
    d.dist = dist3[jw], d.cost = bcost;
    const int id = unit_id[u];
    dec[id] = d;
    mode_map[id] = cands[c].mode;
  }
  __syncthreads();
  const uint16_t* src = rec3 + (long)win_job * L * L;
  const long x0 = units[u].x, y0 = units[u].y;
  for (int i = threadIdx.x; i < L * L; i += 256) {
    const int r = i / L, c = i - r * L;
    recon[(y0 + r) * stride + x0 + c] = (Pix)src[i];
  }
  // the winner's quantized levels (packed min(L, 32)^2, raster), for the
  // entropy coder
  if (levels) {
    const int np = L < 32 ? L * L : 1024;
    const int32_t* qs = q3 + (long)win_job * np;
    int32_t* dst = levels + level_off[unit_id[u]];
    for (int i = threadIdx.x; i < np; i += 256) dst[i] = qs[i];
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host plan
// ---------------------------------------------------------------------------

// Per leaf size: the funnel's candidates, counts and tx types at that size
// and the stage buffers, sized for the size's largest step.
struct SizePlan {
  int L = 0, tx = 0, nt = 0, nc = 0, k1 = 0, k3 = 0, max_u = 0, np = 0;
  uint8_t types[16] = {};
  MdCandList cl{};
  std::vector<WfCand> h_cands;
  int64_t satd_th = 0;
  const uint16_t* fscan = nullptr;
  uint32_t* rate_tab = nullptr;
  int32_t* txt = nullptr;
  WfCand* cands = nullptr;
  uint8_t* d_types = nullptr;
  uint32_t *rate = nullptr, *satd = nullptr;
  uint8_t *sel = nullptr, *m3 = nullptr;
  uint64_t* sel_cost = nullptr;
  TxUnitJob* fs_jobs = nullptr;
  uint16_t *pred = nullptr, *rec3 = nullptr;
  TxUnitRdJob *jobs1 = nullptr, *jobs3 = nullptr;
  uint64_t *dist1 = nullptr, *dist3 = nullptr, *csatd3 = nullptr;
  int32_t *eob1 = nullptr, *rate1 = nullptr, *eob3 = nullptr, *rate3 = nullptr;
  int32_t* q3 = nullptr;  // the plain form's MDS3 levels of one leaf (kept levels only)
  uint8_t *cul1 = nullptr, *cul3 = nullptr;
  void *ws1 = nullptr, *ws3 = nullptr;
  size_t ws3_bytes = 0;  // MDS3's workspace, with room for the trellis's lane form
  int rdoq_form = -1;    // the trellis form (rdoq_batch's form; -1 its defaults), set by the autotuner
};

struct Wavefront {
  zav1_wf_cfg cfg{};
  int stacked_h = 0, n_units = 0, n_steps = 0, max_u = 0, n_layers = 0;
  size_t bytes = 0;
  std::vector<int> step_off;      // per step: first unit (n_steps + 1 entries)
  std::vector<uint8_t> step_sz;   // per step: its leaf size class (8, 16, 32, 64 -> 0..3)
  std::vector<int> step_layer;    // per step: its dependency layer (a layer's steps are independent)
  // a layer's other leaf sizes run on these streams beside the caller's
  // (plans with more than one step in a layer), forked and joined by events
  hipStream_t side[3] = {nullptr, nullptr, nullptr};
  hipEvent_t fork = nullptr, join[3] = {nullptr, nullptr, nullptr};
  std::vector<uint8_t> h_sz;      // per unit: size class
  std::vector<MdUnit> h_units;
  std::vector<int32_t> h_nbr;
  std::vector<uint8_t> leaf_map;  // the caller's partition map (empty: uniform leaves)
  std::vector<uint64_t> h_loff;   // per unit: offset of its kept levels
  std::vector<uint64_t> img_loff; // per image: first kept level (n_images + 1 entries)
  size_t n_levels = 0;
  SizePlan sp[4];
  uint64_t rate_th = 0;
  int64_t rdmult = 0;
  unsigned flags1 = 0, flags3 = 0;
  uint32_t cut_num = 0, cut_den = 0;
  int32_t dq0 = 0, dq1 = 0;
  const QuantRow* row = nullptr;
  const uint16_t *iscan = nullptr, *fscan_base = nullptr;
  // device memory (one allocation, carved)
  char* base = nullptr;
  void *src = nullptr, *recon = nullptr;
  MdUnit* units = nullptr;
  int32_t *nbr = nullptr, *uid = nullptr;
  WfDecision* dec = nullptr;
  uint8_t* mode_map = nullptr;
  int32_t* costs = nullptr;
  int32_t* levels = nullptr;  // kept levels (cfg.keep_levels), per unit at h_loff
  unsigned long long* trellis_runs = nullptr;  // MDS3 jobs handed to the trellis, over every run
  uint64_t* loff = nullptr;
  hipGraphExec_t graph = nullptr;
  hipStream_t graph_stream = nullptr;
  // pinned host staging of the decisions and kept levels (the batch
  // writer's read-back), allocated on its first call
  WfDecision* h_dec_pin = nullptr;
  int32_t* h_lev_pin = nullptr;
};

namespace {

int zorder(int i, int j) {  // bit interleave: row bits odd, column bits even
  int z = 0;
  for (int b = 0; b < 8; b++) z |= (((j >> b) & 1) << (2 * b)) | (((i >> b) & 1) << (2 * b + 1));
  return z;
}

size_t al(size_t v) { return (v + 255) & ~(size_t)255; }
int size_class(int L) { return L == 8 ? 0 : (L == 16 ? 1 : (L == 32 ? 2 : 3)); }

}  // namespace

// The intra ext-tx set of the leaf's transform in TX_TYPE_GROUPS order
// (leaf_funnel/mod.rs:7189: {DCT}, {V_DCT, H_DCT}, {ADST_ADST},
// {ADST_DCT, DCT_ADST}, {FLIPADST_FLIPADST, IDTX}, the rest), limited to the
// first `groups` groups.
static int search_types(int tx_size, int groups, uint8_t* out) {
  static const int8_t kGroups[6][8] = {{0, -1}, {10, 11, -1}, {3, -1}, {1, 2, -1}, {6, 9, -1},
                                       {4, 5, 7, 8, 12, 13, 14, 15}};
  static const int8_t kGroupLen[6] = {1, 2, 1, 2, 2, 8};
  const int w = tx_w(tx_size), h = tx_h(tx_size);
  const bool dct_only = std::max(w, h) >= 32;
  const bool sq16 = w == 16 && h == 16;
  auto used = [&](int t) {
    if (t == 0) return true;
    if (dct_only) return false;
    if (t == 1 || t == 2 || t == 3 || t == 9) return true;
    return !sq16 && (t == 10 || t == 11);
  };
  int n = 0;
  for (int g = 0; g < std::min(groups, 6); g++)
    for (int k = 0; k < kGroupLen[g]; k++)
      if (used(kGroups[g][k])) out[n++] = (uint8_t)kGroups[g][k];
  return n;
}

// The funnel at one leaf size (host part): the configuration's candidates
// without filter intra on 64x64 leaves (filter-intra candidates come last),
// the survivor counts capped by them, the size's tx types.
static void size_setup(const zav1_wf_cfg& c, bool tree, int L, SizePlan* p) {
  static const uint8_t kFiDir[5] = {0, 1, 2, 6, 0};  // FIMODE_TO_INTRADIR
  p->L = L;
  p->np = L < 32 ? L * L : 1024;
  for (int s = 0; s < 19; s++)
    if (tx_w(s) == L && tx_h(s) == L) p->tx = s;
  int nc = c.n_cands;
  while (L > 32 && nc > 0 && c.cands[nc - 1].fi_mode < 5) nc--;
  p->nc = nc;
  p->k1 = std::min(c.k1, nc);
  p->k3 = std::min(c.k3, p->k1);
  p->nt = search_types(p->tx, (tree && L < 16) ? c.txt_groups_lt16 : c.txt_groups, p->types);
  p->satd_th = p->nt > 1 ? c.txt_satd_th : 0;
  p->cl.n = nc;
  p->h_cands.clear();
  for (int i = 0; i < nc; i++) {
    const zav1_md_cand m = c.cands[i];
    p->cl.c[i] = MdCand{m.mode, m.delta, m.fi_mode, 0};
    p->h_cands.push_back(WfCand{m.mode, m.delta, m.fi_mode, (uint8_t)(m.fi_mode < 5 ? kFiDir[m.fi_mode] : m.mode)});
  }
}

// Leaves in the reference's coding order (tiles in order, superblocks in
// raster order within a tile, leaves in Z-order within a superblock: of a
// uniform partition, or of the partition map's 64x64 units in Z-order), each
// with its AV1 neighbour availability; then the steps.  A leaf's step is one
// past the latest step of the leaves whose reconstruction or mode it reads
// (the row above from the top-left sample through the top-right extent, the
// column left through the bottom-left extent), so every step's leaves are
// independent and each leaf's inputs are final when it runs: the decisions
// equal the sequential order's.  A step is split by leaf size, one chain of
// launches per (step, size).
static int plan(Wavefront* w) {
  const zav1_wf_cfg& c = w->cfg;
  const bool tree = !w->leaf_map.empty();
  w->stacked_h = c.height * c.n_images;
  const int sbc = (c.width + c.sb - 1) / c.sb, sbr = (c.height + c.sb - 1) / c.sb;
  const int tcols = std::max(1, c.tile_cols), trows = std::max(1, c.tile_rows);
  // uniform tiles in superblock units (the last ones take the remainder)
  auto tile_edges = [](int n_sb, int n_t, std::vector<int>& e) {
    e.assign(n_t + 1, 0);
    const int per = (n_sb + n_t - 1) / n_t;
    for (int t = 0; t <= n_t; t++) e[t] = std::min(n_sb, t * per);
  };
  std::vector<int> tc, tr;
  tile_edges(sbc, tcols, tc);
  tile_edges(sbr, trows, tr);
  struct Leaf {
    int img, x, y, L, tx0, tx1, ty0, ty1, layer, ntr, nbl;
  };
  std::vector<Leaf> leaves;
  std::vector<int> img_k0(c.n_images + 1, 0);  // each image's leaves are one range
  const int uw = c.width / 64, uh = c.height / 64;
  for (int img = 0; img < c.n_images; img++) {
    img_k0[img] = (int)leaves.size();
    for (int ty = 0; ty < trows; ty++)
      for (int tx = 0; tx < tcols; tx++) {
        const int x0 = tc[tx] * c.sb, x1 = std::min(c.width, tc[tx + 1] * c.sb);
        const int y0 = tr[ty] * c.sb, y1 = std::min(c.height, tr[ty + 1] * c.sb);
        for (int r = tr[ty]; r < tr[ty + 1]; r++)
          for (int q = tc[tx]; q < tc[tx + 1]; q++) {
            if (tree) {
              const int nu = (c.sb / 64) * (c.sb / 64);
              for (int u = 0; u < nu; u++) {
                const int ux = q * c.sb + (u & 1) * 64, uy = r * c.sb + (u >> 1) * 64;
                if (ux >= c.width || uy >= c.height) continue;
                const uint8_t* m = w->leaf_map.data() + ((size_t)img * uw * uh + (size_t)(uy / 64) * uw + ux / 64) * 256;
                // Z-order DFS of the unit's quadtree
                struct Node {
                  int sq, ox, oy;
                };
                Node stack[64];
                int sp = 0;
                stack[sp++] = {64, 0, 0};
                while (sp) {
                  const Node nd = stack[--sp];
                  if ((1 << m[(nd.oy / 4) * 16 + nd.ox / 4]) == nd.sq) {
                    leaves.push_back({img, ux + nd.ox, uy + nd.oy, nd.sq, x0, x1, y0, y1, 0, 0, 0});
                    continue;
                  }
                  const int h = nd.sq / 2;
                  for (int i = 3; i >= 0; i--) stack[sp++] = {h, nd.ox + (i & 1) * h, nd.oy + (i >> 1) * h};
                }
              }
            } else {
              const int L = c.leaf, nl = c.sb / L;
              for (int z = 0; z < nl * nl; z++) {
                int i = 0, j = 0;
                for (int b = 0; b < 8; b++) {
                  j |= ((z >> (2 * b)) & 1) << b;
                  i |= ((z >> (2 * b + 1)) & 1) << b;
                }
                const int x = q * c.sb + j * L, y = r * c.sb + i * L;
                if (x >= c.width || y >= c.height) continue;
                leaves.push_back({img, x, y, L, x0, x1, y0, y1, 0, 0, 0});
              }
            }
          }
      }
  }
  const int n = (int)leaves.size();
  img_k0[c.n_images] = n;
  w->n_units = n;
  // 8x8 cells -> leaf (coding-order index), per image
  const int gw = c.width / 8, gh = c.height / 8;
  std::vector<int32_t> cell((size_t)c.n_images * gw * gh, -1);
  std::vector<MdUnit> units(n);
  std::vector<int32_t> nbr(2 * (size_t)n);
  // images are independent: their leaves' neighbours and layers are found on
  // host threads, one image range each
  const int n_th = std::max(1, std::min({c.n_images, 16, (int)std::max(1u, std::thread::hardware_concurrency())}));
  std::vector<int> th_layers(n_th, 0);
  auto layer_images = [&](int th) {
    int& n_layers = th_layers[th];
    for (int k = img_k0[th * c.n_images / n_th]; k < img_k0[(th + 1) * c.n_images / n_th]; k++) {
      Leaf& f = leaves[k];
      const int L = f.L, nl = c.sb / L;
      const int i = (f.y % c.sb) / L, j = (f.x % c.sb) / L, zi = zorder(i, j);
      const bool up = f.y > f.ty0, lf = f.x > f.tx0;
      // top-right: the row above; inside the superblock only when that square
      // precedes this one in Z-order, never in the superblock to the right
      bool tr_ok;
      if (!up) tr_ok = false;
      else if (i == 0) tr_ok = true;
      else tr_ok = j + 1 < nl && zorder(i - 1, j + 1) < zi;
      // bottom-left: the left superblock's rows of this superblock row, or an
      // earlier square of this superblock
      bool bl_ok;
      if (!lf) bl_ok = false;
      else if (j == 0) bl_ok = i + 1 < nl;
      else bl_ok = i + 1 < nl && zorder(i + 1, j - 1) < zi;
      const int ntr = tr_ok ? std::max(0, std::min(L, f.tx1 - (f.x + L))) : 0;
      const int nbl = bl_ok ? std::max(0, std::min(L, f.ty1 - (f.y + L))) : 0;
      f.ntr = ntr, f.nbl = nbl;
      MdUnit mu{};
      mu.x = f.x;
      mu.y = f.y + f.img * c.height;
      mu.n_top_px = (int16_t)(up ? L : 0);
      mu.n_topright_px = (int16_t)ntr;
      mu.n_left_px = (int16_t)(lf ? L : 0);
      mu.n_bottomleft_px = (int16_t)nbl;
      mu.edge_filter = 1;
      units[k] = mu;
      const size_t g = (size_t)f.img * gw * gh;
      auto at = [&](int x, int y) { return cell[g + (size_t)(y / 8) * gw + x / 8]; };
      nbr[2 * k] = up ? at(f.x, f.y - 1) : -1;
      nbr[2 * k + 1] = lf ? at(f.x - 1, f.y) : -1;
      int layer = 0;
      auto dep = [&](int x, int y) {
        const int id = at(x, y);
        if (id >= 0) layer = std::max(layer, leaves[id].layer + 1);
      };
      if (up)
        for (int x = lf ? f.x - 1 : f.x; x < f.x + L + ntr; x += 8) dep(x, f.y - 1);
      if (lf)
        for (int y = f.y; y < f.y + L + nbl; y += 8) dep(f.x - 1, y);
      f.layer = layer;
      n_layers = std::max(n_layers, layer + 1);
      for (int y = f.y; y < f.y + L; y += 8)
        for (int x = f.x; x < f.x + L; x += 8) cell[g + (size_t)(y / 8) * gw + x / 8] = k;
    }
  };
  if (n_th == 1) {
    layer_images(0);
  } else {
    std::vector<std::thread> pool;
    for (int t = 0; t < n_th; t++) pool.emplace_back(layer_images, t);
    for (auto& t : pool) t.join();
  }
  int n_layers = *std::max_element(th_layers.begin(), th_layers.end());
  if (tree) {
    // the layers above are as soon as possible; a step is one (layer, leaf
    // size) pair, so leaves of a size with slack are moved to a layer that
    // already holds a step of their size.  Each leaf's latest layer (its
    // dependents' latest less one) bounds the move, so no dependency breaks
    // and the layer count does not grow; any order that keeps the
    // dependencies gives the same decisions as coding order.
    auto for_deps = [&](const Leaf& f, auto&& fn) {
      const size_t g = (size_t)f.img * gw * gh;
      auto at = [&](int x, int y) { return cell[g + (size_t)(y / 8) * gw + x / 8]; };
      const bool up = f.y > f.ty0, lf = f.x > f.tx0;
      if (up)
        for (int x = lf ? f.x - 1 : f.x; x < f.x + f.L + f.ntr; x += 8) fn(at(x, f.y - 1));
      if (lf)
        for (int y = f.y; y < f.y + f.L + f.nbl; y += 8) fn(at(f.x - 1, y));
    };
    std::vector<int> late(n, n_layers - 1);
    for (int k = n - 1; k >= 0; k--)
      for_deps(leaves[k], [&](int id) {
        if (id >= 0) late[id] = std::min(late[id], late[k] - 1);
      });
    std::set<int> open[4];  // per size class: the layers holding a step of it
    n_layers = 0;
    for (int k = 0; k < n; k++) {
      Leaf& f = leaves[k];
      int lo = 0;
      for_deps(f, [&](int id) {
        if (id >= 0) lo = std::max(lo, leaves[id].layer + 1);
      });
      const int sc = size_class(f.L), hi = late[k];
      const auto it = open[sc].lower_bound(lo);
      int at;
      if (it != open[sc].end() && *it <= hi) {
        at = *it;
      } else {
        at = hi;
        open[sc].insert(at);
      }
      f.layer = at;
      n_layers = std::max(n_layers, at + 1);
    }
  }
  w->n_layers = n_layers;
  // storage order: by (step, leaf size), coding order within -- a counting
  // sort over the keys layer x 4 + size class (stable)
  std::vector<int> perm(n), pos(n);
  {
    std::vector<int> start((size_t)n_layers * 4 + 1, 0);
    for (int k = 0; k < n; k++) start[(size_t)leaves[k].layer * 4 + size_class(leaves[k].L) + 1]++;
    for (size_t b = 1; b < start.size(); b++) start[b] += start[b - 1];
    for (int k = 0; k < n; k++) perm[start[(size_t)leaves[k].layer * 4 + size_class(leaves[k].L)]++] = k;
  }
  for (int s = 0; s < n; s++) pos[perm[s]] = s;
  w->h_units.resize(n);
  w->h_nbr.resize(2 * (size_t)n);
  w->h_sz.resize(n);
  w->step_off.clear();
  w->step_sz.clear();
  w->step_layer.clear();
  for (int k = 0; k < 4; k++) w->sp[k].max_u = 0;
  for (int s = 0; s < n; s++) {
    const int k = perm[s];
    const Leaf& f = leaves[k];
    w->h_units[s] = units[k];
    w->h_nbr[2 * s] = nbr[2 * k] >= 0 ? pos[nbr[2 * k]] : -1;
    w->h_nbr[2 * s + 1] = nbr[2 * k + 1] >= 0 ? pos[nbr[2 * k + 1]] : -1;
    const int sc = size_class(f.L);
    w->h_sz[s] = (uint8_t)sc;
    if (s == 0 || f.layer != leaves[perm[s - 1]].layer || sc != w->h_sz[s - 1]) {
      w->step_off.push_back(s);
      w->step_sz.push_back((uint8_t)sc);
      w->step_layer.push_back(f.layer);
    }
  }
  w->step_off.push_back(n);
  w->n_steps = (int)w->step_sz.size();
  // kept levels image-major (an image's levels are one range), storage
  // order within an image
  w->h_loff.assign(n, 0);
  w->n_levels = 0;
  w->img_loff.assign(c.n_images + 1, 0);
  if (c.keep_levels) {
    for (int s = 0; s < n; s++) {
      const int L = 8 << w->h_sz[s];
      w->img_loff[w->h_units[s].y / c.height + 1] += L < 32 ? (size_t)L * L : 1024;
    }
    for (int i = 0; i < c.n_images; i++) w->img_loff[i + 1] += w->img_loff[i];
    std::vector<uint64_t> next(w->img_loff.begin(), w->img_loff.end() - 1);
    for (int s = 0; s < n; s++) {
      const int L = 8 << w->h_sz[s];
      uint64_t& o = next[w->h_units[s].y / c.height];
      w->h_loff[s] = o;
      o += L < 32 ? (uint64_t)L * L : 1024;
    }
    w->n_levels = w->img_loff[c.n_images];
  }
  w->max_u = 0;
  for (int s = 0; s < w->n_steps; s++) {
    const int u = w->step_off[s + 1] - w->step_off[s];
    SizePlan& p = w->sp[w->step_sz[s]];
    p.max_u = std::max(p.max_u, u);
    w->max_u = std::max(w->max_u, u);
  }
  for (int k = 0; k < 4; k++)
    if (w->sp[k].max_u) size_setup(c, tree, 8 << k, &w->sp[k]);
  return ZAV1_OK;
}

// MDS3's transform-unit workspace plus the trellis's lane-form workspace,
// so that a batch large enough for the lane form (rdoq_lane_min) can take it
static size_t ws3_bytes(int tx, int n3) {
  return ((txunit_rd_workspace(tx, n3) + 255) & ~(size_t)255) + rdoq_workspace(tx, n3);
}

static size_t size_bytes(const SizePlan& p, bool keep) {
  if (!p.max_u) return 0;
  const size_t U = (size_t)p.max_u, LL = (size_t)p.L * p.L, n3 = U * p.k3 * p.nt;
  size_t b = al(25 * p.nc * 4) + al(13 * 16 * 4) + al(64 * sizeof(WfCand)) + al(16);
  b += 2 * al(U * p.nc * 4) + 2 * al(U * p.k1) + al(U * p.k1 * 8) + al(U * p.k1 * sizeof(TxUnitJob));
  b += al(U * p.k1 * LL * 2) + al(U * p.k1 * sizeof(TxUnitRdJob)) + al(U * p.k1 * 8) + 2 * al(U * p.k1 * 4) +
       al(U * p.k1);
  b += al(n3 * sizeof(TxUnitRdJob)) + 2 * al(n3 * 8) + 2 * al(n3 * 4) + al(n3) + al(n3 * LL * 2);
  b += al(txunit_rd_workspace(p.tx, (int)(U * p.k1))) + al(ws3_bytes(p.tx, (int)n3)) + al(64);
  if (keep) b += al((size_t)p.k3 * p.nt * p.np * 4);
  return b;
}

static size_t footprint(const Wavefront& w) {
  const zav1_wf_cfg& c = w.cfg;
  const size_t es = c.bd > 8 ? 2 : 1, plane = (size_t)c.width * w.stacked_h * es;
  const size_t n = (size_t)w.n_units;
  size_t b = 2 * al(plane) + al(n * sizeof(MdUnit)) + al(n * 8) + al(n * 4) + al(n * sizeof(WfDecision)) + al(n);
  b += al(sizeof(zav1_coeff_costs)) + al(8);
  if (c.keep_levels) b += al(w.n_levels * 4) + al(n * 8);
  for (int k = 0; k < 4; k++) b += size_bytes(w.sp[k], c.keep_levels != 0);
  return b;
}

constexpr int kWfStages = ZAV1_WF_STAGES;
// Events at the stage boundaries of a step (the timed run): ev[k] opens
// stage k, ev[kWfStages] closes the last.
struct StageClock {
  hipEvent_t ev[kWfStages + 1];
};
#define MARK(k)                                                              \
  do {                                                                       \
    if (clk) {                                                               \
      const hipError_t em = hipEventRecord(clk->ev[k], st);                  \
      if (em != hipSuccess) return em;                                       \
    }                                                                        \
  } while (0)

static hipError_t step_chain(Wavefront* w, int s, hipStream_t st, StageClock* clk = nullptr) {
  const int u0 = w->step_off[s], U = w->step_off[s + 1] - u0;
  if (U == 0) return hipSuccess;
  const zav1_wf_cfg& c = w->cfg;
  SizePlan& P = w->sp[w->step_sz[s]];
  const int L = P.L, W = c.width, H = w->stacked_h;
  MdUnit* units = w->units + u0;
  MARK(0);
  hipLaunchKernelGGL(k_wf_prep, dim3((unsigned)((U * P.nc + 255) / 256)), dim3(256), 0, st, units, w->nbr + 2 * u0,
                     U, w->mode_map, P.rate_tab, P.nc, P.rate);
  MARK(1);
  hipError_t e = md_fast_satd(w->src, w->recon, W, W, H, c.bd, L, L, units, U, P.cl, P.satd, st);
  if (e != hipSuccess) return e;
  MARK(2);
  TypeList dct{};
  dct.n = 1;
  e = fast_select(P.satd, P.rate, c.lambda, U, P.nc, P.k1, reinterpret_cast<const int32_t*>(units), 5, L, L, dct, 1,
                  P.sel, P.sel_cost, P.fs_jobs, st);
  if (e != hipSuccess) return e;
  MARK(3);
  e = md_predict_survivors(w->recon, W, W, H, c.bd, L, L, units, U, P.cl, P.sel, P.k1, P.pred, st);
  if (e != hipSuccess) return e;
  MARK(4);
  hipLaunchKernelGGL(k_wf_mds1_jobs, dim3((unsigned)((U * P.k1 + 255) / 256)), dim3(256), 0, st, units, U, P.sel,
                     P.k1, P.cands, P.txt, L, w->rdmult, P.jobs1);
  MARK(5);
  e = txunit_rd_batch(w->src, W, W, H, c.bd, P.tx, P.pred, (long)U * P.k1 * L * L, P.jobs1, U * P.k1, w->row,
                      w->iscan, P.fscan, w->costs, w->flags1, 0, 0, P.ws1, P.dist1, P.eob1, P.rate1, P.cul1, nullptr,
                      nullptr, st, w->dq0, w->dq1, nullptr);
  if (e != hipSuccess) return e;
  MARK(6);
  hipLaunchKernelGGL(k_wf_mds3_jobs, dim3((unsigned)U), dim3(64), 0, st, units, U, P.sel, P.k1, P.k3, P.cands, P.rate,
                     P.nc, P.dist1, P.rate1, c.lambda, P.d_types, P.nt, P.txt, L, w->rdmult, P.m3, P.jobs3);
  const int n3 = U * P.k3 * P.nt;
  MARK(7);
  // types the commit's coefficient-SATD gate skips whatever the rate gate
  // does (more than satd_th percent above the DCT's: the running minimum it
  // compares against is at most the DCT's) skip the trellis and the rest
  const TuSatdGate gate{P.nt, P.satd_th, w->trellis_runs};
  e = txunit_rd_batch(w->src, W, W, H, c.bd, P.tx, P.pred, (long)U * P.k1 * L * L, P.jobs3, n3, w->row, w->iscan,
                      P.fscan, w->costs, w->flags3, w->cut_num, w->cut_den, P.ws3, P.dist3, P.eob3, P.rate3, P.cul3,
                      nullptr, P.rec3, st, w->dq0, w->dq1, P.csatd3, clk ? &clk->ev[8] : nullptr, P.ws3_bytes, &gate,
                      P.rdoq_form);
  if (e != hipSuccess) return e;
  // the commit copies the winners' levels straight from the batch's workspace
  const int32_t* q3 = c.keep_levels ? txunit_rd_levels(P.ws3) : nullptr;
  MARK(10);
  if (c.bd == 8)
    hipLaunchKernelGGL((k_wf_commit<uint8_t>), dim3((unsigned)U), dim3(256), 0, st, units, U, w->uid + u0, P.m3, P.sel,
                       P.k1, P.k3, P.cands, P.rate, P.nc, P.d_types, P.nt, P.txt, P.dist3, P.rate3, P.eob3, P.csatd3,
                       P.rec3, c.lambda, w->rate_th, P.satd_th, (uint8_t*)w->recon, W, L, w->dec, w->mode_map, q3,
                       w->levels, w->loff);
  else
    hipLaunchKernelGGL((k_wf_commit<uint16_t>), dim3((unsigned)U), dim3(256), 0, st, units, U, w->uid + u0, P.m3,
                       P.sel, P.k1, P.k3, P.cands, P.rate, P.nc, P.d_types, P.nt, P.txt, P.dist3, P.rate3, P.eob3,
                       P.csatd3, P.rec3, c.lambda, w->rate_th, P.satd_th, (uint16_t*)w->recon, W, L, w->dec,
                       w->mode_map, q3, w->levels, w->loff);
  MARK(kWfStages);
  return hipGetLastError();
}
#undef MARK

// Steps [s0, s1) in order on `st`; the steps of one layer (one per leaf
// size present) are independent, so when the plan has side streams the
// layer's first step runs on `st` and the others on side streams, forked
// from and joined back into `st` by events (captured as parallel branches
// when `st` is capturing).
static hipError_t run_range(Wavefront* w, int s0, int s1, hipStream_t st) {
  for (int s = s0; s < s1;) {
    int e = s + 1;
    while (e < s1 && w->step_layer[e] == w->step_layer[s]) e++;
    if (e - s == 1 || !w->side[0]) {
      for (; s < e; s++) {
        const hipError_t r = step_chain(w, s, st);
        if (r != hipSuccess) return r;
      }
      continue;
    }
    hipError_t r = hipEventRecord(w->fork, st);
    for (int k = 1; k < e - s && r == hipSuccess; k++) {
      r = hipStreamWaitEvent(w->side[k - 1], w->fork, 0);
      if (r == hipSuccess) r = step_chain(w, s + k, w->side[k - 1]);
      if (r == hipSuccess) r = hipEventRecord(w->join[k - 1], w->side[k - 1]);
    }
    if (r == hipSuccess) r = step_chain(w, s, st);
    for (int k = 1; k < e - s && r == hipSuccess; k++) r = hipStreamWaitEvent(st, w->join[k - 1], 0);
    if (r != hipSuccess) return r;
    s = e;
  }
  return hipSuccess;
}

// The plain form of one unit: one launch per elementary call with one job
// each (every candidate's prediction + SATD, every survivor's prediction and
// transform unit, every MDS3 type's unit), the selections on the host from
// the results read back after each stage -- what a per-call binding of the
// same kernels does.
static hipError_t plain_unit(Wavefront* w, int u, hipStream_t st) {
  const zav1_wf_cfg& c = w->cfg;
  SizePlan& P = w->sp[w->h_sz[u]];
  const int L = P.L, W = c.width, H = w->stacked_h, nc = P.nc, k1 = P.k1, k3 = P.k3, nt = P.nt;
  MdUnit* unit = w->units + u;
  hipLaunchKernelGGL(k_wf_prep, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st, unit, w->nbr + 2 * u, 1,
                     w->mode_map, P.rate_tab, nc, P.rate);
  hipError_t e;
  for (int m = 0; m < nc; m++) {
    MdCandList one{};
    one.n = 1;
    one.c[0] = P.cl.c[m];
    e = md_fast_satd(w->src, w->recon, W, W, H, c.bd, L, L, unit, 1, one, P.satd + m, st);
    if (e != hipSuccess) return e;
  }
  std::vector<uint32_t> satd(nc), rate(nc);
  if ((e = hipMemcpyAsync(satd.data(), P.satd, 4 * nc, hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(rate.data(), P.rate, 4 * nc, hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
  std::vector<uint64_t> fc(nc);
  for (int m = 0; m < nc; m++)
    fc[m] = satd[m] == kSatdInvalid ? ~0ull : rdc(c.lambda, rate[m], (uint64_t)satd[m] << 4);
  uint8_t ord[64];
  exchange_first_k(fc.data(), nc, k1, ord);
  std::vector<uint8_t> sel(ord, ord + k1);
  if ((e = hipMemcpyAsync(P.sel, sel.data(), k1, hipMemcpyHostToDevice, st)) != hipSuccess) return e;
  MdUnit hu = w->h_units[u];
  std::vector<uint64_t> d1(k1);
  std::vector<int32_t> r1(k1);
  for (int r = 0; r < k1; r++) {
    e = md_predict_survivors(w->recon, W, W, H, c.bd, L, L, unit, 1, P.cl, P.sel + r, 1, P.pred + r * L * L, st);
    if (e != hipSuccess) return e;
  }
  // MDS1 jobs on the device as the batched chain builds them, then one
  // transform-unit call per survivor
  hipLaunchKernelGGL(k_wf_mds1_jobs, dim3(1), dim3(256), 0, st, unit, 1, P.sel, k1, P.cands, P.txt, L, w->rdmult,
                     P.jobs1);
  for (int r = 0; r < k1; r++) {
    e = txunit_rd_batch(w->src, W, W, H, c.bd, P.tx, P.pred, (long)k1 * L * L, P.jobs1 + r, 1, w->row, w->iscan,
                        P.fscan, w->costs, w->flags1, 0, 0, P.ws1, P.dist1 + r, P.eob1 + r, P.rate1 + r,
                        P.cul1 + r, nullptr, nullptr, st, w->dq0, w->dq1, nullptr);
    if (e != hipSuccess) return e;
  }
  if ((e = hipMemcpyAsync(d1.data(), P.dist1, 8 * k1, hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(r1.data(), P.rate1, 4 * k1, hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
  std::vector<uint64_t> c1(k1);
  for (int r = 0; r < k1; r++) c1[r] = rdc(c.lambda, (uint64_t)rate[sel[r]] + (uint32_t)r1[r], d1[r]);
  uint8_t ord3[64];
  exchange_first_k(c1.data(), k1, k3, ord3);
  std::vector<TxUnitRdJob> j3((size_t)k3 * nt);
  std::vector<int32_t> htxt(13 * 16);
  if ((e = hipMemcpyAsync(htxt.data(), P.txt, 4 * 13 * 16, hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
  for (int i = 0; i < k3; i++) {
    const WfCand cd = P.h_cands[sel[ord3[i]]];
    for (int t = 0; t < nt; t++) {
      TxUnitRdJob& j = j3[(size_t)i * nt + t];
      j = TxUnitRdJob{};
      j.x = hu.x, j.y = hu.y, j.pred_off = (uint32_t)(ord3[i] * L * L), j.tx_type = P.types[t];
      j.crop_w = (uint8_t)L, j.crop_h = (uint8_t)L, j.txt_rate = htxt[cd.dir * 16 + P.types[t]], j.rdmult = w->rdmult;
    }
  }
  if ((e = hipMemcpyAsync(P.jobs3, j3.data(), j3.size() * sizeof(TxUnitRdJob), hipMemcpyHostToDevice, st)) !=
      hipSuccess)
    return e;
  for (int q = 0; q < k3 * nt; q++) {
    e = txunit_rd_batch(w->src, W, W, H, c.bd, P.tx, P.pred, (long)k1 * L * L, P.jobs3 + q, 1, w->row, w->iscan,
                        P.fscan, w->costs, w->flags3, w->cut_num, w->cut_den, P.ws3, P.dist3 + q, P.eob3 + q,
                        P.rate3 + q, P.cul3 + q, P.q3 ? P.q3 + (size_t)q * P.np : nullptr, P.rec3 + (size_t)q * L * L,
                        st, w->dq0, w->dq1,
                        P.csatd3 + q);
    if (e != hipSuccess) return e;
  }
  // the winner from the results, then one call storing its reconstruction
  std::vector<uint64_t> d3((size_t)k3 * nt), s3((size_t)k3 * nt);
  std::vector<int32_t> r3((size_t)k3 * nt), e3((size_t)k3 * nt);
  if ((e = hipMemcpyAsync(d3.data(), P.dist3, 8 * d3.size(), hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(s3.data(), P.csatd3, 8 * s3.size(), hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(r3.data(), P.rate3, 4 * r3.size(), hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(e3.data(), P.eob3, 4 * e3.size(), hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
  uint64_t best_full = ~0ull;
  int bi = 0;
  for (int i = 0; i < k3; i++) {
    const int cand = sel[ord3[i]];
    uint64_t tc;
    const size_t j0 = (size_t)i * nt;
    const int t = txt_pick(nt, P.types, htxt.data() + P.h_cands[cand].dir * 16, d3.data() + j0, r3.data() + j0,
                           e3.data() + j0, s3.data() + j0, c.lambda, w->rate_th, P.satd_th, &tc);
    const uint64_t full = rdc(c.lambda, (uint64_t)rate[cand] + (uint32_t)r3[j0 + t], d3[j0 + t]);
    if (full < best_full) {
      best_full = full;
      bi = i;
    }
  }
  // the store: the commit kernel restricted to the chosen candidate (its
  // search over that candidate's types repeats the host's)
  uint8_t m3one = ord3[bi];
  if ((e = hipMemcpyAsync(P.m3, &m3one, 1, hipMemcpyHostToDevice, st)) != hipSuccess) return e;
  const size_t j0 = (size_t)bi * nt;
  if (c.bd == 8)
    hipLaunchKernelGGL((k_wf_commit<uint8_t>), dim3(1), dim3(256), 0, st, unit, 1, w->uid + u, P.m3, P.sel, k1, 1,
                       P.cands, P.rate, nc, P.d_types, nt, P.txt, P.dist3 + j0, P.rate3 + j0, P.eob3 + j0,
                       P.csatd3 + j0, P.rec3 + j0 * L * L, c.lambda, w->rate_th, P.satd_th, (uint8_t*)w->recon, W,
                       L, w->dec, w->mode_map, P.q3 ? P.q3 + j0 * P.np : nullptr, w->levels, w->loff);
  else
    hipLaunchKernelGGL((k_wf_commit<uint16_t>), dim3(1), dim3(256), 0, st, unit, 1, w->uid + u, P.m3, P.sel, k1, 1,
                       P.cands, P.rate, nc, P.d_types, nt, P.txt, P.dist3 + j0, P.rate3 + j0, P.eob3 + j0,
                       P.csatd3 + j0, P.rec3 + j0 * L * L, c.lambda, w->rate_th, P.satd_th, (uint16_t*)w->recon,
                       W, L, w->dec, w->mode_map, P.q3 ? P.q3 + j0 * P.np : nullptr, w->levels, w->loff);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return hipStreamSynchronize(st);
}

}  // namespace zav1

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
namespace {
thread_local std::string g_wf_err;
int wf_fail(hipError_t e, const char* where) {
  g_wf_err = std::string(where) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? ZAV1_ENOMEM : ZAV1_EIO;
}
#define WCHECK(call)                                   \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return wf_fail(e_, #call);   \
  } while (0)

// A partition map: per 64x64 unit 16 x 16 cells of log2 leaf sizes forming
// a quadtree of squares 8..64 (ENOTSUP for 4x4 leaves, EINVAL otherwise).
int check_map(const uint8_t* m) {
  struct Node {
    int sq, ox, oy;
  };
  Node st[64];
  int sp = 0;
  st[sp++] = {64, 0, 0};
  while (sp) {
    const Node nd = st[--sp];
    const int l = m[(nd.oy / 4) * 16 + nd.ox / 4];
    if (l < 2 || l > 6 || (1 << l) > nd.sq) return ZAV1_EINVAL;
    if ((1 << l) == nd.sq) {
      if (l < 3) return ZAV1_ENOTSUP;
      for (int r = nd.oy / 4; r < (nd.oy + nd.sq) / 4; r++)
        for (int k = nd.ox / 4; k < (nd.ox + nd.sq) / 4; k++)
          if (m[r * 16 + k] != l) return ZAV1_EINVAL;
      continue;
    }
    const int h = nd.sq / 2;
    for (int i = 3; i >= 0; i--) st[sp++] = {h, nd.ox + (i & 1) * h, nd.oy + (i >> 1) * h};
  }
  return ZAV1_OK;
}

int check_cfg(const zav1_wf_cfg* c) {
  if (!c) return ZAV1_EINVAL;
  const bool tree = c->leaf_map != nullptr;
  if ((c->bd != 8 && c->bd != 10) || (c->sb != 64 && c->sb != 128) || c->width <= 0 || c->height <= 0 ||
      c->n_images < 1 || c->tile_cols < 1 || c->tile_rows < 1 || c->qindex < 0 || c->qindex > 255 ||
      c->n_cands < 1 || c->n_cands > 64 || c->k1 < 1 || c->k1 > c->n_cands || c->k3 < 1 || c->k3 > c->k1 ||
      c->rdoq_level < 0 || c->rdoq_level > 3 || c->txt_groups < 1 || c->txt_groups > 6 || c->txt_rate_th < 0 ||
      c->txt_satd_th < 0 || c->coeff_rate_est_lvl < 0 || c->coeff_rate_est_lvl > 2 || c->lambda == 0 ||
      (long)c->width * c->height * c->n_images > (1L << 31))
    return ZAV1_EINVAL;
  if (c->keep_levels != 0 && c->keep_levels != 1) return ZAV1_EINVAL;
  if (tree) {
    if (c->width % 64 || c->height % 64 || c->txt_groups_lt16 < 1 || c->txt_groups_lt16 > 6) return ZAV1_EINVAL;
  } else if ((c->leaf != 8 && c->leaf != 16 && c->leaf != 32 && c->leaf != 64) || c->width % c->leaf ||
             c->height % c->leaf || c->leaf > c->sb) {
    return ZAV1_EINVAL;
  }
  bool fi_seen = false;
  for (int i = 0; i < c->n_cands; i++) {
    const zav1_md_cand m = c->cands[i];
    if (m.mode > 12 || m.delta < -3 || m.delta > 3 || m.fi_mode > 5) return ZAV1_EINVAL;
    if (m.fi_mode < 5) {
      if (m.mode != 0 || (!tree && c->leaf > 32)) return ZAV1_EINVAL;
      fi_seen = true;
    } else if (fi_seen) {
      return ZAV1_EINVAL;  // filter-intra candidates come last
    }
  }
  if (tree) {
    const size_t nu = (size_t)(c->width / 64) * (c->height / 64) * c->n_images;
    for (size_t u = 0; u < nu; u++) {
      const int rc = check_map(c->leaf_map + u * 256);
      if (rc) return rc;
    }
  }
  return ZAV1_OK;
}

int plan_cfg(zav1::Wavefront* w, const zav1_wf_cfg* cfg) {
  w->cfg = *cfg;
  if (cfg->leaf_map) {
    const size_t nu = (size_t)(cfg->width / 64) * (cfg->height / 64) * cfg->n_images;
    w->leaf_map.assign(cfg->leaf_map, cfg->leaf_map + nu * 256);
  }
  w->cfg.leaf_map = nullptr;  // the plan keeps its own copy
  return zav1::plan(w);
}
}  // namespace

extern "C" {

int zav1_wf_footprint(const zav1_wf_cfg* cfg, size_t* bytes) {
  if (!bytes) return ZAV1_EINVAL;
  int rc = check_cfg(cfg);
  if (rc) return rc;
  zav1::Wavefront w;
  plan_cfg(&w, cfg);
  *bytes = zav1::footprint(w);
  return ZAV1_OK;
}

int zav1_wf_create(const zav1_wf_cfg* cfg, zav1_wf** out) {
  if (!out) return ZAV1_EINVAL;
  *out = nullptr;
  int rc = check_cfg(cfg);
  if (rc) return rc;
  auto* w = new zav1::Wavefront();
  plan_cfg(w, cfg);
  const zav1_wf_cfg& c = w->cfg;
  // tables of the current device (built by zav1_init) and the host-side rates
  rc = zav1::device_tables(&w->row, &w->iscan, &w->fscan_base, c.bd == 10 ? 256 + c.qindex : c.qindex);
  if (rc) {
    delete w;
    return rc;
  }
  zav1_quant_row qr;
  zav1_build_quant_row(c.qindex, c.bd, &qr);
  w->dq0 = qr.dequant[0], w->dq1 = qr.dequant[1];
  w->rdmult = ((int64_t)c.lambda * 17 + 2) >> 2;  // rdoq_rdmult_full, luma, sharpness <= 2
  w->rate_th = (uint64_t)c.txt_rate_th;
  if (c.rdoq_level == 0) w->flags3 = ZAV1_TXU_QUANT_B | ZAV1_TXU_SPATIAL | ZAV1_TXU_RECON;
  else {
    w->flags3 = ZAV1_TXU_RDOQ | ZAV1_TXU_SPATIAL | ZAV1_TXU_RECON;
    w->cut_num = c.rdoq_level == 1 ? 0 : (c.rdoq_level == 2 ? 80 : 60);
    w->cut_den = c.rdoq_level == 1 ? 0 : 100;
  }
  w->flags3 |= zav1::kTxuLuma;  // every MDS3 unit is luma
  w->flags1 = ZAV1_TXU_QUANT_B;
  // the luma rate approximation prices MDS1 and MDS3 alike
  if (c.coeff_rate_est_lvl == 2) {
    w->flags1 |= ZAV1_TXU_RATE_EST2;
    w->flags3 |= ZAV1_TXU_RATE_EST2;
  }
  zav1_coeff_costs costs;
  zav1_build_coeff_costs(c.qindex, &costs);
  // one device allocation, carved
  w->bytes = zav1::footprint(*w);
  if (hipMalloc(&w->base, w->bytes) != hipSuccess) {
    g_wf_err = "zav1_wf_create: device allocation of " + std::to_string(w->bytes) + " bytes failed";
    delete w;
    return ZAV1_ENOMEM;
  }
  char* p = w->base;
  auto take = [&p](size_t b) {
    char* q = p;
    p += zav1::al(b);
    return (void*)q;
  };
  const size_t es = c.bd > 8 ? 2 : 1, plane = (size_t)c.width * w->stacked_h * es;
  const size_t n = (size_t)w->n_units;
  w->src = take(plane);
  w->recon = take(plane);
  w->units = (zav1::MdUnit*)take(n * sizeof(zav1::MdUnit));
  w->nbr = (int32_t*)take(n * 8);
  w->uid = (int32_t*)take(n * 4);
  w->dec = (zav1::WfDecision*)take(n * sizeof(zav1::WfDecision));
  w->mode_map = (uint8_t*)take(n);
  w->costs = (int32_t*)take(sizeof(zav1_coeff_costs));
  w->trellis_runs = (unsigned long long*)take(8);
  if (c.keep_levels) {
    w->levels = (int32_t*)take(w->n_levels * 4);
    w->loff = (uint64_t*)take(n * 8);
  }
  std::vector<int32_t> uid(n);
  for (size_t i = 0; i < n; i++) uid[i] = (int32_t)i;
  hipError_t e = hipSuccess;
  auto up = [&e](void* d, const void* h, size_t b) {
    if (e == hipSuccess && b) e = hipMemcpy(d, h, b, hipMemcpyHostToDevice);
  };
  up(w->units, w->h_units.data(), n * sizeof(zav1::MdUnit));
  up(w->nbr, w->h_nbr.data(), n * 8);
  up(w->uid, uid.data(), n * 4);
  up(w->costs, &costs, sizeof(costs));
  const unsigned long long zero = 0;
  up(w->trellis_runs, &zero, 8);
  if (c.keep_levels) up(w->loff, w->h_loff.data(), n * 8);
  for (int k = 0; k < 4; k++) {
    zav1::SizePlan& P = w->sp[k];
    if (!P.max_u) continue;
    const size_t U = (size_t)P.max_u, LL = (size_t)P.L * P.L, n3 = U * P.k3 * P.nt;
    P.fscan = w->fscan_base + P.tx * 3072;
    P.rate_tab = (uint32_t*)take(25 * P.nc * 4);
    P.txt = (int32_t*)take(13 * 16 * 4);
    P.cands = (zav1::WfCand*)take(64 * sizeof(zav1::WfCand));
    P.d_types = (uint8_t*)take(16);
    P.rate = (uint32_t*)take(U * P.nc * 4);
    P.satd = (uint32_t*)take(U * P.nc * 4);
    P.sel = (uint8_t*)take(U * P.k1);
    P.m3 = (uint8_t*)take(U * P.k1);
    P.sel_cost = (uint64_t*)take(U * P.k1 * 8);
    P.fs_jobs = (zav1::TxUnitJob*)take(U * P.k1 * sizeof(zav1::TxUnitJob));
    P.pred = (uint16_t*)take(U * P.k1 * LL * 2);
    P.jobs1 = (zav1::TxUnitRdJob*)take(U * P.k1 * sizeof(zav1::TxUnitRdJob));
    P.dist1 = (uint64_t*)take(U * P.k1 * 8);
    P.eob1 = (int32_t*)take(U * P.k1 * 4);
    P.rate1 = (int32_t*)take(U * P.k1 * 4);
    P.cul1 = (uint8_t*)take(U * P.k1);
    P.jobs3 = (zav1::TxUnitRdJob*)take(n3 * sizeof(zav1::TxUnitRdJob));
    P.dist3 = (uint64_t*)take(n3 * 8);
    P.csatd3 = (uint64_t*)take(n3 * 8);
    P.eob3 = (int32_t*)take(n3 * 4);
    P.rate3 = (int32_t*)take(n3 * 4);
    P.cul3 = (uint8_t*)take(n3);
    P.rec3 = (uint16_t*)take(n3 * LL * 2);
    P.ws1 = take(zav1::txunit_rd_workspace(P.tx, (int)(U * P.k1)));
    P.ws3_bytes = zav1::ws3_bytes(P.tx, (int)n3);
    P.ws3 = take(P.ws3_bytes);
    if (c.keep_levels) P.q3 = (int32_t*)take((size_t)P.k3 * P.nt * P.np * 4);
    std::vector<uint32_t> rtab(25 * (size_t)P.nc);
    for (int a = 0; a < 5; a++)
      for (int l = 0; l < 5; l++)
        for (int i = 0; i < P.nc; i++)
          rtab[(a * 5 + l) * P.nc + i] = (uint32_t)zav1_intra_mode_rate(P.L, P.L, a, l, c.cands[i].mode,
                                                                         c.cands[i].delta, c.cands[i].fi_mode);
    std::vector<int32_t> txt(13 * 16);
    for (int d = 0; d < 13; d++)
      for (int t = 0; t < 16; t++) txt[d * 16 + t] = zav1_intra_txt_rate(P.tx, d, t);
    up(P.rate_tab, rtab.data(), rtab.size() * 4);
    up(P.txt, txt.data(), txt.size() * 4);
    up(P.cands, P.h_cands.data(), P.h_cands.size() * sizeof(zav1::WfCand));
    up(P.d_types, P.types, 16);
  }
  // side streams when a layer holds steps of several leaf sizes (default
  // priority, non-blocking with respect to the null stream)
  int per_layer = 1;
  for (int s0 = 0, s1 = 0; s0 < w->n_steps; s0 = s1) {
    for (s1 = s0 + 1; s1 < w->n_steps && w->step_layer[s1] == w->step_layer[s0];) s1++;
    per_layer = std::max(per_layer, s1 - s0);
  }
  if (e == hipSuccess && per_layer > 1) {
    e = hipEventCreateWithFlags(&w->fork, hipEventDisableTiming);
    for (int k = 0; k < per_layer - 1 && e == hipSuccess; k++) {
      e = hipStreamCreateWithFlags(&w->side[k], hipStreamNonBlocking);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&w->join[k], hipEventDisableTiming);
    }
  }
  if (e != hipSuccess) {
    const int r = wf_fail(e, "zav1_wf_create upload");
    (void)zav1_wf_destroy(reinterpret_cast<zav1_wf*>(w));
    return r;
  }
  *out = reinterpret_cast<zav1_wf*>(w);
  return ZAV1_OK;
}

int zav1_wf_info(const zav1_wf* wf, zav1_wf_info_t* info) {
  if (!wf || !info) return ZAV1_EINVAL;
  const auto* w = reinterpret_cast<const zav1::Wavefront*>(wf);
  info->d_src = w->src;
  info->d_recon = w->recon;
  info->d_decisions = reinterpret_cast<zav1_wf_decision*>(w->dec);
  info->d_units = reinterpret_cast<const zav1_md_unit*>(w->units);
  info->n_leaves = w->n_units;
  info->n_steps = w->n_steps;
  info->max_leaves_per_step = w->max_u;
  info->stacked_height = w->stacked_h;
  info->bytes = w->bytes;
  // the tx types of the configuration's leaf size (uniform leaves), else of
  // the largest leaf size present
  int k = w->cfg.leaf == 8 ? 0 : (w->cfg.leaf == 16 ? 1 : (w->cfg.leaf == 32 ? 2 : 3));
  if (!w->sp[k].max_u)
    for (k = 3; k > 0 && !w->sp[k].max_u; k--) {
    }
  info->n_types = w->sp[k].nt;
  for (int i = 0; i < 16; i++) info->types[i] = w->sp[k].types[i];
  info->d_levels = w->levels;
  info->d_level_off = w->loff;
  info->n_levels = w->n_levels;
  info->d_trellis_runs = reinterpret_cast<const uint64_t*>(w->trellis_runs);
  return ZAV1_OK;
}

int zav1_wf_run(zav1_wf* wf, int use_graph, zav1_stream_t s) {
  if (!wf || (use_graph != 0 && use_graph != 1)) return ZAV1_EINVAL;
  auto* w = reinterpret_cast<zav1::Wavefront*>(wf);
  hipStream_t st = s ? (hipStream_t)s : hipStreamPerThread;
  if (!use_graph) {
    WCHECK(zav1::run_range(w, 0, w->n_steps, st));
    return ZAV1_OK;
  }
  if (st == hipStreamPerThread) {
    g_wf_err = "zav1_wf_run: graph replay needs a created stream";
    return ZAV1_EINVAL;
  }
  if (!w->graph || w->graph_stream != st) {
    if (w->graph) {
      WCHECK(hipGraphExecDestroy(w->graph));
      w->graph = nullptr;
    }
    hipGraph_t g = nullptr;
    WCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    e = zav1::run_range(w, 0, w->n_steps, st);
    const hipError_t ec = hipStreamEndCapture(st, &g);
    if (e != hipSuccess) return wf_fail(e, "zav1_wf_run capture");
    if (ec != hipSuccess) return wf_fail(ec, "hipStreamEndCapture");
    e = hipGraphInstantiate(&w->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return wf_fail(e, "hipGraphInstantiate");
    w->graph_stream = st;
  }
  WCHECK(hipGraphLaunch(w->graph, st));
  return ZAV1_OK;
}

int zav1_wf_run_steps(zav1_wf* wf, int first_step, int n_steps, zav1_stream_t s) {
  if (!wf) return ZAV1_EINVAL;
  auto* w = reinterpret_cast<zav1::Wavefront*>(wf);
  if (first_step < 0 || n_steps < 0 || first_step + n_steps > w->n_steps) return ZAV1_EINVAL;
  hipStream_t st = s ? (hipStream_t)s : hipStreamPerThread;
  WCHECK(zav1::run_range(w, first_step, first_step + n_steps, st));
  return ZAV1_OK;
}

int zav1_wf_run_timed(zav1_wf* wf, zav1_stream_t s, double* stage_ms, int n_stages) {
  if (!wf || !stage_ms || n_stages < ZAV1_WF_STAGES) return ZAV1_EINVAL;
  auto* w = reinterpret_cast<zav1::Wavefront*>(wf);
  hipStream_t st = s ? (hipStream_t)s : hipStreamPerThread;
  zav1::StageClock clk{};
  int made = 0;
  hipError_t e = hipSuccess;
  for (; made <= ZAV1_WF_STAGES && e == hipSuccess; made++) e = hipEventCreate(&clk.ev[made]);
  for (int k = 0; k < ZAV1_WF_STAGES; k++) stage_ms[k] = 0.0;
  for (int i = 0; i < w->n_steps && e == hipSuccess; i++) {
    if (w->step_off[i + 1] == w->step_off[i]) continue;
    e = zav1::step_chain(w, i, st, &clk);
    if (e == hipSuccess) e = hipEventSynchronize(clk.ev[ZAV1_WF_STAGES]);
    for (int k = 0; k < ZAV1_WF_STAGES && e == hipSuccess; k++) {
      float ms = 0.f;
      e = hipEventElapsedTime(&ms, clk.ev[k], clk.ev[k + 1]);
      stage_ms[k] += ms;
    }
  }
  for (int k = 0; k < made; k++)
    if (clk.ev[k]) (void)hipEventDestroy(clk.ev[k]);
  if (e != hipSuccess) return wf_fail(e, "zav1_wf_run_timed");
  return ZAV1_OK;
}

int zav1_wf_autotune(zav1_wf* wf, int first_step, int n_steps, zav1_stream_t s, int32_t* forms, double* ms) {
  if (!wf || !forms) return ZAV1_EINVAL;
  auto* w = reinterpret_cast<zav1::Wavefront*>(wf);
  if (first_step < 0 || n_steps < 1 || first_step + n_steps > w->n_steps) return ZAV1_EINVAL;
  hipStream_t st = s ? (hipStream_t)s : hipStreamPerThread;
  const bool rdoq = (w->flags3 & ZAV1_TXU_RDOQ) != 0;
  constexpr int kForms = 3;  // queued group, direct latency, lane
  double t[4][kForms] = {};
  zav1::StageClock clk{};
  int made = 0;
  hipError_t e = hipSuccess;
  for (; made <= ZAV1_WF_STAGES && e == hipSuccess; made++) e = hipEventCreate(&clk.ev[made]);
  // the window's trellis time per leaf size under each form (re-deciding the
  // window's leaves gives the results they already have)
  for (int f = 0; f < kForms && rdoq && e == hipSuccess; f++) {
    for (int k = 0; k < 4; k++) w->sp[k].rdoq_form = f;
    for (int i = first_step; i < first_step + n_steps && e == hipSuccess; i++) {
      if (w->step_off[i + 1] == w->step_off[i]) continue;
      e = zav1::step_chain(w, i, st, &clk);
      if (e == hipSuccess) e = hipEventSynchronize(clk.ev[ZAV1_WF_STAGES]);
      float tm = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&tm, clk.ev[8], clk.ev[9]);
      t[w->step_sz[i]][f] += tm;
    }
  }
  for (int k = 0; k < made; k++)
    if (clk.ev[k]) (void)hipEventDestroy(clk.ev[k]);
  for (int k = 0; k < 4; k++) {
    zav1::SizePlan& P = w->sp[k];
    int best = -1;
    if (P.max_u && rdoq && t[k][0] + t[k][1] + t[k][2] > 0.0)  // the size ran in the window
      for (int f = 0; f < kForms; f++)
        if (best < 0 || t[k][f] < t[k][best]) best = f;
    P.rdoq_form = best;
    forms[k] = best;
    if (ms)
      for (int f = 0; f < kForms; f++) ms[k * kForms + f] = t[k][f];
  }
  if (w->graph) {  // captured with the previous forms
    (void)hipGraphExecDestroy(w->graph);
    w->graph = nullptr;
  }
  if (e != hipSuccess) return wf_fail(e, "zav1_wf_autotune");
  return ZAV1_OK;
}

int zav1_wf_run_plain(zav1_wf* wf, int first_leaf, int n_leaves, zav1_stream_t s) {
  if (!wf) return ZAV1_EINVAL;
  auto* w = reinterpret_cast<zav1::Wavefront*>(wf);
  if (first_leaf < 0 || n_leaves < 0 || first_leaf + n_leaves > w->n_units) return ZAV1_EINVAL;
  hipStream_t st = s ? (hipStream_t)s : hipStreamPerThread;
  for (int u = first_leaf; u < first_leaf + n_leaves; u++) WCHECK(zav1::plain_unit(w, u, st));
  return ZAV1_OK;
}

int zav1_wf_destroy(zav1_wf* wf) {
  if (!wf) return ZAV1_EINVAL;
  auto* w = reinterpret_cast<zav1::Wavefront*>(wf);
  if (w->graph) (void)hipGraphExecDestroy(w->graph);
  for (int k = 0; k < 3; k++) {
    if (w->side[k]) (void)hipStreamDestroy(w->side[k]);
    if (w->join[k]) (void)hipEventDestroy(w->join[k]);
  }
  if (w->fork) (void)hipEventDestroy(w->fork);
  if (w->base) (void)hipFree(w->base);
  if (w->h_dec_pin) (void)hipHostFree(w->h_dec_pin);
  if (w->h_lev_pin) (void)hipHostFree(w->h_lev_pin);
  delete w;
  return ZAV1_OK;
}

int zav1_wf_write_stills(zav1_wf* wf, int first, int n, int threads, uint8_t* out, size_t cap, size_t* sizes,
                         zav1_stream_t s) {
  if (!wf || !sizes || first < 0 || n < 1 || threads < 0) return ZAV1_EINVAL;
  auto* w = reinterpret_cast<zav1::Wavefront*>(wf);
  const zav1_wf_cfg& c = w->cfg;
  if (!c.keep_levels || first + n > c.n_images) return ZAV1_EINVAL;
  hipStream_t st = s ? (hipStream_t)s : hipStreamPerThread;
  const int nu = w->n_units;
  if (!w->h_dec_pin) {
    void *pd = nullptr, *pl = nullptr;
    if (hipHostMalloc(&pd, sizeof(zav1::WfDecision) * nu, hipHostMallocDefault) != hipSuccess) {
      g_wf_err = "zav1_wf_write_stills: pinned staging of the decisions";
      return ZAV1_ENOMEM;
    }
    if (hipHostMalloc(&pl, 4 * std::max<size_t>(1, w->n_levels), hipHostMallocDefault) != hipSuccess) {
      (void)hipHostFree(pd);
      g_wf_err = "zav1_wf_write_stills: pinned staging of the levels";
      return ZAV1_ENOMEM;
    }
    w->h_dec_pin = (zav1::WfDecision*)pd, w->h_lev_pin = (int32_t*)pl;
  }
  const zav1::WfDecision* dec = w->h_dec_pin;
  WCHECK(hipMemcpyAsync(w->h_dec_pin, w->dec, sizeof(zav1::WfDecision) * nu, hipMemcpyDeviceToHost, st));
  const uint64_t l0 = w->img_loff[first], l1 = w->img_loff[first + n];
  const int32_t* lev = w->h_lev_pin + l0;
  WCHECK(hipMemcpyAsync(w->h_lev_pin + l0, w->levels + l0, 4 * (l1 - l0), hipMemcpyDeviceToHost, st));
  WCHECK(hipStreamSynchronize(st));
  std::vector<std::vector<zav1_still_leaf>> per(n);
  for (int u = 0; u < nu; u++) {
    const int img = w->h_units[u].y / c.height;
    if (img < first || img >= first + n) continue;
    const zav1::WfDecision& d = dec[u];
    zav1_still_leaf l{};
    l.x = w->h_units[u].x, l.y = w->h_units[u].y - img * c.height;
    l.size = (uint8_t)(8 << w->h_sz[u]);
    l.mode = d.mode, l.delta = c.cands[d.cand].delta, l.fi_mode = d.fi, l.tx_type = d.tx_type;
    l.eob = d.eob, l.level_off = w->h_loff[u] - l0;
    per[img - first].push_back(l);
  }
  // the images' streams on `threads` host threads (0: the host's
  // concurrency): whole images per thread, the tiles of an image split
  // among the threads left over when there are fewer images than threads
  int nth = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
  const int img_threads = std::min(nth, n), tile_threads = std::max(1, nth / img_threads);
  std::vector<std::vector<uint8_t>> streams(n);
  std::vector<int> rcs(n, ZAV1_OK);
  std::vector<std::string> why(n);
  auto one = [&](int i) {
    zav1_still_desc sd{};
    sd.width = c.width, sd.height = c.height, sd.bd = c.bd, sd.sb = c.sb;
    sd.tile_cols = c.tile_cols, sd.tile_rows = c.tile_rows, sd.qindex = c.qindex;
    sd.n_leaves = (int32_t)per[i].size(), sd.threads = tile_threads;
    sd.leaves = per[i].data(), sd.levels = lev;
    rcs[i] = zav1::write_still_vec(&sd, streams[i], &why[i]);
  };
  if (img_threads <= 1) {
    for (int i = 0; i < n; i++) one(i);
  } else {
    std::vector<std::thread> pool;
    for (int k = 0; k < img_threads; k++)
      pool.emplace_back([&, k] {
        for (int i = k; i < n; i += img_threads) one(i);
      });
    for (auto& th : pool) th.join();
  }
  size_t total = 0;
  for (int i = 0; i < n; i++) {
    if (rcs[i]) {
      g_wf_err = "zav1_wf_write_stills: image " + std::to_string(first + i) + ": " + why[i];
      return rcs[i];
    }
    sizes[i] = streams[i].size();
    total += sizes[i];
  }
  if (!out || cap < total) return ZAV1_ENOMEM;
  size_t used = 0;
  for (int i = 0; i < n; i++) {
    std::memcpy(out + used, streams[i].data(), sizes[i]);
    used += sizes[i];
  }
  return ZAV1_OK;
}

const char* zav1_wf_last_error(void) { return g_wf_err.c_str(); }

}  // extern "C"
