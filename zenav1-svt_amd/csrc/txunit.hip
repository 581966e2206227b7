// Fused transform units for gfx950: for every job (one transform unit of one
// tx size with its own type and prediction) residual -> forward 2D ->
// 64-point fold -> quantize (fp or b) -> eob, cul level, frequency-domain
// distortion; and, when reconstructing, inverse 2D -> recon clip -> spatial
// distortion over the cropped extent.  Residual, coefficients and the
// inverse stay in VGPRs / LDS; per job only the results (and optionally the
// packed levels and the reconstruction) reach HBM.
//
// Semantics: rust/crates/svtav1-encoder/src/leaf_funnel/tx_pipeline.rs:268
// tx_unit_inner (8-bit) and :854 tx_unit_hbd (10-bit), without the RDOQ
// trellis, quantization matrices and coefficient rate; compute_cul_level
// :130; recon_add_clamp rust/crates/svtav1-dsp/src/residual.rs:140; spatial
// SSE rust/crates/svtav1-dsp/src/variance.rs:60 / hbd.rs:936.
#include "zav1_dev.h"
#include "zav1_kernels.h"

namespace zav1 {

template <int W, int H>
struct TuGeom {
  static constexpr int T = W > H ? W : H;  // lanes per job
  static constexpr int JPW = 256 / T;      // jobs per workgroup
  static constexpr int LS = W + 1;         // padded LDS row stride
  static constexpr int BLK = H * LS;
  static constexpr int PW = W < 32 ? W : 32, PH = H < 32 ? H : 32;
  static constexpr bool RECT2 = (W == 2 * H) || (H == 2 * W);
};

// Inverse 2D of the dequantized block in LDS (0 where folded) + prediction,
// clipped to bd; rec (w*h, may be null) receives the reconstruction; returns
// this lane's part of the spatial SSE over crop_w x crop_h (0 unless
// `spatial`).  Lanes t < H run the row pass, t < W the column pass; the caller
// reduces the SSE over its lane group and synchronised the LDS block.
template <int W, int H, typename Pix>
__device__ uint64_t tu_recon(int32_t* blk, int t, bool active, const TypeCfg& cfg, int bd, const uint16_t* pr,
                             const Pix* sp, int stride, int crop_w, int crop_h, bool spatial, uint16_t* rec_out) {
  using G = TuGeom<W, H>;
  constexpr InvShift IS = inv_shift(W, H);
  uint64_t sdist = 0;
  const int row_rng = bd + 8, col_rng = (bd + 6) > 16 ? bd + 6 : 16;
  const int32_t wrap_max = (1 << (7 + bd)) - 1 + (914 << (bd - 7));
  if (active && t < H) {
    int32_t v[W];
#pragma unroll
    for (int c = 0; c < W; c++) {
      int32_t x = blk[t * G::LS + c];
      if constexpr (G::RECT2) x = rshift64r((int64_t)x * 2896, 12);
      v[c] = zclamp(x, row_rng);
    }
    inv1d<W>(v, cfg.row, row_rng);
#pragma unroll
    for (int c = 0; c < W; c++) blk[t * G::LS + c] = rshift_round(v[c], IS.r0);
  }
  __syncthreads();
  if (active && t < W) {
    const int cc = cfg.lr ? (W - 1 - t) : t;
    int32_t v[H];
#pragma unroll
    for (int r = 0; r < H; r++) v[r] = zclamp(blk[r * G::LS + cc], col_rng);
    inv1d<H>(v, cfg.col, col_rng);
    const int32_t maxv = (1 << bd) - 1;
#pragma unroll
    for (int r = 0; r < H; r++) {
      int32_t x = rshift_round(v[r], IS.r1);
      x = x < -wrap_max - 1 ? -wrap_max - 1 : (x > wrap_max ? wrap_max : x);
      const int rr = cfg.ud ? (H - 1 - r) : r;
      int32_t rec = (int32_t)pr[rr * W + t] + x;
      rec = rec < 0 ? 0 : (rec > maxv ? maxv : rec);
      if (rec_out) rec_out[rr * W + t] = (uint16_t)rec;
      if (spatial && rr < crop_h && t < crop_w) {
        const int32_t d = (int32_t)sp[rr * stride + t] - rec;
        sdist += (uint64_t)(uint32_t)__mul24(d, d);
      }
    }
  }
  return group_sum_u64<G::T>(sdist);
}

// A job whose (type, size) has no 1D kernel (ADST on a 32/64-point
// dimension) gets eob = -1 and dist = 0 (the reference's `false` return); a
// job whose unit or prediction lies outside the plane / prediction buffer
// gets eob = -22 and is not read.
template <int W, int H, typename Pix>
__global__ __launch_bounds__(256) void k_txunit(const Pix* __restrict__ src, int stride, int plane_w, int plane_h,
                                                int bd, const uint16_t* __restrict__ pred, long pred_len,
                                                const TxUnitJob* __restrict__ jobs, int n_jobs,
                                                const QuantRow* __restrict__ row_p,
                                                const uint16_t* __restrict__ iscan_all, int tx_size, unsigned flags,
                                                uint64_t* __restrict__ out_dist, int32_t* __restrict__ out_eob,
                                                uint8_t* __restrict__ out_cul, int32_t* __restrict__ out_q,
                                                uint16_t* __restrict__ out_rec, const TuKeep keep) {
  using G = TuGeom<W, H>;
  constexpr int CBC = fwd_cb_col(W, H), CBR = fwd_cb_row(W, H);
  constexpr FwdShift FS = fwd_shift(W, H);
  // input bounds of the 1D passes for residuals of <= 10-bit planes: the
  // passes whose bound allows it run with 24-bit-operand butterflies
  constexpr long long kBCol = fwd_col_bound(FS.s0);
  constexpr long long kBRow = fwd_row_bound(H, CBC, FS.s1, kBCol);
  constexpr int LSC = tx_log_scale(W, H);
  __shared__ int32_t lds[G::JPW * G::BLK];

  const int jl = threadIdx.x / G::T, t = threadIdx.x - jl * G::T;
  const int job = blockIdx.x * G::JPW + jl;
  const TxUnitJob jb = jobs[job < n_jobs ? job : 0];
  const TypeCfg cfg = type_cfg(jb.tx_type & 15);
  const bool ok = jb.tx_type < 16 && kind_ok(cfg.col, H) && kind_ok(cfg.row, W);
  const bool inside = jb.x >= 0 && jb.y >= 0 && jb.x + W <= plane_w && jb.y + H <= plane_h &&
                      (long)jb.pred_off + W * H <= pred_len;
  const bool active = job < n_jobs && ok && inside;
  const uint16_t* pr = pred + jb.pred_off;
  const Pix* sp = src + (long)jb.y * stride + jb.x;
  int32_t* blk = lds + jl * G::BLK;

  // ---- forward column pass: lane t < W transforms column t of the residual ----
  if (active && t < W) {
    int32_t v[H];
#pragma unroll
    for (int r = 0; r < H; r++) {
      const int rr = cfg.ud ? (H - 1 - r) : r;
      v[r] = ((int32_t)sp[rr * stride + t] - (int32_t)pr[rr * W + t]) << FS.s0;
    }
    fwd1d<H, CBC, kBCol>(v, cfg.col);
    const int cc = cfg.lr ? (W - 1 - t) : t;
#pragma unroll
    for (int r = 0; r < H; r++) blk[r * G::LS + cc] = rshift_round32(v[r], FS.s1);
  }
  __syncthreads();

  // ---- forward row pass + fold + quantize: lane t < H owns row t ----
  // The arrays the later kernels read (coefficients, levels, dequantized
  // values; packed PW x PH per job) leave through the job's LDS block, so
  // each store instruction writes a contiguous run of the job's array
  // instead of one word per row; the dequantized values stay in LDS (0
  // where folded) as the inverse transform's input.
  uint64_t fdist = 0, fold = 0, csatd = 0;
  int32_t eob = 0, lsum = 0, dcq = 0;
  int32_t v[W];
  const bool row_lane = active && t < H;
  if (row_lane) {
#pragma unroll
    for (int c = 0; c < W; c++) v[c] = blk[t * G::LS + c];
    fwd1d<W, CBR, kBRow>(v, cfg.row);
#pragma unroll
    for (int c = 0; c < W; c++) {
      int32_t x = rshift_round32(v[c], FS.s2);
      // 2:1 rectangles: x NewSqrt2 / 2^12; |x| stays below 2^18 (the pinned
      // coefficient bound before this scaling), so the product fits 32 bits
      if constexpr (G::RECT2) x = (__mul24(x, 5793) + 2048) >> 12;
      v[c] = x;
    }
  }
  const long pbase = (long)job * (G::PW * G::PH);
  // packed element e of the job: lane t moves e = t, t + T, ...
  auto store_packed = [&](int32_t* dst) {
    __syncthreads();
    if (active) {
      for (int e = t; e < G::PW * G::PH; e += G::T) dst[pbase + e] = blk[(e / G::PW) * G::LS + e % G::PW];
    }
    __syncthreads();
  };
  if (keep.t) {
    if (row_lane) {
#pragma unroll
      for (int c = 0; c < G::PW; c++)
        if (t < G::PH) blk[t * G::LS + c] = v[c];
    }
    store_packed(keep.t);
  }
  int32_t dk[G::PW];
  if (row_lane) {
    const QuantRow& qr = *row_p;
    const int cls = (jb.tx_type >= 10) ? (((jb.tx_type - 10) & 1) ? 2 : 1) : 0;
    const uint16_t* iscan = iscan_all + (tx_size * 3 + cls) * 1024 + (t < 32 ? t : 0) * G::PW;
    int32_t sp[G::PW];  // scan positions, loaded unconditionally (no load behind the level test)
#pragma unroll
    for (int c = 0; c < G::PW; c++) sp[c] = (int32_t)iscan[c] + 1;
    const bool use_b = flags & kTxuQuantB;
#pragma unroll
    for (int c = 0; c < W; c++) {
      const int32_t x = v[c];
      if (keep.csatd) csatd += (uint64_t)(x < 0 ? -(int64_t)x : (int64_t)x);
      if (t >= 32 || c >= 32) {
        const uint64_t e2 = (uint64_t)((int64_t)x * x);  // energy of the folded-away coefficients
        fdist += e2;
        fold += e2;
        continue;
      }
      const int iz = (t | c) != 0;
      const int32_t sgn = x >> 31;
      const int32_t a = (x ^ sgn) - sgn;
      const int32_t q = use_b ? quant_b_level(a, qr, iz, LSC, bd) : quant_fp_level(a, qr, iz, LSC, bd);
      int32_t dq = 0, qs = 0;
      if (q != 0) {
        dq = (dequant_level(q, qr, iz, LSC) ^ sgn) - sgn;
        qs = (q ^ sgn) - sgn;
        lsum += q;
        eob = max(eob, sp[c]);
      }
      if (t == 0 && c == 0) dcq = qs;
      const int64_t e = (int64_t)x - dq;
      fdist += (uint64_t)e * (uint64_t)e;
      dk[c] = dq;
      blk[t * G::LS + c] = qs;  // levels first through the block
    }
  }
  if (out_q) store_packed(out_q);
  else __syncthreads();
  if (row_lane) {
#pragma unroll
    for (int c = 0; c < W; c++) blk[t * G::LS + c] = (t >= 32 || c >= 32) ? 0 : dk[c < G::PW ? c : 0];
  }
  // (the dequantized values are not kept: the finish and the trellis
  // recompute them from the levels, sign * ((|q| * dequant) >> log_scale))
  fdist = group_sum_u64<G::T>(fdist);
  if constexpr (W == 64 || H == 64) fold = group_sum_u64<G::T>(fold);
  if (keep.csatd) csatd = group_sum_u64<G::T>(csatd);
  eob = group_max_i32<G::T>(eob);
  lsum = group_sum_i32<G::T>(lsum);

  // ---- reconstruction: inverse 2D of the dequantized block + prediction ----
  uint64_t sdist = 0;
  const bool spatial = flags & kTxuSpatial;
  if (flags & (kTxuRecon | kTxuSpatial)) {
    __syncthreads();
    sdist = tu_recon<W, H>(blk, t, active, cfg, bd, pr, sp, stride, jb.crop_w, jb.crop_h, spatial,
                           out_rec ? out_rec + (long)job * (W * H) : nullptr);
  }

  if (job < n_jobs && t == 0) {
    if (!ok || !inside) {
      out_dist[job] = 0;
      out_eob[job] = !inside ? -22 : -1;
      out_cul[job] = 0;
      if (keep.eob16) {
        keep.eob16[job] = 0;
        keep.fold[job] = 0;
      }
      if (keep.csatd) keep.csatd[job] = 0;
      return;
    }
    constexpr int SHF = (1 - LSC) * 2;
    out_dist[job] = spatial ? (sdist << 4) : (SHF < 0 ? (fdist << (-SHF)) : (fdist >> SHF));
    out_eob[job] = eob;
    if (keep.eob16) {
      keep.eob16[job] = (uint16_t)eob;
      keep.fold[job] = fold;
    }
    if (keep.csatd) keep.csatd[job] = csatd;
    uint32_t cl = (uint32_t)(lsum < 63 ? lsum : 63);
    if (dcq < 0) cl |= 1u << 6;
    else if (dcq > 0) cl += 2u << 6;
    out_cul[job] = (uint8_t)cl;
  }
}

template <int S>
static hipError_t launch_tu(const void* src, int stride, int pw, int ph, int bd, const uint16_t* pred, long pred_len,
                            const TxUnitJob* jobs, int n_jobs, const QuantRow* row, const uint16_t* iscan,
                            unsigned flags, uint64_t* dist, int32_t* eob, uint8_t* cul, int32_t* q, uint16_t* rec,
                            hipStream_t st, TuKeep keep) {
  constexpr int W = tx_w(S), H = tx_h(S);
  const int grid = (n_jobs + TuGeom<W, H>::JPW - 1) / TuGeom<W, H>::JPW;
  if (bd == 8)
    hipLaunchKernelGGL((k_txunit<W, H, uint8_t>), dim3(grid), dim3(256), 0, st, (const uint8_t*)src, stride, pw, ph,
                       bd, pred, pred_len, jobs, n_jobs, row, iscan, S, flags, dist, eob, cul, q, rec, keep);
  else
    hipLaunchKernelGGL((k_txunit<W, H, uint16_t>), dim3(grid), dim3(256), 0, st, (const uint16_t*)src, stride, pw,
                       ph, bd, pred, pred_len, jobs, n_jobs, row, iscan, S, flags, dist, eob, cul, q, rec, keep);
  return hipGetLastError();
}

hipError_t txunit_batch(const void* src, int stride, int plane_w, int plane_h, int bd, int tx_size,
                        const uint16_t* pred, long pred_len, const TxUnitJob* jobs, int n_jobs, const QuantRow* row,
                        const uint16_t* iscan, unsigned flags, uint64_t* dist, int32_t* eob, uint8_t* cul,
                        int32_t* qcoeff, uint16_t* recon, hipStream_t st, TuKeep keep) {
  if (n_jobs <= 0) return hipSuccess;
  switch (tx_size) {
#define X(S)                                                                                                    \
  case S:                                                                                                       \
    return launch_tu<S>(src, stride, plane_w, plane_h, bd, pred, pred_len, jobs, n_jobs, row, iscan, flags, dist, \
                        eob, cul, qcoeff, recon, st, keep);
    X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18)
#undef X
    default: return hipErrorInvalidValue;
  }
}

// ---------------------------------------------------------------------------
// Full-RD transform unit (MDS3): forward path, RDOQ, rate, distortion.
// Semantics: tx_unit_inner with do_rdoq and RateMode::Exact
// (rust/crates/svtav1-encoder/src/leaf_funnel/tx_pipeline.rs:268-640).
// ---------------------------------------------------------------------------

// tx class (TX_TYPE_TO_CLASS): V_* types are the vertical class (2), H_* the
// horizontal one (1), the rest 2D (0)
__device__ inline uint8_t tx_class_of(uint8_t tt) { return tt >= 10 ? ((tt & 1) ? 1 : 2) : 0; }

__global__ __launch_bounds__(256) void k_rd_prep(const TxUnitRdJob* __restrict__ jobs, int n_jobs, int np,
                                                 TxUnitJob* __restrict__ tj, RdoqJob* __restrict__ rj,
                                                 RateJob* __restrict__ cj, unsigned* __restrict__ qctr) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j == 0) qctr[0] = qctr[1] = 0;  // the trellis queue's length and head, for this batch
  if (j >= n_jobs) return;
  const TxUnitRdJob a = jobs[j];
  TxUnitJob u;
  u.x = a.x, u.y = a.y, u.pred_off = a.pred_off, u.tx_type = a.tx_type, u.crop_w = a.crop_w, u.crop_h = a.crop_h;
  u.pad = 0;
  tj[j] = u;
  const uint8_t cls = tx_class_of(a.tx_type & 15);
  RdoqJob r{};
  r.rdmult = a.rdmult, r.coeff = (uint32_t)j * (uint32_t)np, r.tx_class = cls, r.plane_type = a.plane_type;
  r.txb_skip_ctx = a.txb_skip_ctx, r.dc_sign_ctx = a.dc_sign_ctx, r.sharpness = a.sharpness;
  rj[j] = r;
  RateJob c{};
  c.coeff = (uint32_t)j * (uint32_t)np, c.txt_rate = a.txt_rate, c.tx_class = cls, c.plane_type = a.plane_type;
  c.txb_skip_ctx = a.txb_skip_ctx, c.dc_sign_ctx = a.dc_sign_ctx;
  cj[j] = c;
}

// Per job after the trellis and the rate: frequency distortion over the
// final levels (+ the fold's energy), or the spatial one after the inverse;
// the recon; the cul level byte; the skip cost for blocks that end with
// eob 0; the packed levels.  A job the forward pass rejected keeps its
// status (eob -1 / -22) and gets rate 0.
template <int W, int H, typename Pix>
__global__ __launch_bounds__(256) void k_unit_finish(const Pix* __restrict__ src, int stride, int bd,
                                                     const uint16_t* __restrict__ pred,
                                                     const TxUnitRdJob* __restrict__ jobs, int n_jobs,
                                                     const int32_t* __restrict__ tq, const int32_t* __restrict__ tt,
                                                     const QuantRow* __restrict__ row_p,
                                                     const uint16_t* __restrict__ eob16,
                                                     const uint64_t* __restrict__ fold,
                                                     const int32_t* __restrict__ costs, int txs_ctx, unsigned flags,
                                                     uint64_t* __restrict__ out_dist, int32_t* __restrict__ out_eob,
                                                     int32_t* __restrict__ out_rate, uint8_t* __restrict__ out_cul,
                                                     int32_t* __restrict__ out_q, uint16_t* __restrict__ out_rec) {
  using G = TuGeom<W, H>;
  constexpr int LSC = tx_log_scale(W, H);
  constexpr int NP = G::PW * G::PH;
  __shared__ int32_t lds[G::JPW * G::BLK];
  const int jl = threadIdx.x / G::T, t = threadIdx.x - jl * G::T;
  const int job = blockIdx.x * G::JPW + jl;
  const bool live = job < n_jobs;
  const bool active = live && out_eob[job] >= 0;
  const TxUnitRdJob jb = jobs[live ? job : 0];
  const TypeCfg cfg = type_cfg(jb.tx_type & 15);
  int32_t* blk = lds + jl * G::BLK;
  const long base = (long)job * NP;
  uint64_t fd = 0;
  int32_t lsum = 0;
  // Without the trellis and the reconstruction the forward pass's
  // distortion, eob and cul level are final (it wrote them, and the
  // coefficients it would otherwise keep were not written): only the skip
  // cost and the rate estimate are left to apply.
  const bool light = !(flags & (kTxuRdoq | kTxuRecon | kTxuSpatial));
  if (light) {
    if (active && out_q)
      for (int e = t; e < NP; e += G::T) out_q[base + e] = tq[base + e];
    if (!live || t != 0) return;
    if (!active) {
      out_rate[job] = 0;
      return;
    }
    const int eob = eob16[job];
    if (eob == 0) out_rate[job] = costs[(txs_ctx * 2 + (jb.plane_type ? 1 : 0)) * 944 + jb.txb_skip_ctx * 2 + 1];
    if ((flags & kTxuRateEst2) && jb.plane_type == 0 && eob < ((W * H) >> 6)) out_rate[job] = 6000 + 1000 * eob;
    return;
  }
  // the job's packed arrays read as contiguous runs (lane t: elements t,
  // t + T, ...), the dequantized block laid out in LDS for the inverse; the
  // coefficients only for the frequency distortion (not read when the
  // distortion is spatial)
  const bool spatial = flags & kTxuSpatial;
  if (active) {
    const QuantRow& qr = *row_p;
    for (int e = t; e < NP; e += G::T) {
      const int32_t q = tq[base + e];
      const int32_t dm = dequant_level(q < 0 ? -q : q, qr, e != 0, LSC);
      const int32_t d = q < 0 ? -dm : dm;
      if (!spatial) {
        const int64_t df = (int64_t)tt[base + e] - d;
        fd += (uint64_t)(df * df);
      }
      lsum += q < 0 ? -q : q;
      if (out_q) out_q[base + e] = q;
      blk[(e / G::PW) * G::LS + e % G::PW] = d;
    }
    if constexpr (W > G::PW || H > G::PH) {  // the folded region reads as 0
      for (int e = t; e < W * H; e += G::T) {
        const int r = e / W, c = e % W;
        if (r >= G::PH || c >= G::PW) blk[r * G::LS + c] = 0;
      }
    }
  }
  fd = group_sum_u64<G::T>(fd);
  lsum = group_sum_i32<G::T>(lsum);
  uint64_t sdist = 0;
  if (flags & (kTxuRecon | kTxuSpatial)) {
    __syncthreads();
    sdist = tu_recon<W, H>(blk, t, active, cfg, bd, pred + jb.pred_off, src + (long)jb.y * stride + jb.x, stride,
                           jb.crop_w, jb.crop_h, spatial, (out_rec && active) ? out_rec + (long)job * (W * H) : nullptr);
  }
  if (!live || t != 0) return;
  if (!active) {
    out_rate[job] = 0;
    return;
  }
  const int eob = eob16[job];
  constexpr int SHF = (1 - LSC) * 2;
  const uint64_t f = fd + fold[job];
  out_dist[job] = spatial ? (sdist << 4) : (SHF < 0 ? (f << (-SHF)) : (f >> SHF));
  out_eob[job] = eob;
  // cost_skip_txb: txb_skip_cost[ctx][1] of the unit's (txs_ctx, plane) tables
  if (eob == 0) out_rate[job] = costs[(txs_ctx * 2 + (jb.plane_type ? 1 : 0)) * 944 + jb.txb_skip_ctx * 2 + 1];
  // coeff_rate_est_lvl 2 (tx_pipeline.rs:605): small-eob luma units take the
  // closed form in place of the exact rate (eob 0 included)
  if ((flags & kTxuRateEst2) && jb.plane_type == 0 && eob < ((W * H) >> 6)) out_rate[job] = 6000 + 1000 * eob;
  uint32_t cl = (uint32_t)(lsum < 63 ? lsum : 63);
  const int32_t dcq = tq[base];
  if (dcq < 0) cl |= 1u << 6;
  else if (dcq > 0) cl += 2u << 6;
  out_cul[job] = (uint8_t)cl;
}

// The pre-trellis SATD gate (TuSatdGate): per job against its group's first.
// With `queue`, the jobs it hands to the trellis are also appended to the
// trellis's work queue (queue[0 .. qctr[0]); order within it is free).
__global__ __launch_bounds__(256) void k_tu_satd_gate(int n_jobs, int group, const uint64_t* __restrict__ csatd,
                                                      int64_t th, uint16_t* __restrict__ eob16,
                                                      int32_t* __restrict__ eob, unsigned long long* runs,
                                                      int32_t* __restrict__ queue, unsigned* __restrict__ qctr) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  bool run = false;
  if (j < n_jobs && eob[j] >= 0) {
    const int t = j % group;
    bool gated = false;
    if (t != 0 && th > 0) {
      const int64_t s0 = (int64_t)csatd[j - t], sd = (int64_t)csatd[j];
      gated = (sd - s0) * 100 > s0 * th;
    }
    if (gated) {
      eob16[j] = 0;
      eob[j] = kTuGated;
    }
    run = !gated && eob16[j] > 0;
  }
  const uint64_t m = __ballot(run);
  const int lane = threadIdx.x & 63;
  if (runs && lane == 0 && m) atomicAdd(runs, (unsigned long long)__popcll(m));
  if (queue && m) {
    unsigned base = 0;
    if (lane == 0) base = atomicAdd(&qctr[0], (unsigned)__popcll(m));
    base = (unsigned)__shfl((int)base, 0, 64);
    if (run) queue[base + (unsigned)__popcll(m & ((1ull << lane) - 1))] = j;
  }
}

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// the levels are the workspace's first carve (txunit_rd_batch)
int32_t* txunit_rd_levels(void* work) { return static_cast<int32_t*>(work); }

size_t txunit_rd_workspace(int tx_size, int n_jobs) {
  if (tx_size < 0 || tx_size > 18 || n_jobs <= 0) return 0;
  const size_t np = (size_t)std::min(tx_w(tx_size), 32) * std::min(tx_h(tx_size), 32), n = (size_t)n_jobs;
  return 3 * align256(n * np * 4) + align256(n * 2) + align256(n * 8) + align256(n * sizeof(TxUnitJob)) +
         align256(n * sizeof(RdoqJob)) + align256(n * sizeof(RateJob)) + align256(n * 4) + 256;
}

template <int S>
static hipError_t launch_finish(const void* src, int stride, int bd, const uint16_t* pred, const TxUnitRdJob* jobs,
                                int n_jobs, const int32_t* tq, const int32_t* tt, const QuantRow* row,
                                const uint16_t* eob16, const uint64_t* fold, const int32_t* costs, int txs_ctx,
                                unsigned flags, uint64_t* dist, int32_t* eob, int32_t* rate, uint8_t* cul,
                                int32_t* q, uint16_t* rec, hipStream_t st) {
  constexpr int W = tx_w(S), H = tx_h(S);
  const int grid = (n_jobs + TuGeom<W, H>::JPW - 1) / TuGeom<W, H>::JPW;
  if (bd == 8)
    hipLaunchKernelGGL((k_unit_finish<W, H, uint8_t>), dim3(grid), dim3(256), 0, st, (const uint8_t*)src, stride, bd,
                       pred, jobs, n_jobs, tq, tt, row, eob16, fold, costs, txs_ctx, flags, dist, eob, rate, cul, q,
                       rec);
  else
    hipLaunchKernelGGL((k_unit_finish<W, H, uint16_t>), dim3(grid), dim3(256), 0, st, (const uint16_t*)src, stride,
                       bd, pred, jobs, n_jobs, tq, tt, row, eob16, fold, costs, txs_ctx, flags, dist, eob, rate, cul,
                       q, rec);
  return hipGetLastError();
}

hipError_t txunit_rd_batch(const void* src, int stride, int plane_w, int plane_h, int bd, int tx_size,
                           const uint16_t* pred, long pred_len, const TxUnitRdJob* jobs, int n_jobs,
                           const QuantRow* row, const uint16_t* iscan, const uint16_t* scans, const int32_t* costs,
                           unsigned flags, uint32_t cut_num, uint32_t cut_den, void* work, uint64_t* dist,
                           int32_t* eob, int32_t* rate, uint8_t* cul, int32_t* qcoeff, uint16_t* recon,
                           hipStream_t st, int32_t dq0, int32_t dq1, uint64_t* csatd, hipEvent_t* rdoq_ev,
                           size_t work_bytes, const TuSatdGate* gate, int rdoq_form) {
  if (n_jobs <= 0) return hipSuccess;
  if (tx_size < 0 || tx_size > 18) return hipErrorInvalidValue;
  const size_t np = (size_t)std::min(tx_w(tx_size), 32) * std::min(tx_h(tx_size), 32), n = (size_t)n_jobs;
  char* w = static_cast<char*>(work);
  auto carve = [&w](size_t bytes) {
    char* p = w;
    w += align256(bytes);
    return p;
  };
  TuKeep keep;
  int32_t* tq = reinterpret_cast<int32_t*>(carve(n * np * 4));  // first: txunit_rd_levels
  keep.t = reinterpret_cast<int32_t*>(carve(n * np * 4));
  keep.dq = reinterpret_cast<int32_t*>(carve(n * np * 4));
  keep.eob16 = reinterpret_cast<uint16_t*>(carve(n * 2));
  keep.fold = reinterpret_cast<uint64_t*>(carve(n * 8));
  keep.csatd = csatd;  // written by the forward pass straight into the caller's output
  if (!(flags & (kTxuRdoq | kTxuRecon | kTxuSpatial))) keep.t = keep.dq = nullptr;  // the finish does not read them
  TxUnitJob* tj = reinterpret_cast<TxUnitJob*>(carve(n * sizeof(TxUnitJob)));
  RdoqJob* rj = reinterpret_cast<RdoqJob*>(carve(n * sizeof(RdoqJob)));
  RateJob* cj = reinterpret_cast<RateJob*>(carve(n * sizeof(RateJob)));
  int32_t* queue = reinterpret_cast<int32_t*>(carve(n * 4));      // the trellis's work queue
  unsigned* qctr = reinterpret_cast<unsigned*>(carve(16));        // its length and head
  // workspace past the unit's own temporaries goes to the trellis (its lane
  // form when the batch is large enough; rdoq_batch checks the size)
  const size_t used = (size_t)(w - static_cast<char*>(work));
  void* rws = work_bytes > used ? static_cast<void*>(w) : nullptr;
  const size_t rws_bytes = work_bytes > used ? work_bytes - used : 0;
  hipLaunchKernelGGL(k_rd_prep, dim3((unsigned)((n_jobs + 255) / 256)), dim3(256), 0, st, jobs, n_jobs, (int)np, tj,
                     rj, cj, qctr);
  hipError_t e = txunit_batch(src, stride, plane_w, plane_h, bd, tx_size, pred, pred_len, tj, n_jobs, row, iscan,
                              flags & kTxuQuantB, dist, eob, cul, tq, nullptr, st, keep);
  if (e != hipSuccess) return e;
  // levels fit 16 bits when both dequantizers are at least 8: the forward
  // transform's coefficients stay below 2^17 at up to 10 bits, and
  // quantize_fp's level is at most (|c| + dq / 2) / dq
  const bool compact = (flags & kTxuRdoq) && (flags & kTxuLuma) && bd <= 10 && dq0 >= 8 && dq1 >= 8;
  // the gate also fills the trellis's work queue when the trellis drains one
  const bool queued = compact && gate && csatd && gate->group > 0 && rdoq_form != 1 && rdoq_form != 2;
  if (gate && csatd && gate->group > 0 && (queued || gate->runs || (gate->group > 1 && gate->th > 0)))
    hipLaunchKernelGGL(k_tu_satd_gate, dim3((unsigned)((n_jobs + 255) / 256)), dim3(256), 0, st, n_jobs, gate->group,
                       csatd, gate->group > 1 ? gate->th : (int64_t)0, keep.eob16, eob, gate->runs,
                       queued ? queue : nullptr, qctr);
  // rdoq_ev: two events recorded around the trellis (timing; recorded
  // also when the flags skip it)
  if (rdoq_ev && (e = hipEventRecord(rdoq_ev[0], st)) != hipSuccess) return e;
  if (flags & kTxuRdoq) {
    e = rdoq_batch(tx_size, keep.t, tq, keep.dq, keep.eob16, rj, n_jobs, costs, scans, dq0, dq1, cut_num, cut_den,
                   st, rws, rws_bytes, compact, queued ? queue : nullptr, qctr, rdoq_form);
    if (e != hipSuccess) return e;
  }
  if (rdoq_ev && (e = hipEventRecord(rdoq_ev[1], st)) != hipSuccess) return e;
  e = coeff_rate_batch(tx_size, tq, keep.eob16, cj, n_jobs, costs, scans, rate, st);
  if (e != hipSuccess) return e;
  // the tx size's entropy context: mean of its square-down and square-up sizes
  constexpr int kSqr[19] = {0, 1, 2, 3, 4, 0, 0, 1, 1, 2, 2, 3, 3, 0, 0, 1, 1, 2, 2};
  constexpr int kSqrUp[19] = {0, 1, 2, 3, 4, 1, 1, 2, 2, 3, 3, 4, 4, 2, 2, 3, 3, 4, 4};
  const int txs_ctx = (kSqr[tx_size] + kSqrUp[tx_size] + 1) >> 1;
  switch (tx_size) {
#define X(S)                                                                                                        \
  case S:                                                                                                           \
    return launch_finish<S>(src, stride, bd, pred, jobs, n_jobs, tq, keep.t, row, keep.eob16, keep.fold, costs,     \
                            txs_ctx, flags, dist, eob, rate, cul, qcoeff, recon, st);
    X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18)
#undef X
    default: return hipErrorInvalidValue;
  }
}

}  // namespace zav1
