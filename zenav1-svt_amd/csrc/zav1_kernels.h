// Internal launch entry points of the zav1 HIP kernels (device pointers,
// stream-ordered).  The public C-ABI in include/zav1_dsp.h wraps these.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zav1 {

// One quantizer row (DC = [0], AC = [1]) as built by
// rust/crates/svtav1-encoder/src/quant.rs:132-236.
struct QuantRow {
  int32_t zbin[2], round[2], quant[2], quant_shift[2];
  int32_t round_fp[2], quant_fp[2], dequant[2];
};

// Level magnitude of one coefficient magnitude `a` (iz: 0 = DC, 1 = AC) at
// log_scale ls.  fp: quant_coding.rs:80 fp_one (the INT16 clamp applies at 8
// bits only, quant.rs:335); b: quant_coding.rs:365 b_one (quant.rs:391).
__device__ __forceinline__ int32_t quant_fp_level(int32_t a, const QuantRow& t, int iz, int ls, int bd) {
  const int32_t rounding = (t.round_fp[iz] + ((1 << ls) >> 1)) >> ls;
  if (((int64_t)a << (1 + ls)) < t.dequant[iz]) return 0;
  int64_t x = (int64_t)a + rounding;
  if (bd == 8) x = x > 32767 ? 32767 : x;
  return (int32_t)((x * t.quant_fp[iz]) >> (16 - ls));
}
__device__ __forceinline__ int32_t quant_b_level(int32_t a, const QuantRow& t, int iz, int ls, int bd) {
  const int32_t zbin = (t.zbin[iz] + ((1 << ls) >> 1)) >> ls;
  const int32_t rnd = (t.round[iz] + ((1 << ls) >> 1)) >> ls;
  if (a < zbin) return 0;
  int64_t x = (int64_t)a + rnd;
  if (bd == 8) x = x > 32767 ? 32767 : x;
  return (int32_t)(((((x * t.quant[iz]) >> 16) + x) * t.quant_shift[iz]) >> (16 - ls));
}
__device__ __forceinline__ int32_t dequant_level(int32_t q, const QuantRow& t, int iz, int ls) {
  return (int32_t)(((int64_t)q * t.dequant[iz]) >> ls);
}

// The fp quantizer of one row held in registers for a fused kernel (DC at
// [0], AC at [1], rounding already shifted by log_scale).  Contract: the
// coefficient magnitude a < 2^21.  The fused kernels only transform
// residuals of bd <= 10 planes, whose coefficients are bounded by the l1
// norm of the transform rows times 1023: <= 2^17 for every size and type
// (pinned by a CPU test).  Within it, a << 3, a + rounding and q * dequant
// stay below 2^31, and the level (x * quant_fp) >> (16 - LS) equals the high
// word of (x << LS) * (quant_fp << 16): quant_fp = 2^16 / dequant <= 2^14
// (dequant >= 4), so both factors fit 32 bits.  One 32-bit high multiply
// for the level and one 24-bit multiply for the dequantized magnitude, no
// branches.  The error a - |dq| stays below 2^15 in magnitude (one step plus
// the truncation of quant_fp; pinned), so its square is one 24-bit multiply.
struct QuantFp {
  int32_t dq[2], rnd[2];
  uint32_t qf16[2];  // quant_fp << 16
};
__host__ __device__ inline QuantFp make_quant_fp(const QuantRow& t, int ls) {
  QuantFp f;
#pragma unroll
  for (int i = 0; i < 2; i++) {
    f.dq[i] = t.dequant[i];
    f.rnd[i] = (t.round_fp[i] + ((1 << ls) >> 1)) >> ls;
    f.qf16[i] = (uint32_t)t.quant_fp[i] << 16;
  }
  return f;
}
// The 24-bit multiplier forms, written out: x * y (x, y < 2^24, product
// < 2^32) and x * y + a (|x|, |y| < 2^23, result in 32 bits).  Instruction
// selection only picks them when it sees the operands' range in the same
// block, which a loop-invariant factor hoisted out of the loop hides.
__device__ __forceinline__ uint32_t mul_u24(uint32_t x, uint32_t y) {
  uint32_t r;
  asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "v"(x), "v"(y));
  return r;
}
__device__ __forceinline__ int32_t mad_i24(int32_t x, int32_t y, int32_t a) {
  int32_t r;
  asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(y), "v"(a));
  return r;
}

// Level (>= 0) of |coefficient| a; *adq receives its dequantized magnitude.
template <int LS>
__device__ __forceinline__ int32_t quant_fp_fast(int32_t a, int iz, const QuantFp& f, int bd, int32_t* adq) {
  const int32_t dqv = iz ? f.dq[1] : f.dq[0];
  int32_t xx = a + (iz ? f.rnd[1] : f.rnd[0]);
  if (bd == 8) xx = xx > 32767 ? 32767 : xx;
  const uint32_t q = __umulhi((uint32_t)xx << LS, iz ? f.qf16[1] : f.qf16[0]);
  const int32_t qq = ((a << (1 + LS)) >= dqv) ? (int32_t)q : 0;
  *adq = (int32_t)mul_u24((uint32_t)qq, (uint32_t)dqv) >> LS;  // both < 2^24, product < 2^31
  return qq;
}
// fp quantizer of the fused search kernels (log scale 0), level on the
// 24-bit multiplier when Q24: level = (xx * quant_fp) >> 16 = high word of
// (xx << S) * (quant_fp << (16 - S)), exact when both factors are below
// 2^24.  8 bits: xx <= 32767 (the INT16 clamp), S = 8.  10 bits: xx = a +
// round_fp < 2^18 (|coefficient| < 2^17, pinned), S = 6, which needs
// quant_fp < 2^14, i.e. dequant > 4: the caller takes the Q24 form only then
// (everywhere but qindex 0).  Otherwise the 32-bit high multiply, S = 0.
// The reference's dead zone (level 0 when 2a < dequant) needs no test here:
// the rows hold round_fp = floor(dequant / 2) and quant_fp =
// floor(2^16 / dequant), so 2a < dequant gives a + round_fp <= dequant - 1
// and (a + round_fp) * quant_fp < 2^16, a zero level (pinned on the CPU for
// every row).  UNI: the quantizer operands are wave-uniform (scalar
// registers feed the multiplies directly).
struct QuantFpL {
  int32_t ndq[2];
  uint32_t rs[2], qt[2];
};
__host__ __device__ constexpr int quant_fpl_shift(int bd, bool q24) { return q24 ? (bd == 8 ? 8 : 6) : 0; }
__host__ __device__ inline QuantFpL make_quant_fpl(const QuantRow& t, int bd, bool q24) {
  const int S = quant_fpl_shift(bd, q24);
  QuantFpL f;
  for (int i = 0; i < 2; i++) {
    f.ndq[i] = -t.dequant[i];
    f.rs[i] = (uint32_t)t.round_fp[i] << S;
    f.qt[i] = (uint32_t)t.quant_fp[i] << (16 - S);
  }
  return f;
}
// Up to two quantizers evaluated by one launch of the fused search kernels
// (kernel argument: scalar registers).
struct QuantFpSet {
  QuantFpL q[2];
};
struct QuantFpPair {
  QuantFp q[2];
};
template <bool UNI>
__device__ __forceinline__ uint32_t mul_hi_u24(uint32_t x, uint32_t y) {
  uint32_t r;
  if constexpr (UNI) asm("v_mul_hi_u32_u24 %0, %1, %2" : "=v"(r) : "s"(y), "v"(x));
  else asm("v_mul_hi_u32_u24 %0, %1, %2" : "=v"(r) : "v"(x), "v"(y));
  return r;
}
// High 16 bits (sign-extended) of the 48-bit product of two signed 24-bit
// values.
__device__ __forceinline__ int32_t mul_hi_i24(int32_t x, int32_t y) {
  int32_t r;
  asm("v_mul_hi_i32_i24 %0, %1, %2" : "=v"(r) : "v"(x), "v"(y));
  return r;
}
template <bool UNI>
__device__ __forceinline__ int32_t mad_i24u(int32_t x, int32_t y, int32_t a) {
  int32_t r;
  if constexpr (UNI) asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(x), "s"(y), "v"(a));
  else asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(y), "v"(a));
  return r;
}
template <int BD, bool Q24, bool UNI>
__device__ __forceinline__ int32_t quant_fpl(int32_t a, int iz, const QuantFpL& f, int32_t* err) {
  constexpr int S = quant_fpl_shift(BD, Q24);
  uint32_t xs = ((uint32_t)a << S) + (iz ? f.rs[1] : f.rs[0]);
  if constexpr (BD == 8) xs = xs < (32767u << S) ? xs : (32767u << S);
  const uint32_t qt = iz ? f.qt[1] : f.qt[0];
  uint32_t q;
  if constexpr (Q24) q = mul_hi_u24<UNI>(xs, qt);
  else q = __umulhi(xs, qt);
  *err = mad_i24u<UNI>((int32_t)q, iz ? f.ndq[1] : f.ndq[0], a);
  return (int32_t)q;
}

// The same level; *err receives a - |dequantized| (|err| < 2^15), for
// LS == 0 in one multiply-add.
template <int LS, int BD>
__device__ __forceinline__ int32_t quant_fp_err(int32_t a, int iz, const QuantFp& f, int32_t* err) {
  const int32_t dqv = iz ? f.dq[1] : f.dq[0];
  int32_t xx = a + (iz ? f.rnd[1] : f.rnd[0]);
  if constexpr (BD == 8) xx = xx > 32767 ? 32767 : xx;
  const uint32_t q = __umulhi((uint32_t)xx << LS, iz ? f.qf16[1] : f.qf16[0]);
  const int32_t qq = ((a << (1 + LS)) >= dqv) ? (int32_t)q : 0;
  if constexpr (LS == 0) *err = mad_i24(qq, -dqv, a);
  else *err = a - ((int32_t)mul_u24((uint32_t)qq, (uint32_t)dqv) >> LS);
  return qq;
}

// Plain-form quantizer parameters: kind 0 uses round (= pre-shifted _fp
// rounding), quant (= quant_fp), dequant; kind 1 uses zbin, round, quant,
// quant_shift, dequant (zbin/round pre-shifted).
struct PlainQuant {
  int32_t kind, log_scale;
  int32_t zbin[2], round[2], quant[2], quant_shift[2], dequant[2];
};
hipError_t quantize_plain(const int32_t* coeff, int32_t* qcoeff, int32_t* dqcoeff, int n,
                          const PlainQuant& p, hipStream_t st);
hipError_t quantize_hbd_plain(const int32_t* coeff, const uint16_t* scan, int n, const QuantRow& row, int ls,
                              int kind, int32_t* qcoeff, int32_t* dqcoeff, int32_t* eob, hipStream_t st);
hipError_t cdef_dist_plain(const uint16_t* src, const uint16_t* packed, int n_samples, int bw, int sub,
                           int coeff_shift, uint64_t* out, hipStream_t st);

hipError_t fwd_txfm2d_batch(const int32_t* in, int32_t* out, const uint8_t* types, int tx_size,
                            int n_jobs, hipStream_t st);
hipError_t inv_txfm2d_batch(const int32_t* in, int32_t* out, const uint8_t* types, int tx_size,
                            int n_jobs, int bd, hipStream_t st);

// kind: 0 = fp (RDOQ initial), 1 = b (dead-zone).  rows: device array of
// QuantRow indexed by qidx[job].
// One block of the RDOQ batch (same layout as zav1_rdoq_job).
struct RdoqJob {
  int64_t rdmult;
  uint32_t coeff;
  uint8_t tx_class, plane_type, txb_skip_ctx, dc_sign_ctx, sharpness;
  uint8_t pad[7];
};
static_assert(sizeof(RdoqJob) == 24, "RdoqJob layout");
// One block of the coefficient-rate batch (same layout as zav1_rate_job).
struct RateJob {
  uint32_t coeff;
  int32_t txt_rate;
  uint8_t tx_class, plane_type, txb_skip_ctx, dc_sign_ctx;
};
static_assert(sizeof(RateJob) == 12, "RateJob layout");
hipError_t coeff_rate_batch(int tx_size, const int32_t* qcoeff, const uint16_t* eobs, const RateJob* jobs,
                            int n_jobs, const int32_t* costs, const uint16_t* scans, int32_t* rate, hipStream_t st);
// The trellis: the group / wave form, or with a workspace of rdoq_workspace
// bytes and at least rdoq_lane_min(tx_size) jobs the lane form.  compact:
// the caller guarantees every job is luma and every |level| < 2^15 (the
// group form then stages one plane's tables and 16-bit levels).  queue /
// qctr (compact only): a device work queue of the jobs to walk (queue[0 ..
// qctr[0]), head qctr[1] = 0), which the persistent form drains; jobs
// outside it are left as they are.  form: -1 the defaults above; 0 the
// queued form where it applies; 1 the direct latency form; 2 the lane form
// (with the workspace).  Every form gives the same levels.
hipError_t rdoq_batch(int tx_size, const int32_t* tcoeff, int32_t* qcoeff, int32_t* dqcoeff, uint16_t* eobs,
                      const RdoqJob* jobs, int n_jobs, const int32_t* costs, const uint16_t* scans, int32_t dq0,
                      int32_t dq1, uint32_t cut_num, uint32_t cut_den, hipStream_t st, void* ws = nullptr,
                      size_t ws_bytes = 0, bool compact = false, const int32_t* queue = nullptr,
                      unsigned* qctr = nullptr, int form = -1);
size_t rdoq_workspace(int tx_size, int n_jobs);
int rdoq_lane_min(int tx_size);
void rdoq_set_lane_min(int tx_size, int n);
hipError_t quantize_batch(const int32_t* coeff, int32_t* qcoeff, int32_t* dqcoeff,
                          const QuantRow* rows, const uint8_t* qidx, int n_per_job, int n_jobs,
                          int log_scale, int bd, int kind, hipStream_t st);
// out[job] = sum (a-b)^2 (wrapping u64); b == nullptr gives sum a^2.
hipError_t sse_i32_batch(const int32_t* a, const int32_t* b, uint64_t* out, int n_per_job,
                         int n_jobs, hipStream_t st);
// Hadamard of n x n i16 residual tiles (packed) -> SATD per tile; coeff may
// be nullptr.  n in {4, 8, 16, 32}.
hipError_t satd_batch(const int16_t* res, int32_t* coeff, int32_t* satd, int n, int n_jobs,
                      hipStream_t st);

// Candidate tx types of one search launch (passed by value).
struct TypeList {
  int32_t n;
  uint8_t t[16];
};
// Processing plan of a candidate type list (built on the host): types in
// processing order with their index in the caller's list, and the distinct
// column transforms (kind 0 DCT / 1 ADST / 3 identity, up-down flip).
// Coefficient scan of a packed (<= 32x32) block of tx_size, class 0
// default / 1 raster / 2 column-major: scan[i] = raster index (host).
void build_scan(int tx_size, int cls, uint16_t* scan);

struct TxPlan {
  int32_t n_types, n_var;
  uint8_t type[16], out_k[16], var_of[16];
  uint8_t row_kind[16], scan_cls[16];  // per type: row kind 0..3, scan class
  uint8_t var_kind[4], var_ud[4];
  uint8_t var_start[5];  // types of variant v: [var_start[v], var_start[v+1]) when grouped by variant
  // the same, one dword per entry (uniform loads of whole dwords stay scalar):
  // lane_type[i] = out_k | row_kind << 8 | scan_cls << 16 (types grouped by
  // variant), lane_var[v] = column kind 0..3 | first type << 8 | end << 16
  uint32_t lane_type[16], lane_var[4];
};
// Fused tx-type search over blocks of one tx size on a 16-bit plane (8-bit
// content is widened by the caller).  iscan: device table of inverse scans,
// [tx_size][class][1024] u16.  Outputs per (block, type) pair.
// Raster grid of n blocks (nx per row) of bw x bh: xy[2i], xy[2i+1].
hipError_t grid_xy(int32_t* xy, int nx, int n, int bw, int bh, hipStream_t st);

// n_q quantizers: dev_rows[q] (device rows) and host_rows[q] (the same
// rows on the host); outputs per (q, block, type) at
// q * n_blocks * types.n + block * types.n + k.
hipError_t txsearch_batch(const uint16_t* plane, int stride, int bd, int tx_size, const int32_t* xy,
                          int n_blocks, const TypeList& types, const QuantRow* const* dev_rows,
                          const QuantRow* host_rows, int n_q, const uint16_t* iscan, uint64_t* dist, int32_t* eob,
                          uint32_t* lsum, hipStream_t st);

// One transform unit of the fused unit kernel (layout of zav1_txunit_job).
struct TxUnitJob {
  int32_t x, y;
  uint32_t pred_off;
  uint8_t tx_type, crop_w, crop_h, pad;
};
// kTxuLuma: every job is a luma unit (lets the trellis stage one plane's tables)
enum : unsigned { kTxuQuantB = 1u, kTxuSpatial = 2u, kTxuRecon = 4u, kTxuRdoq = 8u, kTxuRateEst2 = 16u, kTxuLuma = 32u };
// Intermediates a transform unit keeps for the RDOQ trellis and the
// finishing pass (all null for the plain fused unit): the packed pre-quant
// coefficients and dequantized values (min(w,32) x min(h,32) per job), the
// eob as u16 and the 64-point fold's discarded energy per job.
struct TuKeep {
  int32_t* t = nullptr;   // coefficients and dequantized values (nullable: the
  int32_t* dq = nullptr;  // eob and fold are still kept when eob16 is set)
  uint16_t* eob16 = nullptr;
  uint64_t* fold = nullptr;
  uint64_t* csatd = nullptr;  // sum of |coefficient| over the whole transform (nullable)
};
// One full-RD transform unit (layout of zav1_txunit_rd_job).
struct TxUnitRdJob {
  int32_t x, y;
  uint32_t pred_off;
  uint8_t tx_type, crop_w, crop_h, plane_type;
  uint8_t txb_skip_ctx, dc_sign_ctx, sharpness, pad;
  int32_t txt_rate;
  int64_t rdmult;
};
// A w x h unit job reads plane rows y..y+h-1, columns x..x+w-1 and
// prediction samples pred_off..pred_off+w*h-1.
__host__ __device__ inline bool unit_inside(const TxUnitJob& j, int w, int h, int pw, int ph, long pred_len) {
  return j.x >= 0 && j.y >= 0 && j.x <= pw - w && j.y <= ph - h && (long)j.pred_off + (long)w * h <= pred_len;
}
// SATD reported for a job outside the plane or the prediction buffer.
constexpr uint32_t kSatdInvalid = 0xffffffffu;
// MDS0 SATD of (unit, prediction) jobs of one w x h unit size: square
// tiles of min(w, h, 32), Hadamard + SATD per tile, summed per job.
hipError_t pred_satd_batch(const void* src, int stride, int pw, int ph, int bd, int w, int h, const uint16_t* pred,
                           long pred_len, const TxUnitJob* jobs, int n_jobs, uint32_t* satd, hipStream_t st);
// MDS0 selection: per unit, fast cost rdcost(lambda, rate, satd << 4) of
// each of its n_cands candidates (kSatdInvalid -> UINT64_MAX), the first k
// positions of the exchange sort over them into sel / sel_cost
// [unit * k + rank], and the transform-unit jobs of those candidates for
// every type of `types` into jobs[(unit * k + rank) * types.n + ti]
// (x, y = unit_xy[unit * unit_stride + 0 / 1]; prediction at
// (u * n_cands + m) * w * h, or (u * k + rank) * w * h when survivors_only).
hipError_t fast_select(const uint32_t* satd, const uint32_t* rate, uint64_t lambda, int n_units, int n_cands, int k,
                       const int32_t* unit_xy, int unit_stride, int w, int h, const TypeList& types,
                       int survivors_only, uint8_t* sel, uint64_t* sel_cost, TxUnitJob* jobs, hipStream_t st);
// Fused transform units of one tx size: residual (src - pred) -> forward ->
// fold -> quantize -> eob / cul level / frequency distortion, and, when
// reconstructing, inverse -> recon clip -> spatial distortion.
hipError_t txunit_batch(const void* src, int stride, int plane_w, int plane_h, int bd, int tx_size,
                        const uint16_t* pred, long pred_len, const TxUnitJob* jobs, int n_jobs, const QuantRow* row,
                        const uint16_t* iscan, unsigned flags, uint64_t* dist, int32_t* eob, uint8_t* cul,
                        int32_t* qcoeff, uint16_t* recon, hipStream_t st, TuKeep keep = TuKeep{});
// Full-RD transform units (the MDS3 unit): the fused unit's forward path
// (quantize_b, or quantize_fp + the RDOQ trellis with kTxuRdoq), then the
// coefficient rate (with the tx-type rate, skip cost at eob 0) and the
// distortion (frequency, or spatial after inverse + recon).  work holds
// txunit_rd_workspace(tx_size, n_jobs) bytes; scans are tx_size's three
// forward scans, costs the frame's coefficient cost tables.
size_t txunit_rd_workspace(int tx_size, int n_jobs);
// A pre-trellis gate over jobs in groups of `group` (one candidate's tx
// types in search order, the reference type first): a job whose coefficient
// SATD exceeds its group's first by more than `th` percent gets no trellis,
// rate, reconstruction or distortion (its eob becomes kTuGated).  `runs`
// (may be null) accumulates the jobs then handed to the trellis with a
// nonzero eob (measurement).
struct TuSatdGate {
  int group;
  int64_t th;
  unsigned long long* runs;
};
constexpr int32_t kTuGated = -3;
// The fixed per-device scratch of the no-allocation entry points
// (noalloc.hip): set up once per device with that device current.
hipError_t noalloc_setup(int dev);
void* noalloc_scratch(int dev);
size_t noalloc_scratch_bytes();
hipError_t txunit_rd_batch(const void* src, int stride, int plane_w, int plane_h, int bd, int tx_size,
                           const uint16_t* pred, long pred_len, const TxUnitRdJob* jobs, int n_jobs,
                           const QuantRow* row, const uint16_t* iscan, const uint16_t* scans, const int32_t* costs,
                           unsigned flags, uint32_t cut_num, uint32_t cut_den, void* work, uint64_t* dist,
                           int32_t* eob, int32_t* rate, uint8_t* cul, int32_t* qcoeff, uint16_t* recon,
                           hipStream_t st, int32_t dq0, int32_t dq1, uint64_t* csatd = nullptr,
                           hipEvent_t* rdoq_ev = nullptr, size_t work_bytes = 0,
                           const TuSatdGate* gate = nullptr, int rdoq_form = -1);
// The final levels of a txunit_rd_batch call stay in its workspace until the
// workspace is used again: job j's packed levels at txunit_rd_levels(work) +
// j * min(w, 32) * min(h, 32).
int32_t* txunit_rd_levels(void* work);
// The current device's quantizer row `row_index` (bd index * 256 + qindex),
// inverse scans and forward-scan base, as built by zav1_init (no building).
int device_tables(const QuantRow** row, const uint16_t** iscan, const uint16_t** fscan_base, int row_index);

// One intra transform unit (layout of zav1_intra_job in include/zav1_dsp.h).
struct IntraJob {
  int32_t x, y;
  uint8_t w, h, mode;
  int8_t delta;
  uint8_t fi_mode, edge_filter, filt_type, pad0;
  int16_t n_top_px, n_topright_px, n_left_px, n_bottomleft_px;
  uint32_t out_off, pad1;
};
struct CflJob {
  int32_t lx, ly;       // luma origin of the 2cw x 2ch co-located luma
  int32_t cw, ch;       // chroma block dims (4..32)
  int32_t alpha_q3;
  uint32_t dc_off, out_off, pad;
};
hipError_t intra_pred_batch(const void* plane, int stride, int bd, const IntraJob* jobs, int n_jobs, uint16_t* out,
                            hipStream_t st);
// Mode-decision units and candidates (C-ABI zav1_md_unit / zav1_md_cand).
struct MdUnit {
  int32_t x, y;
  int16_t n_top_px, n_topright_px, n_left_px, n_bottomleft_px;
  uint8_t edge_filter, filt_type, pad0, pad1;
};
struct MdCand {
  uint8_t mode;
  int8_t delta;
  uint8_t fi_mode, pad;
};
struct MdCandList {
  int32_t n;
  MdCand c[64];
};
// Fused MDS0 fast loop: per unit, every candidate predicted from recon and
// its Hadamard SATD against src; satd[unit * n + cand] (kSatdInvalid for a
// unit whose reference reads would leave the plane).
hipError_t md_fast_satd(const void* src, const void* recon, int stride, int pw, int ph, int bd, int w, int h,
                        const MdUnit* units, int n_units, const MdCandList& cl, uint32_t* satd, hipStream_t st);
// Predictions of the selected candidates: out + (unit * k + rank) * w * h
// (zeros for a unit whose reference reads would leave the plane).
hipError_t md_predict_survivors(const void* recon, int stride, int pw, int ph, int bd, int w, int h,
                                const MdUnit* units, int n_units, const MdCandList& cl, const uint8_t* sel, int k,
                                uint16_t* out, hipStream_t st);
hipError_t cfl_batch(const void* luma, int stride, int bd, const CflJob* jobs, int n_jobs, const uint16_t* dc,
                     uint16_t* out, hipStream_t st);

// Per-4x4 (mi) frame geometry for deblocking (device pointers), the
// reference's DeblockGeom (rust/crates/svtav1-encoder/src/deblock.rs:532).
struct DeblockGeomDev {
  int32_t mi_cols, mi_rows, true_w, true_h;
  const uint32_t* block_id;
  const uint8_t *bw, *bh, *tw, *th, *skip;  // skip may be nullptr (all-intra)
};
// Vertical then horizontal edges of one plane, in place (buf is u8 at
// bd 8, u16 at bd 10).
hipError_t deblock_plane(void* buf, int stride, int pw, int ph, int bd, int plane, int ss, int level_v, int level_h,
                         int sharpness, const DeblockGeomDev& g, hipStream_t st);
// Deblocking level trials: for each candidate level (host array of
// n_levels <= 64 values in 0..63, both directions) a tight copy of recon in
// work[i] is filtered and its SSE against src written to sse[i].  work
// holds n_levels planes.
hipError_t deblock_trials(const void* recon, const void* src, void* work, int pw, int ph, int bd, int plane, int ss,
                          const uint8_t* levels_host, int n_levels, int sharpness, const DeblockGeomDev& g,
                          uint64_t* sse, hipStream_t st);
// CDEF strength search over every 64x64 filter block: mse_y / mse_uv rows of
// 64 strengths per block (raster of blocks); dirvar [fb][64][2] (dir, var).
// skip: per-mi signaled skip (mi_rows x mi_cols) or nullptr.
hipError_t wiener_apply(const void* dgd, void* out, int pw, int ph, int bd, int unit_size, int ss_y, int nh, int nv,
                        const void* units, hipStream_t st);
hipError_t wiener_refine(const void* dgd, const void* src, int pw, int ph, int bd, int win, int unit_size, int ss_y,
                         int nh, int nv, int use_refinement, int max_one_step, void* units, int64_t* err,
                         hipStream_t st);
hipError_t cdef_apply(const void* in_y, const void* in_u, const void* in_v, void* out_y, void* out_u, void* out_v,
                      int width, int height, int bd, int damping, const uint8_t* fb_str, const uint8_t* skip,
                      int mi_cols, int mi_rows, uint64_t* stats, hipStream_t st);
hipError_t cdef_search(const void* rec_y, const void* rec_u, const void* rec_v, const void* src_y,
                       const void* src_u, const void* src_v, int width, int height, int bd, int damping, int sub_y,
                       int sub_uv, const uint8_t* skip, int mi_cols, int mi_rows, int32_t* dirvar, uint64_t* mse_y,
                       uint64_t* mse_uv, hipStream_t st);
// Restoration-unit statistics over an nh x nv unit grid (raster order).
hipError_t wiener_stats(const void* dgd, const void* src, int pw, int ph, int bd, int win, int unit_size, int ss_y,
                        int nh, int nv, int64_t* M, int64_t* H, hipStream_t st);
hipError_t unit_sse(const void* a, const void* b, int pw, int ph, int bd, int unit_size, int ss_y, int nh, int nv,
                    uint64_t* out, hipStream_t st);
hipError_t sgr_stats(const void* dgd, const void* src, int pw, int ph, int bd, int unit_size, int ss_y, int nh,
                     int nv, int64_t* out, hipStream_t st);

// ---- plain form: single calls of the reference's per-block kernels ----
struct PlainCdef {
  int32_t pri, sec, dir, pdamp, sdamp, rows, cols, cs, sub;
};
// residual (src - pred), recon (clip(pred + inv) to bd), SSE over n packed samples
hipError_t plain_residual(const void* src, const void* pred, int n, int hbd, int32_t* out, hipStream_t st);
hipError_t plain_recon(const void* pred, const int32_t* inv, int n, int bd, int hbd, void* out, hipStream_t st);
hipError_t plain_sse(const void* a, const void* b, int n, int hbd, uint64_t* out, hipStream_t st);
// 4 lines of one deblocking call (line-major, len samples each), in place
hipError_t plain_lpf(void* seg, int hbd, int len, int mblim, int lim, int hev, int bd, hipStream_t st);
hipError_t plain_cdef_dir(const uint16_t* blk8x8, int cs, int32_t* out_dir_var, hipStream_t st);
// tile: padded input of one block (3 rows / 8 columns border, row stride 80)
hipError_t plain_cdef_filter(const uint16_t* tile, const PlainCdef& p, int32_t* out, hipStream_t st);
// compute_stats over one explicit window [hs, he) x [vs, ve) of pw x ph planes
hipError_t plain_wiener_window(const void* dgd, const void* src, int pw, int ph, int bd, int win, int hs, int he,
                               int vs, int ve, int64_t* M, int64_t* H, hipStream_t st);
struct PlainIntra {
  int32_t kind;  // 0: predictor from cached edges (non-directional / filter intra), 1: directional from final edges
  int32_t mode, fi_mode, w, h, bd, has_above, has_left, angle, ua, ul;
};
// kind 0: row[0] = top-left, row[1 + i] = above[i], col[i] = left[i] (int32);
// kind 1: above / left int16 edge buffers of 160 samples, origin 16.
hipError_t plain_intra(const PlainIntra& p, const int32_t* row, const int32_t* col, const int16_t* above,
                       const int16_t* left, void* out, hipStream_t st);
// CfL: out = clip(pred + round(alpha * ac / 64)) over n packed samples
hipError_t plain_cfl(const int16_t* ac, const void* pred, int alpha_q3, int n, int bd, void* out, hipStream_t st);


// Partition pre-pass (pd0.hip).  Layouts of zav1_pd0_sb / zav1_pd0_node.
struct Pd0SbInfo {
  uint16_t var[85];
  uint8_t mode, max_sq, subres, pad[3];
};
struct Pd0Node {
  uint64_t cost;
  uint8_t flags, pad[7];
};
constexpr uint8_t kPd0Visited = 1, kPd0Tested = 2, kPd0Split = 8;
struct Pd0Params {
  int width, height, n_sb, bd, level, min_sq, max_tx, cap_max_block, coeff_rate_est_lvl, tile_w, tile_h, ires, nps;
  uint32_t qw, qwd;
  int32_t split_bits[4], none_bits, skip0_bits, txt_dc[3];
  uint64_t lambda;
};
hipError_t pd0_run(const void* plane, const Pd0Params& p, const QuantRow* row, const uint16_t* iscan,
                   const uint16_t* fscan, const int32_t* costs, Pd0SbInfo* info, Pd0Node* nodes, uint8_t* leaf_map,
                   hipStream_t st);

}  // namespace zav1
