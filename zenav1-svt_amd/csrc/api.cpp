// C-ABI of the zav1 gfx950 backend: runtime plumbing, quantizer rows, the
// plain per-call function table and the batched entry points.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/zav1_dsp.h"
#include "av1_qlookup_tables.h"
#include "zav1_dev.h"
#include "zav1_kernels.h"

namespace {

thread_local std::string g_err;

}  // namespace

namespace zav1 {
void set_last_error(const std::string& what) { g_err = what; }
}  // namespace zav1

namespace {

int fail(hipError_t e, const char* where) {
  g_err = std::string(where) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? ZAV1_ENOMEM : ZAV1_EIO;
}
#define ZCHECK(call)                                   \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return fail(e_, #call);      \
  } while (0)

hipStream_t S(zav1_stream_t s) { return s ? (hipStream_t)s : hipStreamPerThread; }

// Per-device table of quantizer rows: [bd index][qindex], bd index 0 = 8, 1 = 10.
struct DeviceState {
  zav1::QuantRow* rows = nullptr;
  uint16_t* iscan = nullptr;  // [19][3][1024] inverse scans
  uint16_t* fscan = nullptr;  // [19][3 * 1024]: per tx size its three scans, class c at c * n
};
std::mutex g_mu;
std::vector<DeviceState> g_dev;

void build_row_sharp(int qindex, int bd, int sharpness, zav1::QuantRow* t) {
  // rust/crates/svtav1-encoder/src/quant.rs:195 build_quant_table_bd_sharp;
  // qzbin per rust/crates/svtav1-encoder/src/bd10.rs:52; invert_quant per quant.rs:69.
  const int32_t dc = bd == 10 ? DC_QLOOKUP_10[qindex] : DC_QLOOKUP_8[qindex];
  const int32_t ac = bd == 10 ? AC_QLOOKUP_10[qindex] : AC_QLOOKUP_8[qindex];
  const int th = bd == 10 ? 592 : 148;
  int qzbin = qindex == 0 ? 64 : (dc < th ? 84 : 80);
  int qround = qindex == 0 ? 64 : 48;
  // sharpness (apply_quant_sharpness_factors, quant.rs:105): a positive
  // sharpness below qindex 31 narrows the dead zone and raises the rounding,
  // a negative one above it does the opposite, by at least 2 |sharpness| and
  // at most the qindex distance
  const int dist = qindex - 31;
  if (sharpness != 0 && ((sharpness > 0) == (dist < 0)) && dist != 0) {
    const int off = sharpness > 0 ? std::max(2 * sharpness, -dist) : std::min(-2 * sharpness, dist);
    const int dir = sharpness > 0 ? 1 : -1;
    qzbin = std::clamp(qzbin - dir * off, 1, 256);
    qround = std::clamp(qround + dir * off, 1, 256);
  }
  const int32_t q[2] = {dc, ac};
  for (int i = 0; i < 2; i++) {
    uint32_t tt = (uint32_t)q[i];
    int l = 0;
    while (tt > 1) { tt >>= 1; l++; }
    const int64_t m = 1 + ((int64_t)1 << (16 + l)) / q[i];
    t->quant[i] = (int32_t)(m - (1 << 16));
    t->quant_shift[i] = 1 << (16 - l);
    t->zbin[i] = (qzbin * q[i] + 64) >> 7;
    t->round[i] = (qround * q[i]) >> 7;
    t->quant_fp[i] = (1 << 16) / q[i];
    t->round_fp[i] = (64 * q[i]) >> 7;
    t->dequant[i] = q[i];
  }
}

void build_row(int qindex, int bd, zav1::QuantRow* t) { build_row_sharp(qindex, bd, 0, t); }

// Device pointer to the 2 x 256 quantizer-row table of the current device.
int quant_rows(const zav1::QuantRow** out) {
  int dev = 0;
  ZCHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_mu);
  if ((int)g_dev.size() <= dev) g_dev.resize(dev + 1);
  if (!g_dev[dev].rows) {
    std::vector<zav1::QuantRow> h(512);
    for (int b = 0; b < 2; b++)
      for (int q = 0; q < 256; q++) build_row(q, b ? 10 : 8, &h[b * 256 + q]);
    zav1::QuantRow* d = nullptr;
    ZCHECK(hipMalloc(&d, sizeof(zav1::QuantRow) * 512));
    ZCHECK(hipMemcpy(d, h.data(), sizeof(zav1::QuantRow) * 512, hipMemcpyHostToDevice));
    g_dev[dev].rows = d;
  }
  *out = g_dev[dev].rows;
  return ZAV1_OK;
}

int iscan_table(const uint16_t** out) {
  int dev = 0;
  ZCHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_mu);
  if ((int)g_dev.size() <= dev) g_dev.resize(dev + 1);
  if (!g_dev[dev].iscan) {
    std::vector<uint16_t> h(19 * 3 * 1024, 0), scan(1024);
    for (int s = 0; s < 19; s++)
      for (int c = 0; c < 3; c++) {
        zav1::build_scan(s, c, scan.data());
        int w = zav1::tx_w(s), hh = zav1::tx_h(s);
        const int n = (w > 32 ? 32 : w) * (hh > 32 ? 32 : hh);
        for (int i = 0; i < n; i++) h[(s * 3 + c) * 1024 + scan[i]] = (uint16_t)i;
      }
    uint16_t* d = nullptr;
    ZCHECK(hipMalloc(&d, h.size() * sizeof(uint16_t)));
    ZCHECK(hipMemcpy(d, h.data(), h.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    g_dev[dev].iscan = d;
  }
  *out = g_dev[dev].iscan;
  return ZAV1_OK;
}

// Forward scans of tx_size on the current device, as the trellis and the
// coefficient rate read them: the three classes back to back, each of the
// coded block's length.
int fscan_table(int tx_size, const uint16_t** out) {
  int dev = 0;
  ZCHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_mu);
  if ((int)g_dev.size() <= dev) g_dev.resize(dev + 1);
  if (!g_dev[dev].fscan) {
    std::vector<uint16_t> h(19 * 3 * 1024, 0);
    for (int sz = 0; sz < 19; sz++) {
      const int n = std::min(zav1::tx_w(sz), 32) * std::min(zav1::tx_h(sz), 32);
      for (int c = 0; c < 3; c++) zav1::build_scan(sz, c, h.data() + sz * 3072 + c * n);
    }
    uint16_t* d = nullptr;
    ZCHECK(hipMalloc(&d, h.size() * sizeof(uint16_t)));
    ZCHECK(hipMemcpy(d, h.data(), h.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    g_dev[dev].fscan = d;
  }
  *out = g_dev[dev].fscan + tx_size * 3072;
  return ZAV1_OK;
}

// The current device's constant tables if they were built (zav1_init
// builds them), without building them: for entry points that must not
// allocate.
int tables_ready(const zav1::QuantRow** rows, const uint16_t** iscan, const uint16_t** fscan_base) {
  int dev = 0;
  ZCHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_mu);
  if ((int)g_dev.size() <= dev || !g_dev[dev].rows || !g_dev[dev].iscan || !g_dev[dev].fscan) {
    g_err = "device not initialised: call zav1_init(device) before an entry point that does not allocate";
    return ZAV1_EINVAL;
  }
  *rows = g_dev[dev].rows;
  *iscan = g_dev[dev].iscan;
  *fscan_base = g_dev[dev].fscan;
  return ZAV1_OK;
}

}  // namespace

namespace zav1 {
int device_tables(const QuantRow** row, const uint16_t** iscan, const uint16_t** fscan_base, int row_index) {
  const QuantRow* rows = nullptr;
  const int rc = tables_ready(&rows, iscan, fscan_base);
  if (rc) return rc;
  *row = rows + row_index;
  return ZAV1_OK;
}
}  // namespace zav1

namespace {

// Per-thread device staging for the plain form (grown on demand, kept).
struct Staging {
  char* d = nullptr;
  size_t cap = 0;
  int dev = -1;
  ~Staging() {
    if (d) (void)hipFree(d);
  }
};
thread_local Staging t_stage;

int staging(size_t bytes, char** out) {
  int dev = 0;
  ZCHECK(hipGetDevice(&dev));
  if (t_stage.cap < bytes || t_stage.dev != dev) {
    if (t_stage.d) (void)hipFree(t_stage.d);
    t_stage.d = nullptr;
    t_stage.cap = 0;
    size_t cap = bytes < (1u << 20) ? (1u << 20) : bytes;
    ZCHECK(hipMalloc(&t_stage.d, cap));
    t_stage.cap = cap;
    t_stage.dev = dev;
  }
  *out = t_stage.d;
  return ZAV1_OK;
}

bool valid_tx(int tx_size, int tx_type) { return tx_size >= 0 && tx_size < 19 && tx_type >= 0 && tx_type < 16; }
bool tx_supported(int tx_size, int tx_type) {
  const zav1::TypeCfg c = zav1::type_cfg(tx_type);
  return zav1::kind_ok(c.col, zav1::tx_h(tx_size)) && zav1::kind_ok(c.row, zav1::tx_w(tx_size));
}

// ---- plain form ------------------------------------------------------------
int plain_fwd_txfm2d(const int32_t* in, int32_t* out, size_t stride, uint8_t tx_size, uint8_t tx_type) {
  if (!in || !out || !valid_tx(tx_size, tx_type)) return ZAV1_EINVAL;
  if (!tx_supported(tx_size, tx_type)) return 0;
  const int w = zav1::tx_w(tx_size), h = zav1::tx_h(tx_size), n = w * h;
  if (stride < (size_t)w) return ZAV1_EINVAL;
  std::vector<int32_t> packed((size_t)n);
  for (int r = 0; r < h; r++) std::memcpy(&packed[(size_t)r * w], in + r * stride, sizeof(int32_t) * w);
  char* st = nullptr;
  int rc = staging(sizeof(int32_t) * 2 * n + 64, &st);
  if (rc) return rc;
  int32_t* din = (int32_t*)st;
  int32_t* dout = din + n;
  uint8_t* dty = (uint8_t*)(dout + n);
  hipStream_t s = hipStreamPerThread;
  ZCHECK(hipMemcpyAsync(din, packed.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
  ZCHECK(hipMemcpyAsync(dty, &tx_type, 1, hipMemcpyHostToDevice, s));
  ZCHECK(zav1::fwd_txfm2d_batch(din, dout, dty, tx_size, 1, s));
  ZCHECK(hipMemcpyAsync(out, dout, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
  ZCHECK(hipStreamSynchronize(s));
  return 1;
}

int plain_inv_txfm2d_bd(const int32_t* in, int32_t* out, size_t stride, uint8_t tx_size, uint8_t tx_type,
                        uint8_t bd) {
  if (!in || !out || !valid_tx(tx_size, tx_type) || (bd != 8 && bd != 10)) return ZAV1_EINVAL;
  if (!tx_supported(tx_size, tx_type)) return 0;
  const int w = zav1::tx_w(tx_size), h = zav1::tx_h(tx_size), n = w * h;
  if (stride < (size_t)(w < 32 ? w : 32)) return ZAV1_EINVAL;
  std::vector<int32_t> packed((size_t)n, 0);
  for (int r = 0; r < h && r < 32; r++)
    for (int c = 0; c < w && c < 32; c++) packed[(size_t)r * w + c] = in[r * stride + c];
  char* st = nullptr;
  int rc = staging(sizeof(int32_t) * 2 * n + 64, &st);
  if (rc) return rc;
  int32_t* din = (int32_t*)st;
  int32_t* dout = din + n;
  uint8_t* dty = (uint8_t*)(dout + n);
  hipStream_t s = hipStreamPerThread;
  ZCHECK(hipMemcpyAsync(din, packed.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
  ZCHECK(hipMemcpyAsync(dty, &tx_type, 1, hipMemcpyHostToDevice, s));
  ZCHECK(zav1::inv_txfm2d_batch(din, dout, dty, tx_size, 1, bd, s));
  ZCHECK(hipMemcpyAsync(packed.data(), dout, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
  ZCHECK(hipStreamSynchronize(s));
  for (int r = 0; r < h; r++) std::memcpy(out + r * stride, &packed[(size_t)r * w], sizeof(int32_t) * w);
  return 1;
}

int plain_quant(const int32_t* coeffs, int32_t* q, int32_t* dq, size_t n, const zav1::PlainQuant& p) {
  if (!coeffs || !q || !dq) return ZAV1_EINVAL;
  if (n == 0) return ZAV1_OK;
  if (p.log_scale < 0 || p.log_scale > 2) return ZAV1_EINVAL;
  char* st = nullptr;
  int rc = staging(sizeof(int32_t) * 3 * n, &st);
  if (rc) return rc;
  int32_t* dc = (int32_t*)st;
  int32_t* dqq = dc + n;
  int32_t* ddq = dqq + n;
  hipStream_t s = hipStreamPerThread;
  ZCHECK(hipMemcpyAsync(dc, coeffs, sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
  ZCHECK(zav1::quantize_plain(dc, dqq, ddq, (int)n, p, s));
  ZCHECK(hipMemcpyAsync(q, dqq, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
  ZCHECK(hipMemcpyAsync(dq, ddq, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
  ZCHECK(hipStreamSynchronize(s));
  return ZAV1_OK;
}

int plain_quantize_fp_raster(const int32_t* coeffs, int32_t* qcoeff, int32_t* dqcoeff, size_t n,
                             const int32_t rounding[2], const int32_t quant_fp[2], const int32_t dequant[2],
                             int32_t log_scale) {
  if (!rounding || !quant_fp || !dequant) return ZAV1_EINVAL;
  zav1::PlainQuant p{};
  p.kind = 0;
  p.log_scale = log_scale;
  for (int i = 0; i < 2; i++) {
    p.round[i] = rounding[i];
    p.quant[i] = quant_fp[i];
    p.dequant[i] = dequant[i];
  }
  return plain_quant(coeffs, qcoeff, dqcoeff, n, p);
}

int plain_quantize_b_raster(const int32_t* coeffs, int32_t* qcoeff, int32_t* dqcoeff, size_t n,
                            const int32_t zbins[2], const int32_t round[2], const int32_t quant[2],
                            const int32_t quant_shift[2], const int32_t dequant[2], int32_t log_scale) {
  if (!zbins || !round || !quant || !quant_shift || !dequant) return ZAV1_EINVAL;
  zav1::PlainQuant p{};
  p.kind = 1;
  p.log_scale = log_scale;
  for (int i = 0; i < 2; i++) {
    p.zbin[i] = zbins[i];
    p.round[i] = round[i];
    p.quant[i] = quant[i];
    p.quant_shift[i] = quant_shift[i];
    p.dequant[i] = dequant[i];
  }
  return plain_quant(coeffs, qcoeff, dqcoeff, n, p);
}

int plain_quant_hbd(const int32_t* coeffs, const uint16_t* scan, size_t n, const zav1_quant_row* t,
                    int32_t log_scale, int32_t* qcoeff, int32_t* dqcoeff, int kind) {
  if (!coeffs || !scan || !t || !qcoeff || !dqcoeff || n == 0 || n > 1024 || log_scale < 0 || log_scale > 2)
    return ZAV1_EINVAL;
  for (size_t i = 0; i < n; i++)
    if (scan[i] >= n) return ZAV1_EINVAL;
  zav1::QuantRow row;
  static_assert(sizeof(row) == sizeof(*t), "quant row layout");
  std::memcpy(&row, t, sizeof(row));
  for (int i = 0; i < 2; i++)
    if (row.dequant[i] <= 0) return ZAV1_EINVAL;
  char* st = nullptr;
  int rc = staging(sizeof(int32_t) * 3 * n + sizeof(uint16_t) * n + 64, &st);
  if (rc) return rc;
  int32_t* dc = (int32_t*)st;
  int32_t* dq = dc + n;
  int32_t* ddq = dq + n;
  int32_t* deob = ddq + n;
  uint16_t* dscan = (uint16_t*)(deob + 4);
  hipStream_t s = hipStreamPerThread;
  ZCHECK(hipMemcpyAsync(dc, coeffs, sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
  ZCHECK(hipMemcpyAsync(dscan, scan, sizeof(uint16_t) * n, hipMemcpyHostToDevice, s));
  ZCHECK(zav1::quantize_hbd_plain(dc, dscan, (int)n, row, log_scale, kind, dq, ddq, deob, s));
  int32_t eob = 0;
  ZCHECK(hipMemcpyAsync(qcoeff, dq, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
  ZCHECK(hipMemcpyAsync(dqcoeff, ddq, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
  ZCHECK(hipMemcpyAsync(&eob, deob, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  ZCHECK(hipStreamSynchronize(s));
  return eob;
}

int plain_quantize_fp_hbd(const int32_t* coeffs, const uint16_t* scan, size_t n, const zav1_quant_row* t,
                          int32_t log_scale, int32_t* qcoeff, int32_t* dqcoeff) {
  return plain_quant_hbd(coeffs, scan, n, t, log_scale, qcoeff, dqcoeff, 0);
}

int plain_quantize_b_hbd(const int32_t* coeffs, const uint16_t* scan, size_t n, const zav1_quant_row* t,
                         int32_t log_scale, int32_t* qcoeff, int32_t* dqcoeff) {
  return plain_quant_hbd(coeffs, scan, n, t, log_scale, qcoeff, dqcoeff, 1);
}

int plain_cdef_dist(const void* plane, size_t plane_len, size_t plane_off, size_t pstride, const void* packed,
                    size_t packed_len, const uint8_t* dlist, size_t n_dlist, int bsize, int coeff_shift,
                    size_t subsampling_factor, int bd, uint64_t* out) {
  if (!plane || !packed || !out || (n_dlist && !dlist) || bsize < 0 || bsize > 3 || coeff_shift < 0 ||
      coeff_shift > 4 || (bd != 8 && bd != 10) || n_dlist > 64)
    return ZAV1_EINVAL;
  const int bw = (bsize == 3 || bsize == 2) ? 8 : 4, bh = (bsize == 3 || bsize == 1) ? 8 : 4;
  if (subsampling_factor < 1 || bh % (int)subsampling_factor) return ZAV1_EINVAL;
  if (packed_len < n_dlist * bw * bh) return ZAV1_EINVAL;
  const size_t per = (size_t)bw * bh;
  std::vector<uint16_t> src(n_dlist * per), pk(n_dlist * per);
  for (size_t b = 0; b < n_dlist; b++) {
    const size_t p0 = plane_off + (size_t)dlist[2 * b] * bh * pstride + (size_t)dlist[2 * b + 1] * bw;
    if (p0 + (bh - 1) * pstride + bw > plane_len) return ZAV1_EINVAL;
    for (int i = 0; i < bh; i++)
      for (int j = 0; j < bw; j++) {
        const size_t k = b * per + i * bw + j, pi = p0 + i * pstride + j;
        src[k] = bd > 8 ? ((const uint16_t*)plane)[pi] : ((const uint8_t*)plane)[pi];
        pk[k] = bd > 8 ? ((const uint16_t*)packed)[b * per + i * bw + j] : ((const uint8_t*)packed)[b * per + i * bw + j];
      }
  }
  if (n_dlist == 0) {
    *out = 0;
    return ZAV1_OK;
  }
  char* st = nullptr;
  int rc = staging(sizeof(uint16_t) * 2 * src.size() + 64, &st);
  if (rc) return rc;
  uint64_t* dout = (uint64_t*)st;
  uint16_t* dsrc = (uint16_t*)(st + 64);
  uint16_t* dpk = dsrc + src.size();
  hipStream_t s = hipStreamPerThread;
  ZCHECK(hipMemcpyAsync(dsrc, src.data(), sizeof(uint16_t) * src.size(), hipMemcpyHostToDevice, s));
  ZCHECK(hipMemcpyAsync(dpk, pk.data(), sizeof(uint16_t) * pk.size(), hipMemcpyHostToDevice, s));
  ZCHECK(zav1::cdef_dist_plain(dsrc, dpk, (int)src.size(), bw, (int)subsampling_factor, coeff_shift, dout, s));
  ZCHECK(hipMemcpyAsync(out, dout, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  ZCHECK(hipStreamSynchronize(s));
  return ZAV1_OK;
}

int plain_sse_i32(const int32_t* a, const int32_t* b, size_t n, uint64_t* out) {
  if (!a || !out) return ZAV1_EINVAL;
  char* st = nullptr;
  int rc = staging(sizeof(int32_t) * 2 * n + 64, &st);
  if (rc) return rc;
  int32_t* da = (int32_t*)st;
  int32_t* db = b ? da + n : nullptr;
  uint64_t* dout = (uint64_t*)(((uintptr_t)(da + 2 * n) + 7) & ~(uintptr_t)7);
  hipStream_t s = hipStreamPerThread;
  ZCHECK(hipMemcpyAsync(da, a, sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
  if (b) ZCHECK(hipMemcpyAsync(db, b, sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
  ZCHECK(zav1::sse_i32_batch(da, db, dout, (int)n, 1, s));
  ZCHECK(hipMemcpyAsync(out, dout, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  ZCHECK(hipStreamSynchronize(s));
  return ZAV1_OK;
}

int plain_hadamard_satd(const int16_t* src, size_t stride, int n, int32_t* coeff, int32_t* satd) {
  if (!src || !satd || (n != 4 && n != 8 && n != 16 && n != 32) || stride < (size_t)n) return ZAV1_EINVAL;
  std::vector<int16_t> packed((size_t)n * n);
  for (int r = 0; r < n; r++) std::memcpy(&packed[(size_t)r * n], src + r * stride, sizeof(int16_t) * n);
  char* st = nullptr;
  const size_t nn = (size_t)n * n;
  int rc = staging(sizeof(int16_t) * nn + sizeof(int32_t) * (nn + 2) + 64, &st);
  if (rc) return rc;
  int16_t* dres = (int16_t*)st;
  int32_t* dco = (int32_t*)(st + ((sizeof(int16_t) * nn + 15) & ~(size_t)15));
  int32_t* dsatd = dco + nn;
  hipStream_t s = hipStreamPerThread;
  ZCHECK(hipMemcpyAsync(dres, packed.data(), sizeof(int16_t) * nn, hipMemcpyHostToDevice, s));
  ZCHECK(zav1::satd_batch(dres, dco, dsatd, n, 1, s));
  if (coeff) ZCHECK(hipMemcpyAsync(coeff, dco, sizeof(int32_t) * nn, hipMemcpyDeviceToHost, s));
  ZCHECK(hipMemcpyAsync(satd, dsatd, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  ZCHECK(hipStreamSynchronize(s));
  return ZAV1_OK;
}

// ---- plain form, pixel-domain and loop-filter entries --------------------
// Staging layout helper: consecutive 256-byte aligned regions of one buffer.
struct Carve {
  char* base;
  size_t off = 0;
  template <typename T>
  T* take(size_t count) {
    T* p = (T*)(base + off);
    off += (sizeof(T) * count + 255) & ~(size_t)255;
    return p;
  }
};
constexpr size_t kCarveMax = 1u << 20;

// Packs a strided w x h block of `es`-byte samples into dst.
void pack2d(void* dst, const void* src, size_t stride, size_t w, size_t h, size_t es) {
  for (size_t r = 0; r < h; r++) std::memcpy((char*)dst + r * w * es, (const char*)src + r * stride * es, w * es);
}
void unpack2d(void* dst, size_t stride, const void* src, size_t w, size_t h, size_t es) {
  for (size_t r = 0; r < h; r++) std::memcpy((char*)dst + r * stride * es, (const char*)src + r * w * es, w * es);
}
bool plain_dims(size_t w, size_t h) { return w >= 1 && h >= 1 && w <= 128 && h <= 128; }

int residual_common(const void* src, size_t ss, const void* pred, size_t ps, size_t w, size_t h, int32_t* out,
                    int hbd) {
  if (!src || !pred || !out || !plain_dims(w, h) || ss < w || ps < w) return ZAV1_EINVAL;
  const size_t n = w * h, es = hbd ? 2 : 1;
  std::vector<char> a(n * es), b(n * es);
  pack2d(a.data(), src, ss, w, h, es);
  pack2d(b.data(), pred, ps, w, h, es);
  char* st = nullptr;
  int rc = staging(kCarveMax, &st);
  if (rc) return rc;
  Carve c{st};
  char* da = c.take<char>(n * es);
  char* db = c.take<char>(n * es);
  int32_t* dout = c.take<int32_t>(n);
  hipStream_t s = hipStreamPerThread;
  ZCHECK(hipMemcpyAsync(da, a.data(), n * es, hipMemcpyHostToDevice, s));
  ZCHECK(hipMemcpyAsync(db, b.data(), n * es, hipMemcpyHostToDevice, s));
  ZCHECK(zav1::plain_residual(da, db, (int)n, hbd, dout, s));
  ZCHECK(hipMemcpyAsync(out, dout, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
  ZCHECK(hipStreamSynchronize(s));
  return ZAV1_OK;
}
int plain_residual_i32(const uint8_t* src, size_t ss, const uint8_t* pred, size_t ps, size_t w, size_t h,
                       int32_t* out) {
  return residual_common(src, ss, pred, ps, w, h, out, 0);
}
int plain_residual_i32_u16(const uint16_t* src, size_t ss, const uint16_t* pred, size_t ps, size_t w, size_t h,
                           int32_t* out) {
  return residual_common(src, ss, pred, ps, w, h, out, 1);
}

int recon_common(const void* pred, size_t ps, const int32_t* inv, size_t w, size_t h, void* out, int bd, int hbd) {
  if (!pred || !inv || !out || !plain_dims(w, h) || ps < w || (bd != 8 && bd != 10) || (!hbd && bd != 8))
    return ZAV1_EINVAL;
  const size_t n = w * h, es = hbd ? 2 : 1;
  std::vector<char> a(n * es), o(n * es);
  pack2d(a.data(), pred, ps, w, h, es);
  char* st = nullptr;
  int rc = staging(kCarveMax, &st);
  if (rc) return rc;
  Carve c{st};
  char* dp = c.take<char>(n * es);
  int32_t* di = c.take<int32_t>(n);
  char* dout = c.take<char>(n * es);
  hipStream_t s = hipStreamPerThread;
  ZCHECK(hipMemcpyAsync(dp, a.data(), n * es, hipMemcpyHostToDevice, s));
  ZCHECK(hipMemcpyAsync(di, inv, sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
  ZCHECK(zav1::plain_recon(dp, di, (int)n, bd, hbd, dout, s));
  ZCHECK(hipMemcpyAsync(out, dout, n * es, hipMemcpyDeviceToHost, s));
  ZCHECK(hipStreamSynchronize(s));
  return ZAV1_OK;
}
int plain_recon_add_clamp(const uint8_t* pred, size_t ps, const int32_t* inv, size_t w, size_t h, uint8_t* out) {
  return recon_common(pred, ps, inv, w, h, out, 8, 0);
}
int plain_recon_add_clamp_u16(const uint16_t* pred, size_t ps, const int32_t* inv, size_t w, size_t h, uint16_t* out,
                              int bd) {
  return recon_common(pred, ps, inv, w, h, out, bd, 1);
}

int sse_common(const void* a, size_t as, const void* b, size_t bs, size_t w, size_t h, uint64_t* out, int hbd) {
  if (!a || !b || !out || !plain_dims(w, h) || as < w || bs < w) return ZAV1_EINVAL;
  const size_t n = w * h, es = hbd ? 2 : 1;
  std::vector<char> pa(n * es), pb(n * es);
  pack2d(pa.data(), a, as, w, h, es);
  pack2d(pb.data(), b, bs, w, h, es);
  char* st = nullptr;
  int rc = staging(kCarveMax, &st);
  if (rc) return rc;
  Carve c{st};
  char* da = c.take<char>(n * es);
  char* db = c.take<char>(n * es);
  uint64_t* dout = c.take<uint64_t>(1);
  hipStream_t s = hipStreamPerThread;
  ZCHECK(hipMemcpyAsync(da, pa.data(), n * es, hipMemcpyHostToDevice, s));
  ZCHECK(hipMemcpyAsync(db, pb.data(), n * es, hipMemcpyHostToDevice, s));
  ZCHECK(zav1::plain_sse(da, db, (int)n, hbd, dout, s));
  ZCHECK(hipMemcpyAsync(out, dout, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  ZCHECK(hipStreamSynchronize(s));
  return ZAV1_OK;
}
int plain_sse_u8(const uint8_t* a, size_t as, const uint8_t* b, size_t bs, size_t w, size_t h, uint64_t* out) {
  return sse_common(a, as, b, bs, w, h, out, 0);
}
int plain_sse_u16(const uint16_t* a, size_t as, const uint16_t* b, size_t bs, size_t w, size_t h, uint64_t* out) {
  return sse_common(a, as, b, bs, w, h, out, 1);
}

bool pred_dims(int w, int h) {
  auto ok = [](int v) { return v == 4 || v == 8 || v == 16 || v == 32 || v == 64; };
  return ok(w) && ok(h);
}

// Runs one plain prediction and writes the w x h result into dst (stride).
int run_plain_intra(const zav1::PlainIntra& p, const std::vector<int32_t>& row, const std::vector<int32_t>& col,
                    const std::vector<int16_t>& above, const std::vector<int16_t>& left, void* dst,
                    size_t dst_stride) {
  const size_t n = (size_t)p.w * p.h, es = p.bd > 8 ? 2 : 1;
  char* st = nullptr;
  int rc = staging(kCarveMax, &st);
  if (rc) return rc;
  Carve c{st};
  int32_t* drow = c.take<int32_t>(160);
  int32_t* dcol = c.take<int32_t>(160);
  int16_t* dab = c.take<int16_t>(160);
  int16_t* dlf = c.take<int16_t>(160);
  char* dout = c.take<char>(n * es);
  hipStream_t s = hipStreamPerThread;
  if (!row.empty()) ZCHECK(hipMemcpyAsync(drow, row.data(), sizeof(int32_t) * row.size(), hipMemcpyHostToDevice, s));
  if (!col.empty()) ZCHECK(hipMemcpyAsync(dcol, col.data(), sizeof(int32_t) * col.size(), hipMemcpyHostToDevice, s));
  if (!above.empty())
    ZCHECK(hipMemcpyAsync(dab, above.data(), sizeof(int16_t) * above.size(), hipMemcpyHostToDevice, s));
  if (!left.empty()) ZCHECK(hipMemcpyAsync(dlf, left.data(), sizeof(int16_t) * left.size(), hipMemcpyHostToDevice, s));
  ZCHECK(zav1::plain_intra(p, drow, dcol, dab, dlf, dout, s));
  std::vector<char> out(n * es);
  ZCHECK(hipMemcpyAsync(out.data(), dout, n * es, hipMemcpyDeviceToHost, s));
  ZCHECK(hipStreamSynchronize(s));
  unpack2d(dst, dst_stride, out.data(), p.w, p.h, es);
  return ZAV1_OK;
}

int32_t sample_at(const void* p, int hbd, size_t i) {
  return hbd ? ((const uint16_t*)p)[i] : ((const uint8_t*)p)[i];
}

// predict_dc / _v / _h / _smooth / _smooth_v / _smooth_h / _paeth
// (intra_pred.rs:18-154, hbd.rs:169-361): mode 0, 1, 2, 9, 10, 11, 12.
int plain_intra_pred(int mode, void* dst, size_t dst_stride, const void* above, const void* left, int top_left, int w,
                     int h, int has_above, int has_left, int bd) {
  if (!dst || !pred_dims(w, h) || dst_stride < (size_t)w || (bd != 8 && bd != 10)) return ZAV1_EINVAL;
  if (!(mode == 0 || mode == 1 || mode == 2 || (mode >= 9 && mode <= 12))) return ZAV1_EINVAL;
  const bool dc = mode == 0;
  const bool need_a = dc ? has_above != 0 : mode != 2, need_l = dc ? has_left != 0 : mode != 1;
  if ((need_a && !above) || (need_l && !left)) return ZAV1_EINVAL;
  const int hbd = bd > 8;
  std::vector<int32_t> row(1 + w, 0), col(h, 0);
  row[0] = top_left;
  for (int i = 0; i < w && above; i++) row[1 + i] = sample_at(above, hbd, i);
  for (int i = 0; i < h && left; i++) col[i] = sample_at(left, hbd, i);
  zav1::PlainIntra p{};
  p.kind = 0;
  p.mode = mode;
  p.fi_mode = 5;
  p.w = w;
  p.h = h;
  p.bd = bd;
  p.has_above = dc ? (has_above != 0) : 1;
  p.has_left = dc ? (has_left != 0) : 1;
  return run_plain_intra(p, row, col, {}, {}, dst, dst_stride);
}

// predict_filter_intra (intra_pred.rs:988, hbd.rs:784): above[0] = top-left,
// above[1..=w] the row above; left[0..h).
int plain_filter_intra(void* dst, size_t dst_stride, const void* above, const void* left, int w, int h, int mode,
                       int bd) {
  if (!dst || !above || !left || !pred_dims(w, h) || w > 32 || h > 32 || dst_stride < (size_t)w || mode < 0 ||
      mode > 4 || (bd != 8 && bd != 10))
    return ZAV1_EINVAL;
  const int hbd = bd > 8;
  std::vector<int32_t> row(1 + w), col(h);
  for (int i = 0; i <= w; i++) row[i] = sample_at(above, hbd, i);
  for (int i = 0; i < h; i++) col[i] = sample_at(left, hbd, i);
  zav1::PlainIntra p{};
  p.kind = 0;
  p.mode = 0;
  p.fi_mode = mode;
  p.w = w;
  p.h = h;
  p.bd = bd;
  p.has_above = 1;
  p.has_left = 1;
  return run_plain_intra(p, row, col, {}, {}, dst, dst_stride);
}

// dr_predictor_edged (intra_pred.rs:810, hbd.rs:700): above / left edge
// buffers of edge_len samples, `origin` = index of sample 0 of the row /
// column (the corner at origin - 1).
int plain_dr_predictor_edged(void* dst, size_t dst_stride, const void* above, const void* left, size_t edge_len,
                             size_t origin, int upsample_above, int upsample_left, int w, int h, int angle, int bd) {
  if (!dst || !above || !left || !pred_dims(w, h) || dst_stride < (size_t)w || origin >= edge_len ||
      (bd != 8 && bd != 10))
    return ZAV1_EINVAL;
  // the angles the encoder forms: a directional mode's nominal angle +- 3, 6, 9
  // degrees; upsampled edges exist only for w + h <= 16 (use_intra_edge_upsample)
  static const int kNominal[8] = {45, 67, 90, 113, 135, 157, 180, 203};
  bool angle_ok = false;
  for (int m = 0; m < 8; m++)
    for (int d = -3; d <= 3; d++) angle_ok |= angle == kNominal[m] + 3 * d;
  if (!angle_ok || ((upsample_above || upsample_left) && w + h > 16)) return ZAV1_EINVAL;
  const int hbd = bd > 8;
  std::vector<int16_t> ab(160, 0), lf(160, 0);
  for (int k = 0; k < 160; k++) {
    const long i = (long)origin - 16 + k;
    if (i >= 0 && i < (long)edge_len) {
      ab[k] = (int16_t)sample_at(above, hbd, (size_t)i);
      lf[k] = (int16_t)sample_at(left, hbd, (size_t)i);
    }
  }
  zav1::PlainIntra p{};
  p.kind = 1;
  p.w = w;
  p.h = h;
  p.bd = bd;
  p.angle = angle;
  p.ua = upsample_above ? 1 : 0;
  p.ul = upsample_left ? 1 : 0;
  return run_plain_intra(p, {}, {}, ab, lf, dst, dst_stride);
}

// cfl_predict_lbd (intra_pred.rs:1137) / cfl_predict_hbd (hbd.rs:889): ac in
// Q3 with row stride 32 (CFL_BUF_LINE).
int plain_cfl_predict(const int16_t* ac_q3, const void* pred, size_t pred_stride, void* dst, size_t dst_stride,
                      int alpha_q3, int w, int h, int bd) {
  if (!ac_q3 || !pred || !dst || w < 4 || h < 4 || w > 32 || h > 32 || pred_stride < (size_t)w ||
      dst_stride < (size_t)w || (bd != 8 && bd != 10))
    return ZAV1_EINVAL;
  const size_t n = (size_t)w * h, es = bd > 8 ? 2 : 1;
  std::vector<int16_t> ac(n);
  for (int r = 0; r < h; r++) std::memcpy(&ac[(size_t)r * w], ac_q3 + r * 32, sizeof(int16_t) * w);
  std::vector<char> pp(n * es), out(n * es);
  pack2d(pp.data(), pred, pred_stride, w, h, es);
  char* st = nullptr;
  int rc = staging(kCarveMax, &st);
  if (rc) return rc;
  Carve c{st};
  int16_t* dac = c.take<int16_t>(n);
  char* dp = c.take<char>(n * es);
  char* dout = c.take<char>(n * es);
  hipStream_t s = hipStreamPerThread;
  ZCHECK(hipMemcpyAsync(dac, ac.data(), sizeof(int16_t) * n, hipMemcpyHostToDevice, s));
  ZCHECK(hipMemcpyAsync(dp, pp.data(), n * es, hipMemcpyHostToDevice, s));
  ZCHECK(zav1::plain_cfl(dac, dp, alpha_q3, (int)n, bd, dout, s));
  ZCHECK(hipMemcpyAsync(out.data(), dout, n * es, hipMemcpyDeviceToHost, s));
  ZCHECK(hipStreamSynchronize(s));
  unpack2d(dst, dst_stride, out.data(), w, h, es);
  return ZAV1_OK;
}

// lpf_{horizontal,vertical}_{4,6,8,14} (loop_filter.rs:307-405) and the
// _hbd twins (hbd.rs:1335-1433): the 4 lines at sample offset `off`.
int plain_lpf(void* buf, size_t buf_len, size_t off, size_t pitch, int vertical, int taps, uint8_t mblim,
              uint8_t lim, uint8_t hev_thr, int bd) {
  if (!buf || (taps != 4 && taps != 6 && taps != 8 && taps != 14) || (bd != 8 && bd != 10) || pitch == 0)
    return ZAV1_EINVAL;
  const int hbd = bd > 8;
  const size_t es = hbd ? 2 : 1;
  const long half = taps / 2;
  std::vector<long> idx(4 * (size_t)taps);
  for (int i = 0; i < 4; i++) {
    const long base = vertical ? (long)off + (long)i * (long)pitch : (long)off + i;
    const long step = vertical ? 1 : (long)pitch;
    for (int k = 0; k < taps; k++) {
      const long j = base + (k - half) * step;
      if (j < 0 || j >= (long)buf_len) return ZAV1_EINVAL;
      idx[(size_t)i * taps + k] = j;
    }
  }
  std::vector<char> seg(idx.size() * es);
  for (size_t k = 0; k < idx.size(); k++) std::memcpy(&seg[k * es], (char*)buf + idx[k] * es, es);
  char* st = nullptr;
  int rc = staging(kCarveMax, &st);
  if (rc) return rc;
  hipStream_t s = hipStreamPerThread;
  ZCHECK(hipMemcpyAsync(st, seg.data(), seg.size(), hipMemcpyHostToDevice, s));
  ZCHECK(zav1::plain_lpf(st, hbd, taps, mblim, lim, hev_thr, bd, s));
  ZCHECK(hipMemcpyAsync(seg.data(), st, seg.size(), hipMemcpyDeviceToHost, s));
  ZCHECK(hipStreamSynchronize(s));
  for (size_t k = 0; k < idx.size(); k++) std::memcpy((char*)buf + idx[k] * es, &seg[k * es], es);
  return ZAV1_OK;
}

// cdef_find_dir (cdef.rs:143): 8x8 of u16 samples; returns the direction
// 0..7 (negative on error), *var the variance.
int plain_cdef_find_dir(const uint16_t* img, size_t stride, int coeff_shift, int32_t* var) {
  if (!img || !var || stride < 8 || coeff_shift < 0 || coeff_shift > 4) return ZAV1_EINVAL;
  uint16_t blk[64];
  for (int r = 0; r < 8; r++) std::memcpy(blk + r * 8, img + r * stride, 16);
  char* st = nullptr;
  int rc = staging(kCarveMax, &st);
  if (rc) return rc;
  Carve c{st};
  uint16_t* db = c.take<uint16_t>(64);
  int32_t* dout = c.take<int32_t>(2);
  hipStream_t s = hipStreamPerThread;
  int32_t out[2];
  ZCHECK(hipMemcpyAsync(db, blk, sizeof(blk), hipMemcpyHostToDevice, s));
  ZCHECK(zav1::plain_cdef_dir(db, coeff_shift, dout, s));
  ZCHECK(hipMemcpyAsync(out, dout, sizeof(out), hipMemcpyDeviceToHost, s));
  ZCHECK(hipStreamSynchronize(s));
  *var = out[1];
  return out[0];
}

// cdef_filter_block (cdef.rs:236) / cdef_filter_block_hbd (hbd.rs:1533):
// `in` is the padded u16 source with row stride 144 (CDEF_BSTRIDE), the
// block at in + ioff; dst u8 at bd 8, u16 at bd 10.  bsize 0 4x4, 1 4x8,
// 2 8x4, 3 8x8; rows stepped by subsampling_factor.
int plain_cdef_filter_block(void* dst, size_t dst_stride, const uint16_t* in, size_t in_len, size_t ioff,
                            int pri_strength, int sec_strength, int dir, int pri_damping, int sec_damping, int bsize,
                            int coeff_shift, size_t subsampling_factor, int bd) {
  constexpr long kBS = 144, kTS = 80, kVB = 3, kHB = 8;
  if (!dst || !in || bsize < 0 || bsize > 3 || dir < 0 || dir > 7 || pri_strength < 0 || sec_strength < 0 ||
      subsampling_factor < 1 || coeff_shift < 0 || coeff_shift > 4 || (bd != 8 && bd != 10))
    return ZAV1_EINVAL;
  if ((pri_strength && pri_damping < 0) || (sec_strength && sec_damping < 0)) return ZAV1_EINVAL;
  const int rows = (bsize == 3 || bsize == 1) ? 8 : 4, cols = (bsize == 3 || bsize == 2) ? 8 : 4;
  if (dst_stride < (size_t)cols) return ZAV1_EINVAL;
  const long lo = (long)ioff - kVB * kBS - kHB, hi = (long)ioff + (rows + kVB - 1) * kBS + cols + kHB - 1;
  if (lo < 0 || hi >= (long)in_len) return ZAV1_EINVAL;
  std::vector<uint16_t> tile((size_t)(rows + 2 * kVB) * kTS, 0);
  for (long r = 0; r < rows + 2 * kVB; r++)
    std::memcpy(&tile[(size_t)(r * kTS)], in + ioff + (r - kVB) * kBS - kHB, sizeof(uint16_t) * (cols + 2 * kHB));
  zav1::PlainCdef p{};
  p.pri = pri_strength;
  p.sec = sec_strength;
  p.dir = dir;
  p.pdamp = pri_damping;
  p.sdamp = sec_damping;
  p.rows = rows;
  p.cols = cols;
  p.cs = coeff_shift;
  p.sub = (int)subsampling_factor;
  char* st = nullptr;
  int rc = staging(kCarveMax, &st);
  if (rc) return rc;
  Carve c{st};
  uint16_t* dt = c.take<uint16_t>(tile.size());
  int32_t* dout = c.take<int32_t>(64);
  int32_t out[64];
  hipStream_t s = hipStreamPerThread;
  ZCHECK(hipMemcpyAsync(dt, tile.data(), sizeof(uint16_t) * tile.size(), hipMemcpyHostToDevice, s));
  ZCHECK(zav1::plain_cdef_filter(dt, p, dout, s));
  ZCHECK(hipMemcpyAsync(out, dout, sizeof(int32_t) * rows * cols, hipMemcpyDeviceToHost, s));
  ZCHECK(hipStreamSynchronize(s));
  for (int i = 0; i < rows; i += (int)subsampling_factor)
    for (int j = 0; j < cols; j++) {
      if (bd > 8) ((uint16_t*)dst)[i * dst_stride + j] = (uint16_t)out[i * cols + j];
      else ((uint8_t*)dst)[i * dst_stride + j] = (uint8_t)out[i * cols + j];
    }
  return ZAV1_OK;
}

// compute_stats (restoration.rs:232) / compute_stats_hbd (:1738): dgd needs
// wiener_win / 2 samples of margin around the window (as the reference's
// search provides); samples u8 at bd 8, u16 at bd 10.
int plain_compute_stats(int wiener_win, const void* dgd, size_t dgd_len, size_t dgd_origin, size_t dgd_stride,
                        const void* src, size_t src_len, size_t src_origin, size_t src_stride, int h_start, int h_end,
                        int v_start, int v_end, int64_t* M, int64_t* H, int bd) {
  if (!dgd || !src || !M || !H || (wiener_win != 5 && wiener_win != 7) || (bd != 8 && bd != 10) ||
      h_end <= h_start || v_end <= v_start || h_end - h_start > 384 || v_end - v_start > 384)
    return ZAV1_EINVAL;
  const int half = wiener_win / 2, uw = h_end - h_start, uh = v_end - v_start;
  const int pw = uw + 2 * half, ph = uh + 2 * half;
  const size_t es = bd > 8 ? 2 : 1;
  auto at = [](size_t origin, size_t stride, long r, long c) { return (long)origin + r * (long)stride + c; };
  if (at(dgd_origin, dgd_stride, v_start - half, h_start - half) < 0 ||
      at(dgd_origin, dgd_stride, v_end + half - 1, h_end + half - 1) >= (long)dgd_len ||
      at(src_origin, src_stride, v_start, h_start) < 0 ||
      at(src_origin, src_stride, v_end - 1, h_end - 1) >= (long)src_len)
    return ZAV1_EINVAL;
  std::vector<char> pd((size_t)pw * ph * es), psrc((size_t)pw * ph * es, 0);
  for (int r = 0; r < ph; r++)
    std::memcpy(&pd[(size_t)r * pw * es],
                (const char*)dgd + at(dgd_origin, dgd_stride, v_start - half + r, h_start - half) * es, pw * es);
  for (int r = 0; r < uh; r++)
    std::memcpy(&psrc[((size_t)(r + half) * pw + half) * es],
                (const char*)src + at(src_origin, src_stride, v_start + r, h_start) * es, uw * es);
  const int w2 = wiener_win * wiener_win;
  char* st = nullptr;
  int rc = staging(kCarveMax + 2 * pd.size(), &st);
  if (rc) return rc;
  Carve c{st};
  char* dd = c.take<char>(pd.size());
  char* ds = c.take<char>(psrc.size());
  int64_t* dm = c.take<int64_t>(w2);
  int64_t* dh = c.take<int64_t>((size_t)w2 * w2);
  hipStream_t s = hipStreamPerThread;
  ZCHECK(hipMemcpyAsync(dd, pd.data(), pd.size(), hipMemcpyHostToDevice, s));
  ZCHECK(hipMemcpyAsync(ds, psrc.data(), psrc.size(), hipMemcpyHostToDevice, s));
  ZCHECK(zav1::plain_wiener_window(dd, ds, pw, ph, bd, wiener_win, half, half + uw, half, half + uh, dm, dh, s));
  ZCHECK(hipMemcpyAsync(M, dm, sizeof(int64_t) * w2, hipMemcpyDeviceToHost, s));
  ZCHECK(hipMemcpyAsync(H, dh, sizeof(int64_t) * w2 * w2, hipMemcpyDeviceToHost, s));
  ZCHECK(hipStreamSynchronize(s));
  return ZAV1_OK;
}

const zav1_dsp_v1 g_table = {
    ZAV1_ABI_VERSION,
    1u,
    plain_fwd_txfm2d,
    plain_inv_txfm2d_bd,
    plain_quantize_fp_raster,
    plain_quantize_b_raster,
    plain_sse_i32,
    plain_hadamard_satd,
    plain_residual_i32,
    plain_residual_i32_u16,
    plain_recon_add_clamp,
    plain_recon_add_clamp_u16,
    plain_sse_u8,
    plain_sse_u16,
    plain_intra_pred,
    plain_dr_predictor_edged,
    plain_filter_intra,
    plain_cfl_predict,
    plain_lpf,
    plain_cdef_find_dir,
    plain_cdef_filter_block,
    plain_compute_stats,
    plain_quantize_fp_hbd,
    plain_quantize_b_hbd,
    plain_cdef_dist,
};

}  // namespace

extern "C" {

int zav1_init(int device) {
  int n = 0;
  ZCHECK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return ZAV1_EINVAL;
  ZCHECK(hipSetDevice(device));
  ZCHECK(hipFree(nullptr));
  // the device's constant tables and the address of its fixed scratch, so
  // that entry points that must not allocate find them ready
  const zav1::QuantRow* rows = nullptr;
  const uint16_t* tab = nullptr;
  int rc = quant_rows(&rows);
  if (!rc) rc = iscan_table(&tab);
  if (!rc) rc = fscan_table(0, &tab);
  if (rc) return rc;
  ZCHECK(zav1::noalloc_setup(device));
  return ZAV1_OK;
}

int zav1_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int zav1_malloc(void** p, size_t bytes) {
  if (!p) return ZAV1_EINVAL;
  ZCHECK(hipMalloc(p, bytes ? bytes : 1));
  return ZAV1_OK;
}
int zav1_free(void* p) {
  ZCHECK(hipFree(p));
  return ZAV1_OK;
}
int zav1_host_alloc(void** p, size_t bytes) {
  if (!p) return ZAV1_EINVAL;
  ZCHECK(hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault));
  return ZAV1_OK;
}
int zav1_host_free(void* p) {
  ZCHECK(hipHostFree(p));
  return ZAV1_OK;
}
int zav1_memcpy_h2d(void* dst, const void* src, size_t bytes, zav1_stream_t s) {
  ZCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(s)));
  return ZAV1_OK;
}
int zav1_memcpy_d2h(void* dst, const void* src, size_t bytes, zav1_stream_t s) {
  ZCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(s)));
  ZCHECK(hipStreamSynchronize(S(s)));
  return ZAV1_OK;
}
int zav1_memcpy_d2d(void* dst, const void* src, size_t bytes, zav1_stream_t s) {
  ZCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(s)));
  return ZAV1_OK;
}
int zav1_memset(void* dst, int value, size_t bytes, zav1_stream_t s) {
  ZCHECK(hipMemsetAsync(dst, value, bytes, S(s)));
  return ZAV1_OK;
}
int zav1_stream_create(zav1_stream_t* s) {
  if (!s) return ZAV1_EINVAL;
  hipStream_t h = nullptr;
  ZCHECK(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));
  *s = (zav1_stream_t)h;
  return ZAV1_OK;
}
int zav1_stream_destroy(zav1_stream_t s) {
  if (s) ZCHECK(hipStreamDestroy((hipStream_t)s));
  return ZAV1_OK;
}
int zav1_stream_sync(zav1_stream_t s) {
  ZCHECK(hipStreamSynchronize(S(s)));
  return ZAV1_OK;
}
int zav1_device_sync(void) {
  ZCHECK(hipDeviceSynchronize());
  return ZAV1_OK;
}
int zav1_event_create(void** ev) {
  if (!ev) return ZAV1_EINVAL;
  hipEvent_t e;
  ZCHECK(hipEventCreate(&e));
  *ev = (void*)e;
  return ZAV1_OK;
}
int zav1_event_destroy(void* ev) {
  ZCHECK(hipEventDestroy((hipEvent_t)ev));
  return ZAV1_OK;
}
int zav1_event_record(void* ev, zav1_stream_t s) {
  ZCHECK(hipEventRecord((hipEvent_t)ev, S(s)));
  return ZAV1_OK;
}
int zav1_event_elapsed_ms(void* a, void* b, float* ms) {
  if (!ms) return ZAV1_EINVAL;
  ZCHECK(hipEventSynchronize((hipEvent_t)b));
  ZCHECK(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
  return ZAV1_OK;
}

int zav1_event_query(void* ev) {
  if (!ev) return ZAV1_EINVAL;
  const hipError_t e = hipEventQuery((hipEvent_t)ev);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return fail(e, "event_query");
}

int zav1_grid_xy(int32_t* d_xy, int plane_w, int plane_h, int bw, int bh, zav1_stream_t s) {
  if (!d_xy || plane_w <= 0 || plane_h <= 0 || bw <= 0 || bh <= 0) return ZAV1_EINVAL;
  const long n = (long)(plane_w / bw) * (plane_h / bh);
  if (n > INT_MAX) return ZAV1_EINVAL;
  if (n > 0) ZCHECK(zav1::grid_xy(d_xy, plane_w / bw, (int)n, bw, bh, S(s)));
  return (int)n;
}

int zav1_stream_wait_event(zav1_stream_t s, void* ev) {
  if (!ev) return ZAV1_EINVAL;
  ZCHECK(hipStreamWaitEvent(S(s), (hipEvent_t)ev, 0));
  return ZAV1_OK;
}
const char* zav1_last_error(void) { return g_err.c_str(); }

int zav1_build_quant_row_sharp(int qindex, int bd, int sharpness, zav1_quant_row* row) {
  if (qindex < 0 || qindex > 255 || (bd != 8 && bd != 10) || sharpness < -7 || sharpness > 7 || !row)
    return ZAV1_EINVAL;
  zav1::QuantRow r;
  build_row_sharp(qindex, bd, sharpness, &r);
  std::memcpy(row, &r, sizeof(r));
  return ZAV1_OK;
}

int zav1_build_quant_row(int qindex, int bd, zav1_quant_row* row) {
  if (!row || qindex < 0 || qindex > 255 || (bd != 8 && bd != 10)) return ZAV1_EINVAL;
  static_assert(sizeof(zav1_quant_row) == sizeof(zav1::QuantRow), "row layout");
  build_row(qindex, bd, reinterpret_cast<zav1::QuantRow*>(row));
  return ZAV1_OK;
}

const zav1_dsp_v1* zav1_dsp_get(int backend) {
  if (backend != 1) return nullptr;
  if (zav1_device_count() <= 0) return nullptr;
  return &g_table;
}

int zav1_fwd_txfm2d_batch(const int32_t* d_in, int32_t* d_out, const uint8_t* d_types, int tx_size, int n_jobs,
                          zav1_stream_t s) {
  if (tx_size < 0 || tx_size >= 19 || n_jobs < 0 || (n_jobs && (!d_in || !d_out || !d_types))) return ZAV1_EINVAL;
  ZCHECK(zav1::fwd_txfm2d_batch(d_in, d_out, d_types, tx_size, n_jobs, S(s)));
  return ZAV1_OK;
}

int zav1_inv_txfm2d_batch(const int32_t* d_in, int32_t* d_out, const uint8_t* d_types, int tx_size, int n_jobs,
                          int bd, zav1_stream_t s) {
  if (tx_size < 0 || tx_size >= 19 || n_jobs < 0 || (bd != 8 && bd != 10) ||
      (n_jobs && (!d_in || !d_out || !d_types)))
    return ZAV1_EINVAL;
  ZCHECK(zav1::inv_txfm2d_batch(d_in, d_out, d_types, tx_size, n_jobs, bd, S(s)));
  return ZAV1_OK;
}

int zav1_quantize_batch(const int32_t* d_coeff, int32_t* d_qcoeff, int32_t* d_dqcoeff, const uint8_t* d_qidx,
                        int n_per_job, int n_jobs, int log_scale, int bd, int kind, zav1_stream_t s) {
  if (n_per_job <= 0 || n_jobs < 0 || log_scale < 0 || log_scale > 2 || (bd != 8 && bd != 10) ||
      (kind != 0 && kind != 1))
    return ZAV1_EINVAL;
  const zav1::QuantRow* rows = nullptr;
  int rc = quant_rows(&rows);
  if (rc) return rc;
  ZCHECK(zav1::quantize_batch(d_coeff, d_qcoeff, d_dqcoeff, rows + (bd == 10 ? 256 : 0), d_qidx, n_per_job,
                              n_jobs, log_scale, bd, kind, S(s)));
  return ZAV1_OK;
}

static_assert(sizeof(zav1_rdoq_job) == sizeof(zav1::RdoqJob), "rdoq job layout");
static_assert(sizeof(zav1_txb_costs) == 944 * 4 && sizeof(zav1_eob_costs) == 22 * 4, "cost table layout");
static_assert(sizeof(zav1_coeff_costs) == (10 * 944 + 14 * 22) * 4, "cost tables layout");

static_assert(sizeof(zav1_rate_job) == sizeof(zav1::RateJob), "rate job layout");

int zav1_coeff_rate_batch(int tx_size, const int32_t* d_qcoeff, const uint16_t* d_eob, const zav1_rate_job* d_jobs,
                          int n_jobs, const zav1_coeff_costs* d_costs, const uint16_t* d_scans, int32_t* d_rate,
                          zav1_stream_t s) {
  if (tx_size < 0 || tx_size >= 19 || n_jobs < 0 ||
      (n_jobs > 0 && (!d_qcoeff || !d_eob || !d_jobs || !d_costs || !d_scans || !d_rate)))
    return ZAV1_EINVAL;
  ZCHECK(zav1::coeff_rate_batch(tx_size, d_qcoeff, d_eob, reinterpret_cast<const zav1::RateJob*>(d_jobs), n_jobs,
                                reinterpret_cast<const int32_t*>(d_costs), d_scans, d_rate, S(s)));
  return ZAV1_OK;
}

int zav1_rdoq_batch_ws(int tx_size, int qindex, int bd, const int32_t* d_tcoeff, int32_t* d_qcoeff,
                       int32_t* d_dqcoeff, uint16_t* d_eob, const zav1_rdoq_job* d_jobs, int n_jobs,
                       const zav1_coeff_costs* d_costs, const uint16_t* d_scans, uint32_t cut_off_num,
                       uint32_t cut_off_denum, void* d_work, size_t work_bytes, zav1_stream_t s) {
  if (tx_size < 0 || tx_size >= 19 || qindex < 0 || qindex > 255 || (bd != 8 && bd != 10) || n_jobs < 0 ||
      (cut_off_num != 0 && cut_off_denum == 0) || (n_jobs > 0 && (!d_tcoeff || !d_qcoeff || !d_dqcoeff || !d_eob ||
                                                                  !d_jobs || !d_costs || !d_scans)))
    return ZAV1_EINVAL;
  zav1::QuantRow row;
  build_row(qindex, bd, &row);
  ZCHECK(zav1::rdoq_batch(tx_size, d_tcoeff, d_qcoeff, d_dqcoeff, d_eob, reinterpret_cast<const zav1::RdoqJob*>(d_jobs),
                          n_jobs, reinterpret_cast<const int32_t*>(d_costs), d_scans, row.dequant[0],
                          row.dequant[1], cut_off_num, cut_off_denum, S(s), d_work, d_work ? work_bytes : 0));
  return ZAV1_OK;
}

int zav1_rdoq_batch(int tx_size, int qindex, int bd, const int32_t* d_tcoeff, int32_t* d_qcoeff, int32_t* d_dqcoeff,
                    uint16_t* d_eob, const zav1_rdoq_job* d_jobs, int n_jobs, const zav1_coeff_costs* d_costs,
                    const uint16_t* d_scans, uint32_t cut_off_num, uint32_t cut_off_denum, zav1_stream_t s) {
  return zav1_rdoq_batch_ws(tx_size, qindex, bd, d_tcoeff, d_qcoeff, d_dqcoeff, d_eob, d_jobs, n_jobs, d_costs,
                            d_scans, cut_off_num, cut_off_denum, nullptr, 0, s);
}

size_t zav1_rdoq_workspace(int tx_size, int n_jobs) { return zav1::rdoq_workspace(tx_size, n_jobs); }

int zav1_rdoq_lane_min(int tx_size, int n_jobs) {
  if (tx_size < 0 || tx_size >= 19) return ZAV1_EINVAL;
  const int prev = zav1::rdoq_lane_min(tx_size);
  if (n_jobs >= 0) zav1::rdoq_set_lane_min(tx_size, n_jobs);
  return prev;
}

int zav1_sse_i32_batch(const int32_t* d_a, const int32_t* d_b, uint64_t* d_out, int n_per_job, int n_jobs,
                       zav1_stream_t s) {
  if (n_per_job <= 0 || n_jobs < 0) return ZAV1_EINVAL;
  ZCHECK(zav1::sse_i32_batch(d_a, d_b, d_out, n_per_job, n_jobs, S(s)));
  return ZAV1_OK;
}

int zav1_satd_batch(const int16_t* d_res, int32_t* d_coeff, int32_t* d_satd, int n, int n_jobs, zav1_stream_t s) {
  if ((n != 4 && n != 8 && n != 16 && n != 32) || n_jobs < 0) return ZAV1_EINVAL;
  ZCHECK(zav1::satd_batch(d_res, d_coeff, d_satd, n, n_jobs, S(s)));
  return ZAV1_OK;
}

int zav1_scan_order(int tx_size, int scan_class, uint16_t* scan) {
  if (tx_size < 0 || tx_size >= 19 || scan_class < 0 || scan_class > 2 || !scan) return ZAV1_EINVAL;
  zav1::build_scan(tx_size, scan_class, scan);
  int w = zav1::tx_w(tx_size), h = zav1::tx_h(tx_size);
  return (w > 32 ? 32 : w) * (h > 32 ? 32 : h);
}

int zav1_txsearch_batch_q(const uint16_t* d_plane, int stride, int bd, int tx_size, const int32_t* d_xy,
                          int n_blocks, const uint8_t* types, int n_types, const int* qindices, int n_q,
                          uint64_t* d_dist, int32_t* d_eob, uint32_t* d_lsum, zav1_stream_t s) {
  if (tx_size < 0 || tx_size >= 19 || (bd != 8 && bd != 10) || n_blocks < 0 || n_types < 1 || n_types > 16 ||
      !types || !qindices || n_q < 1 || n_q > 8 || stride <= 0 ||
      (n_blocks > 0 && (!d_plane || !d_xy || !d_dist || !d_eob || !d_lsum)))
    return ZAV1_EINVAL;
  for (int q = 0; q < n_q; q++)
    if (qindices[q] < 0 || qindices[q] > 255) return ZAV1_EINVAL;
  zav1::TypeList tl{};
  tl.n = n_types;
  for (int i = 0; i < n_types; i++) {
    if (types[i] > 15 || !tx_supported(tx_size, types[i])) return ZAV1_ENOTSUP;
    tl.t[i] = types[i];
  }
  const zav1::QuantRow* rows = nullptr;
  const uint16_t* iscan = nullptr;
  int rc = quant_rows(&rows);
  if (rc) return rc;
  rc = iscan_table(&iscan);
  if (rc) return rc;
  const zav1::QuantRow* dev_rows[8];
  zav1::QuantRow host_rows[8];
  for (int q = 0; q < n_q; q++) {
    dev_rows[q] = rows + (bd == 10 ? 256 : 0) + qindices[q];
    build_row(qindices[q], bd, &host_rows[q]);
  }
  ZCHECK(zav1::txsearch_batch(d_plane, stride, bd, tx_size, d_xy, n_blocks, tl, dev_rows, host_rows, n_q, iscan,
                              d_dist, d_eob, d_lsum, S(s)));
  return ZAV1_OK;
}

int zav1_txsearch_batch(const uint16_t* d_plane, int stride, int bd, int tx_size, const int32_t* d_xy,
                        int n_blocks, const uint8_t* types, int n_types, int qindex, uint64_t* d_dist,
                        int32_t* d_eob, uint32_t* d_lsum, zav1_stream_t s) {
  return zav1_txsearch_batch_q(d_plane, stride, bd, tx_size, d_xy, n_blocks, types, n_types, &qindex, 1, d_dist,
                               d_eob, d_lsum, s);
}

int zav1_txunit_batch(const void* d_src, int stride, int plane_w, int plane_h, int bd, int tx_size,
                      const uint16_t* d_pred, size_t pred_len, const zav1_txunit_job* d_jobs, int n_jobs, int qindex,
                      unsigned flags, uint64_t* d_dist, int32_t* d_eob, uint8_t* d_cul, int32_t* d_qcoeff,
                      uint16_t* d_recon, zav1_stream_t s) {
  static_assert(sizeof(zav1_txunit_job) == sizeof(zav1::TxUnitJob), "txunit job layout");
  if (tx_size < 0 || tx_size >= 19 || (bd != 8 && bd != 10) || n_jobs < 0 || qindex < 0 || qindex > 255 ||
      plane_w <= 0 || plane_h <= 0 || stride < plane_w || (flags & ~7u) ||
      (n_jobs && (!d_src || !d_pred || !d_jobs || !d_dist || !d_eob || !d_cul)))
    return ZAV1_EINVAL;
  const zav1::QuantRow* rows = nullptr;
  const uint16_t* iscan = nullptr;
  int rc = quant_rows(&rows);
  if (rc) return rc;
  rc = iscan_table(&iscan);
  if (rc) return rc;
  ZCHECK(zav1::txunit_batch(d_src, stride, plane_w, plane_h, bd, tx_size, d_pred, (long)pred_len,
                            reinterpret_cast<const zav1::TxUnitJob*>(d_jobs), n_jobs,
                            rows + (bd == 10 ? 256 : 0) + qindex, iscan, flags, d_dist, d_eob, d_cul, d_qcoeff,
                            d_recon, S(s)));
  return ZAV1_OK;
}

int zav1_txunit_rd_workspace(int tx_size, int n_jobs, size_t* bytes) {
  if (tx_size < 0 || tx_size >= 19 || n_jobs < 0 || !bytes) return ZAV1_EINVAL;
  *bytes = n_jobs ? zav1::txunit_rd_workspace(tx_size, n_jobs) : 0;
  return ZAV1_OK;
}

int zav1_txunit_rd_batch(const void* d_src, int stride, int plane_w, int plane_h, int bd, int tx_size,
                         const uint16_t* d_pred, size_t pred_len, const zav1_txunit_rd_job* d_jobs, int n_jobs,
                         int qindex, unsigned flags, uint32_t cut_off_num, uint32_t cut_off_denum,
                         const zav1_coeff_costs* d_costs, void* d_work, size_t work_bytes, uint64_t* d_dist,
                         int32_t* d_eob, int32_t* d_rate, uint8_t* d_cul, int32_t* d_qcoeff, uint16_t* d_recon,
                         zav1_stream_t s) {
  static_assert(sizeof(zav1_txunit_rd_job) == sizeof(zav1::TxUnitRdJob), "rd job layout");
  if (tx_size < 0 || tx_size >= 19 || (bd != 8 && bd != 10) || n_jobs < 0 || qindex < 0 || qindex > 255 ||
      plane_w <= 0 || plane_h <= 0 || stride < plane_w || (flags & ~31u) ||
      ((flags & ZAV1_TXU_RDOQ) && (flags & ZAV1_TXU_QUANT_B)) || (cut_off_num != 0 && cut_off_denum == 0) ||
      (n_jobs && (!d_src || !d_pred || !d_jobs || !d_costs || !d_work || !d_dist || !d_eob || !d_rate || !d_cul)))
    return ZAV1_EINVAL;
  if (n_jobs == 0) return ZAV1_OK;
  if (work_bytes < zav1::txunit_rd_workspace(tx_size, n_jobs)) return ZAV1_EINVAL;
  const zav1::QuantRow* rows = nullptr;
  const uint16_t *iscan = nullptr, *fscan = nullptr;
  int rc = quant_rows(&rows);
  if (rc) return rc;
  rc = iscan_table(&iscan);
  if (rc) return rc;
  rc = fscan_table(tx_size, &fscan);
  if (rc) return rc;
  // the device row for the kernels, its dequantizers by value for the trellis
  zav1::QuantRow row;
  build_row(qindex, bd, &row);
  const zav1::QuantRow* drow = rows + (bd == 10 ? 256 : 0) + qindex;
  ZCHECK(zav1::txunit_rd_batch(d_src, stride, plane_w, plane_h, bd, tx_size, d_pred, (long)pred_len,
                               reinterpret_cast<const zav1::TxUnitRdJob*>(d_jobs), n_jobs, drow, iscan, fscan,
                               reinterpret_cast<const int32_t*>(d_costs), flags, cut_off_num, cut_off_denum, d_work,
                               d_dist, d_eob, d_rate, d_cul, d_qcoeff, d_recon, S(s), row.dequant[0],
                               row.dequant[1], nullptr, nullptr, work_bytes));
  return ZAV1_OK;
}

int zav1_txunit_rd_batch_noalloc(const void* d_src, int stride, int plane_w, int plane_h, int bd, int tx_size,
                                 const uint16_t* d_pred, size_t pred_len, const zav1_txunit_rd_job* d_jobs,
                                 int n_jobs, int qindex, unsigned flags, uint32_t cut_off_num,
                                 uint32_t cut_off_denum, const zav1_coeff_costs* d_costs, uint64_t* d_dist,
                                 int32_t* d_eob, int32_t* d_rate, uint8_t* d_cul, int32_t* d_qcoeff,
                                 uint16_t* d_recon, zav1_stream_t s) {
  if (tx_size < 0 || tx_size >= 19 || (bd != 8 && bd != 10) || n_jobs < 0 || qindex < 0 || qindex > 255 ||
      plane_w <= 0 || plane_h <= 0 || stride < plane_w || (flags & ~31u) ||
      ((flags & ZAV1_TXU_RDOQ) && (flags & ZAV1_TXU_QUANT_B)) || (cut_off_num != 0 && cut_off_denum == 0) ||
      (n_jobs && (!d_src || !d_pred || !d_jobs || !d_costs || !d_dist || !d_eob || !d_rate || !d_cul)))
    return ZAV1_EINVAL;
  if (n_jobs == 0) return ZAV1_OK;
  const size_t need = zav1::txunit_rd_workspace(tx_size, n_jobs);
  if (need > zav1::noalloc_scratch_bytes()) {
    g_err = "zav1_txunit_rd_batch_noalloc: " + std::to_string(n_jobs) + " jobs need " + std::to_string(need) +
            " bytes of scratch, more than the fixed " + std::to_string(zav1::noalloc_scratch_bytes());
    return ZAV1_EINVAL;
  }
  int dev = 0;
  ZCHECK(hipGetDevice(&dev));
  void* work = zav1::noalloc_scratch(dev);
  const zav1::QuantRow* rows = nullptr;
  const uint16_t *iscan = nullptr, *fscan = nullptr;
  int rc = tables_ready(&rows, &iscan, &fscan);
  if (rc) return rc;
  if (!work) {
    g_err = "device not initialised: call zav1_init(device) before zav1_txunit_rd_batch_noalloc";
    return ZAV1_EINVAL;
  }
  zav1::QuantRow row;
  build_row(qindex, bd, &row);
  ZCHECK(zav1::txunit_rd_batch(d_src, stride, plane_w, plane_h, bd, tx_size, d_pred, (long)pred_len,
                               reinterpret_cast<const zav1::TxUnitRdJob*>(d_jobs), n_jobs,
                               rows + (bd == 10 ? 256 : 0) + qindex, iscan, fscan + tx_size * 3072,
                               reinterpret_cast<const int32_t*>(d_costs), flags, cut_off_num, cut_off_denum, work,
                               d_dist, d_eob, d_rate, d_cul, d_qcoeff, d_recon, S(s), row.dequant[0],
                               row.dequant[1]));
  return ZAV1_OK;
}

size_t zav1_noalloc_scratch_bytes(void) { return zav1::noalloc_scratch_bytes(); }

static bool funnel_dim(int v) { return v == 4 || v == 8 || v == 16 || v == 32 || v == 64; }

int zav1_pred_satd_batch(const void* d_src, int stride, int plane_w, int plane_h, int bd, int w, int h,
                         const uint16_t* d_pred, size_t pred_len, const zav1_txunit_job* d_jobs, int n_jobs,
                         uint32_t* d_satd, zav1_stream_t s) {
  if (!funnel_dim(w) || !funnel_dim(h) || (bd != 8 && bd != 10) || n_jobs < 0 || plane_w <= 0 || plane_h <= 0 ||
      stride < plane_w || (n_jobs && (!d_src || !d_pred || !d_jobs || !d_satd)))
    return ZAV1_EINVAL;
  const int t = std::min(std::min(w, h), 32);
  if ((long)n_jobs * (w / t) * (h / t) > INT32_MAX) return ZAV1_EINVAL;
  ZCHECK(zav1::pred_satd_batch(d_src, stride, plane_w, plane_h, bd, w, h, d_pred, (long)pred_len,
                               reinterpret_cast<const zav1::TxUnitJob*>(d_jobs), n_jobs, d_satd, S(s)));
  return ZAV1_OK;
}

int zav1_md_fast_select(const uint32_t* d_satd, const uint32_t* d_rate, uint64_t lambda, int n_units, int n_cands,
                        int k, const int32_t* d_unit_xy, int xy_stride, int w, int h, const uint8_t* types,
                        int n_types, int survivors_only, uint8_t* d_sel, uint64_t* d_sel_cost,
                        zav1_txunit_job* d_jobs, zav1_stream_t s) {
  if (n_units < 0 || n_cands < 1 || n_cands > 64 || k < 1 || k > n_cands || !funnel_dim(w) || !funnel_dim(h) ||
      n_types < 1 || n_types > 16 || !types || xy_stride < 2 || (survivors_only != 0 && survivors_only != 1) ||
      (n_units && (!d_satd || !d_unit_xy || !d_sel || !d_jobs)))
    return ZAV1_EINVAL;
  if ((long)n_units * (survivors_only ? k : n_cands) * w * h > UINT32_MAX) return ZAV1_EINVAL;  // pred_off range
  zav1::TypeList tl{};
  tl.n = n_types;
  for (int i = 0; i < n_types; i++) {
    if (types[i] > 15) return ZAV1_EINVAL;
    tl.t[i] = types[i];
  }
  ZCHECK(zav1::fast_select(d_satd, d_rate, lambda, n_units, n_cands, k, d_unit_xy, xy_stride, w, h, tl,
                           survivors_only, d_sel, d_sel_cost, reinterpret_cast<zav1::TxUnitJob*>(d_jobs), S(s)));
  return ZAV1_OK;
}

static int md_cand_list(const zav1_md_cand* cands, int n_cands, int w, int h, zav1::MdCandList* cl) {
  static_assert(sizeof(zav1_md_cand) == sizeof(zav1::MdCand), "md cand layout");
  if (!cands || n_cands < 1 || n_cands > 64) return ZAV1_EINVAL;
  *cl = zav1::MdCandList{};
  cl->n = n_cands;
  for (int i = 0; i < n_cands; i++) {
    const zav1_md_cand c = cands[i];
    if (c.mode > 12 || c.delta < -3 || c.delta > 3 || c.fi_mode > 5) return ZAV1_EINVAL;
    if (c.fi_mode < 5 && (c.mode != 0 || w > 32 || h > 32)) return ZAV1_EINVAL;
    cl->c[i] = zav1::MdCand{c.mode, c.delta, c.fi_mode, 0};
  }
  return ZAV1_OK;
}

int zav1_md_fast_satd(const void* d_src, const void* d_recon, int stride, int plane_w, int plane_h, int bd, int w,
                      int h, const zav1_md_unit* d_units, int n_units, const zav1_md_cand* cands, int n_cands,
                      uint32_t* d_satd, zav1_stream_t s) {
  static_assert(sizeof(zav1_md_unit) == sizeof(zav1::MdUnit), "md unit layout");
  if (!funnel_dim(w) || !funnel_dim(h) || w > 4 * h || h > 4 * w || (bd != 8 && bd != 10) || n_units < 0 ||
      plane_w <= 0 || plane_h <= 0 ||
      stride < plane_w || (n_units && (!d_src || !d_recon || !d_units || !d_satd)))
    return ZAV1_EINVAL;
  zav1::MdCandList cl;
  const int rc = md_cand_list(cands, n_cands, w, h, &cl);
  if (rc) return rc;
  ZCHECK(zav1::md_fast_satd(d_src, d_recon, stride, plane_w, plane_h, bd, w, h,
                            reinterpret_cast<const zav1::MdUnit*>(d_units), n_units, cl, d_satd, S(s)));
  return ZAV1_OK;
}

int zav1_md_predict_survivors(const void* d_recon, int stride, int plane_w, int plane_h, int bd, int w, int h,
                              const zav1_md_unit* d_units, int n_units, const zav1_md_cand* cands, int n_cands,
                              const uint8_t* d_sel, int k, uint16_t* d_pred, zav1_stream_t s) {
  if (!funnel_dim(w) || !funnel_dim(h) || (bd != 8 && bd != 10) || n_units < 0 || plane_w <= 0 || plane_h <= 0 ||
      stride < plane_w || k < 1 ||
      k > n_cands || (n_units && (!d_recon || !d_units || !d_sel || !d_pred)))
    return ZAV1_EINVAL;
  if ((long)n_units * k > INT32_MAX || (long)n_units * k * w * h > UINT32_MAX) return ZAV1_EINVAL;
  zav1::MdCandList cl;
  const int rc = md_cand_list(cands, n_cands, w, h, &cl);
  if (rc) return rc;
  ZCHECK(zav1::md_predict_survivors(d_recon, stride, plane_w, plane_h, bd, w, h,
                                    reinterpret_cast<const zav1::MdUnit*>(d_units), n_units, cl, d_sel, k, d_pred,
                                    S(s)));
  return ZAV1_OK;
}

int zav1_intra_pred_batch(const void* d_plane, int stride, int bd, const zav1_intra_job* d_jobs, int n_jobs,
                          uint16_t* d_out, zav1_stream_t s) {
  static_assert(sizeof(zav1_intra_job) == sizeof(zav1::IntraJob), "intra job layout");
  if ((bd != 8 && bd != 10) || n_jobs < 0 || stride <= 0 || (n_jobs && (!d_plane || !d_jobs || !d_out)))
    return ZAV1_EINVAL;
  ZCHECK(zav1::intra_pred_batch(d_plane, stride, bd, reinterpret_cast<const zav1::IntraJob*>(d_jobs), n_jobs, d_out,
                                S(s)));
  return ZAV1_OK;
}

int zav1_cfl_batch(const void* d_luma, int stride, int bd, const zav1_cfl_job* d_jobs, int n_jobs,
                   const uint16_t* d_dc, uint16_t* d_out, zav1_stream_t s) {
  static_assert(sizeof(zav1_cfl_job) == sizeof(zav1::CflJob), "cfl job layout");
  if ((bd != 8 && bd != 10) || n_jobs < 0 || stride <= 0 || (n_jobs && (!d_luma || !d_jobs || !d_dc || !d_out)))
    return ZAV1_EINVAL;
  ZCHECK(zav1::cfl_batch(d_luma, stride, bd, reinterpret_cast<const zav1::CflJob*>(d_jobs), n_jobs, d_dc, d_out,
                         S(s)));
  return ZAV1_OK;
}

static int geom_dev(const zav1_deblock_geom* geom, int plane, int ss, int plane_w, int plane_h,
                    zav1::DeblockGeomDev* g) {
  if (!geom || !geom->block_id || !geom->bw || !geom->bh || !geom->tw || !geom->th || geom->mi_cols <= 0 ||
      geom->mi_rows <= 0 || ((plane_w << ss) + 3) / 4 > geom->mi_cols || ((plane_h << ss) + 3) / 4 > geom->mi_rows)
    return ZAV1_EINVAL;
  // A 14-tap edge reaches 6 samples past a 16-aligned coordinate: luma planes
  // padded to 8 samples (chroma to 4) keep every filter inside the plane.
  const int align = plane == 0 ? 8 : 4;
  if (plane_w % align || plane_h % align) return ZAV1_EINVAL;
  *g = zav1::DeblockGeomDev{geom->mi_cols, geom->mi_rows, geom->true_w, geom->true_h, geom->block_id,
                            geom->bw,      geom->bh,      geom->tw,     geom->th,     geom->skip};
  return ZAV1_OK;
}

int zav1_deblock_plane(void* d_plane, int stride, int plane_w, int plane_h, int bd, int plane, int ss, int level_v,
                       int level_h, int sharpness, const zav1_deblock_geom* geom, zav1_stream_t s) {
  zav1::DeblockGeomDev g{};
  if ((bd != 8 && bd != 10) || plane < 0 || plane > 2 || ss < 0 || ss > 1 || (plane == 0 && ss != 0) ||
      plane_w <= 0 || plane_h <= 0 || stride < plane_w || level_v < 0 || level_v > 63 || level_h < 0 ||
      level_h > 63 || sharpness < 0 || sharpness > 7 || !d_plane ||
      geom_dev(geom, plane, ss, plane_w, plane_h, &g) != ZAV1_OK)
    return ZAV1_EINVAL;
  ZCHECK(zav1::deblock_plane(d_plane, stride, plane_w, plane_h, bd, plane, ss, level_v, level_h, sharpness, g, S(s)));
  return ZAV1_OK;
}

int zav1_deblock_level_sse(const void* d_recon, const void* d_src, int plane_w, int plane_h, int bd, int plane, int ss,
                           int sharpness, const zav1_deblock_geom* geom, const uint8_t* levels, int n_levels,
                           void* d_work, size_t work_bytes, uint64_t* d_sse, zav1_stream_t s) {
  zav1::DeblockGeomDev g{};
  if ((bd != 8 && bd != 10) || plane < 0 || plane > 2 || ss < 0 || ss > 1 || (plane == 0 && ss != 0) ||
      sharpness < 0 || sharpness > 7 || !d_recon || !d_src || !d_work || !d_sse || !levels || n_levels < 1 ||
      n_levels > 64 || plane_w <= 0 || plane_h <= 0 || geom_dev(geom, plane, ss, plane_w, plane_h, &g) != ZAV1_OK)
    return ZAV1_EINVAL;
  for (int i = 0; i < n_levels; i++)
    if (levels[i] > 63) return ZAV1_EINVAL;
  const size_t need = (size_t)n_levels * plane_w * plane_h * (bd > 8 ? 2 : 1);
  if (work_bytes < need) return ZAV1_EINVAL;
  ZCHECK(zav1::deblock_trials(d_recon, d_src, d_work, plane_w, plane_h, bd, plane, ss, levels, n_levels, sharpness, g,
                              d_sse, S(s)));
  return ZAV1_OK;
}

// search_filter_level (rust/crates/svtav1-encoder/src/deblock.rs:337) on the
// still path: start level 0, step 4, bias against raising the level, step
// halving on convergence and an early exit after conv_th convergences (0 =
// never).  Each round's (at most two) untried levels run as one GPU trial
// batch; the decisions read the SSE table exactly as the sequential search
// does, so the pick is identical.
int zav1_deblock_search_level(const void* d_recon, const void* d_src, int plane_w, int plane_h, int bd, int plane,
                              int ss, int sharpness, const zav1_deblock_geom* geom, int conv_th, void* d_work,
                              size_t work_bytes, uint64_t* d_sse, int* best_level, int64_t* zero_sse,
                              int64_t* best_sse, zav1_stream_t s) {
  if (!best_level || !zero_sse || !best_sse || conv_th < 0) return ZAV1_EINVAL;
  int64_t ss_err[64];
  for (int i = 0; i < 64; i++) ss_err[i] = -1;
  auto trial = [&](int a, int b) -> int {  // evaluate levels a (and b >= 0) not yet in the table
    uint8_t lv[2];
    int n = 0;
    if (a >= 0 && ss_err[a] < 0) lv[n++] = (uint8_t)a;
    if (b >= 0 && b != a && ss_err[b] < 0) lv[n++] = (uint8_t)b;
    if (!n) return ZAV1_OK;
    int rc = zav1_deblock_level_sse(d_recon, d_src, plane_w, plane_h, bd, plane, ss, sharpness, geom, lv, n, d_work,
                                    work_bytes, d_sse, s);
    if (rc) return rc;
    uint64_t out[2];
    ZCHECK(hipMemcpyAsync(out, d_sse, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, S(s)));
    ZCHECK(hipStreamSynchronize(S(s)));
    for (int i = 0; i < n; i++) ss_err[lv[i]] = (int64_t)out[i];
    return ZAV1_OK;
  };
  int filt_mid = 0, filt_dir = 0, conv = 0;
  int step = filt_mid < 16 ? 4 : filt_mid / 4;
  int rc = trial(filt_mid, -1);
  if (rc) return rc;
  int64_t best_err = ss_err[filt_mid];
  int filt_best = filt_mid;
  while (step > 0) {
    const int hi = filt_mid + step > 63 ? 63 : filt_mid + step;
    const int lo = filt_mid - step < 0 ? 0 : filt_mid - step;
    int64_t bias = (best_err >> (15 - (filt_mid / 8))) * (int64_t)step;
    bias >>= 1;
    const bool try_lo = filt_dir <= 0 && lo != filt_mid, try_hi = filt_dir >= 0 && hi != filt_mid;
    rc = trial(try_lo ? lo : -1, try_hi ? hi : -1);
    if (rc) return rc;
    if (try_lo && ss_err[lo] < best_err + bias) {
      if (ss_err[lo] < best_err) best_err = ss_err[lo];
      filt_best = lo;
    }
    if (try_hi && ss_err[hi] < best_err - bias) {
      best_err = ss_err[hi];
      filt_best = hi;
    }
    if (filt_best == filt_mid) {
      conv++;
      step = (conv == conv_th) ? 0 : step / 2;
      filt_dir = 0;
    } else {
      filt_dir = filt_best < filt_mid ? -1 : 1;
      filt_mid = filt_best;
    }
  }
  *best_level = filt_best;
  *zero_sse = ss_err[0];
  *best_sse = ss_err[filt_best];
  return ZAV1_OK;
}

int zav1_cdef_search(const void* d_rec_y, const void* d_rec_u, const void* d_rec_v, const void* d_src_y,
                     const void* d_src_u, const void* d_src_v, int width, int height, int bd, int qindex,
                     int subsampling, const uint8_t* d_skip, int mi_cols, int mi_rows, int32_t* d_dirvar,
                     uint64_t* d_mse_y, uint64_t* d_mse_uv, zav1_stream_t s) {
  if ((bd != 8 && bd != 10) || width <= 0 || height <= 0 || qindex < 0 || qindex > 255 || subsampling < 1 ||
      !d_rec_y || !d_src_y || !d_mse_y)
    return ZAV1_EINVAL;
  if (d_skip && (mi_cols < (width + 3) / 4 || mi_rows < (height + 3) / 4)) return ZAV1_EINVAL;
  if (d_mse_uv && (!d_rec_u || !d_rec_v || !d_src_u || !d_src_v || !d_dirvar)) return ZAV1_EINVAL;
  const int damping = 3 + (qindex >> 6);
  ZCHECK(zav1::cdef_search(d_rec_y, d_rec_u, d_rec_v, d_src_y, d_src_u, d_src_v, width, height, bd, damping,
                           subsampling < 4 ? subsampling : 4, 1, d_skip, mi_cols, mi_rows, d_dirvar, d_mse_y,
                           d_mse_uv, S(s)));
  return ZAV1_OK;
}

int zav1_cdef_apply(const void* d_in_y, const void* d_in_u, const void* d_in_v, void* d_out_y, void* d_out_u,
                    void* d_out_v, int width, int height, int bd, int damping, const uint8_t* d_fb_strength,
                    const uint8_t* d_skip, int mi_cols, int mi_rows, uint64_t* d_stats, zav1_stream_t s) {
  if ((bd != 8 && bd != 10) || width <= 0 || height <= 0 || (width & 7) || (height & 7) || damping < 3 ||
      damping > 6 || !d_in_y || !d_out_y || !d_fb_strength || d_in_y == d_out_y)
    return ZAV1_EINVAL;
  if ((d_in_u == nullptr) != (d_in_v == nullptr) || (d_in_u && (!d_out_u || !d_out_v || d_in_u == d_out_u ||
                                                                d_in_v == d_out_v)))
    return ZAV1_EINVAL;
  if (d_skip && (mi_cols < (width + 3) / 4 || mi_rows < (height + 3) / 4)) return ZAV1_EINVAL;
  ZCHECK(zav1::cdef_apply(d_in_y, d_in_u, d_in_v, d_out_y, d_out_u, d_out_v, width, height, bd, damping,
                          d_fb_strength, d_skip, mi_cols, mi_rows, d_stats, S(s)));
  return ZAV1_OK;
}

static bool lr_unit_size(int u) { return u == 32 || u == 64 || u == 128 || u == 256; }

int zav1_wiener_apply(const void* d_dgd, void* d_out, int plane_w, int plane_h, int bd, int unit_size, int ss_y,
                      const zav1_wiener_unit* d_units, zav1_stream_t s) {
  if (!d_dgd || !d_out || d_dgd == d_out || !d_units || plane_w <= 0 || plane_h <= 0 || (bd != 8 && bd != 10) ||
      !lr_unit_size(unit_size) || (ss_y != 0 && ss_y != 1))
    return ZAV1_EINVAL;
  int nh = 0, nv = 0;
  zav1_rest_unit_grid(unit_size, plane_w, plane_h, &nh, &nv);
  ZCHECK(zav1::wiener_apply(d_dgd, d_out, plane_w, plane_h, bd, unit_size, ss_y, nh, nv, d_units, S(s)));
  return ZAV1_OK;
}

int zav1_wiener_refine(const void* d_dgd, const void* d_src, int plane_w, int plane_h, int bd, int win, int unit_size,
                       int ss_y, int use_refinement, int max_one_step, zav1_wiener_unit* d_units, int64_t* d_err,
                       zav1_stream_t s) {
  if (!d_dgd || !d_src || !d_units || !d_err || plane_w <= 0 || plane_h <= 0 || (bd != 8 && bd != 10) ||
      (win != 5 && win != 7) || !lr_unit_size(unit_size) || (ss_y != 0 && ss_y != 1) ||
      (use_refinement != 0 && use_refinement != 1) || (max_one_step != 0 && max_one_step != 1))
    return ZAV1_EINVAL;
  int nh = 0, nv = 0;
  zav1_rest_unit_grid(unit_size, plane_w, plane_h, &nh, &nv);
  ZCHECK(zav1::wiener_refine(d_dgd, d_src, plane_w, plane_h, bd, win, unit_size, ss_y, nh, nv, use_refinement,
                             max_one_step, d_units, d_err, S(s)));
  return ZAV1_OK;
}

int zav1_rest_unit_grid(int unit_size, int plane_w, int plane_h, int* n_h, int* n_v) {
  if (unit_size < 8 || unit_size > 256 || (unit_size & (unit_size - 1)) || plane_w <= 0 || plane_h <= 0 || !n_h ||
      !n_v)
    return ZAV1_EINVAL;
  auto count = [&](int size) {
    const int n = (size + (unit_size >> 1)) / unit_size;
    return n < 1 ? 1 : n;
  };
  *n_h = count(plane_w);
  *n_v = count(plane_h);
  return ZAV1_OK;
}

int zav1_wiener_stats(const void* d_dgd, const void* d_src, int plane_w, int plane_h, int bd, int win, int unit_size,
                      int ss_y, int64_t* d_M, int64_t* d_H, zav1_stream_t s) {
  int nh = 0, nv = 0;
  if ((bd != 8 && bd != 10) || (win != 5 && win != 7) || ss_y < 0 || ss_y > 1 || !d_dgd || !d_src || !d_M || !d_H ||
      zav1_rest_unit_grid(unit_size, plane_w, plane_h, &nh, &nv) != ZAV1_OK)
    return ZAV1_EINVAL;
  ZCHECK(zav1::wiener_stats(d_dgd, d_src, plane_w, plane_h, bd, win, unit_size, ss_y, nh, nv, d_M, d_H, S(s)));
  return ZAV1_OK;
}

int zav1_unit_sse(const void* d_a, const void* d_b, int plane_w, int plane_h, int bd, int unit_size, int ss_y,
                  uint64_t* d_out, zav1_stream_t s) {
  int nh = 0, nv = 0;
  if ((bd != 8 && bd != 10) || ss_y < 0 || ss_y > 1 || !d_a || !d_b || !d_out ||
      zav1_rest_unit_grid(unit_size, plane_w, plane_h, &nh, &nv) != ZAV1_OK)
    return ZAV1_EINVAL;
  ZCHECK(zav1::unit_sse(d_a, d_b, plane_w, plane_h, bd, unit_size, ss_y, nh, nv, d_out, S(s)));
  return ZAV1_OK;
}

int zav1_sgr_stats(const void* d_dgd, const void* d_src, int plane_w, int plane_h, int bd, int unit_size, int ss_y,
                   int64_t* d_out, zav1_stream_t s) {
  int nh = 0, nv = 0;
  if ((bd != 8 && bd != 10) || ss_y < 0 || ss_y > 1 || !d_dgd || !d_src || !d_out ||
      zav1_rest_unit_grid(unit_size, plane_w, plane_h, &nh, &nv) != ZAV1_OK)
    return ZAV1_EINVAL;
  ZCHECK(zav1::sgr_stats(d_dgd, d_src, plane_w, plane_h, bd, unit_size, ss_y, nh, nv, d_out, S(s)));
  return ZAV1_OK;
}

// ---- contexts and MD waves -------------------------------------------------
}  // extern "C"

struct zav1_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  char* arena = nullptr;
  size_t cap = 0, used = 0;
  int err = ZAV1_OK;  // first error of the submitted waves
};

namespace {
int ctx_take(zav1_ctx* c, size_t bytes, void** p) {
  const size_t b = (bytes + 255) & ~(size_t)255;
  if (b > c->cap - c->used) return ZAV1_ENOMEM;
  *p = c->arena + c->used;
  c->used += b;
  return ZAV1_OK;
}
int tx_size_of(int w, int h) {
  for (int s = 0; s < 19; s++)
    if (zav1::tx_w(s) == w && zav1::tx_h(s) == h) return s;
  return -1;
}
}  // namespace

extern "C" {

int zav1_ctx_create(int device, size_t hbm_budget, zav1_ctx** ctx) {
  if (!ctx || hbm_budget == 0) return ZAV1_EINVAL;
  *ctx = nullptr;
  int n = 0;
  ZCHECK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return ZAV1_EINVAL;
  ZCHECK(hipSetDevice(device));
  zav1_ctx* c = new zav1_ctx();
  c->device = device;
  c->cap = (hbm_budget + 255) & ~(size_t)255;
  hipError_t e = hipMalloc((void**)&c->arena, c->cap);
  if (e != hipSuccess) {
    delete c;
    (void)hipGetLastError();
    return e == hipErrorOutOfMemory ? ZAV1_ENOMEM : fail(e, "hipMalloc(arena)");
  }
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    (void)hipFree(c->arena);
    delete c;
    return fail(e, "hipStreamCreateWithFlags");
  }
  *ctx = c;
  return ZAV1_OK;
}

int zav1_ctx_destroy(zav1_ctx* c) {
  if (!c) return ZAV1_EINVAL;
  (void)hipSetDevice(c->device);
  hipError_t e1 = hipStreamSynchronize(c->stream);
  hipError_t e2 = hipStreamDestroy(c->stream);
  hipError_t e3 = hipFree(c->arena);
  delete c;
  ZCHECK(e1);
  ZCHECK(e2);
  ZCHECK(e3);
  return ZAV1_OK;
}

int zav1_txsearch_frame(zav1_ctx* c, const uint16_t* d_plane, int w, int h, int bd, const int* qindices, int n_q,
                        zav1_frame_search* out) {
  if (!c || !d_plane || !out || w <= 0 || h <= 0 || (bd != 8 && bd != 10) || !qindices || n_q < 1 || n_q > 8)
    return ZAV1_EINVAL;
  for (int q = 0; q < n_q; q++)
    if (qindices[q] < 0 || qindices[q] > 255) return ZAV1_EINVAL;
  // reserve every grid and output first: all or nothing
  const size_t mark = c->used;
  std::memset(out, 0, sizeof(*out));
  void* xy[19] = {};
  for (int ts = 0; ts < 19; ts++) {
    const int bw = zav1::tx_w(ts), bh = zav1::tx_h(ts), mx = bw > bh ? bw : bh;
    const long nb = (long)(w / bw) * (h / bh);
    if (nb > INT_MAX / 16) {
      c->used = mark;
      return ZAV1_EINVAL;
    }
    // the intra ext-tx set of the size (coeff_c.rs:181-212): DCT only once a
    // side reaches 32 (square-up size >= 32x32), five types at 16x16, seven
    // (with V_DCT / H_DCT) at every other size
    int nt = 0;
    if (mx >= 32) out->types[ts][nt++] = 0;
    else {
      const uint8_t set[7] = {0, 1, 2, 3, 9, 10, 11};
      for (int i = 0; i < (bw == 16 && bh == 16 ? 5 : 7); i++) out->types[ts][nt++] = set[i];
    }
    out->n_blocks[ts] = (int)nb;
    out->n_types[ts] = nt;
    if (nb == 0) continue;
    const size_t n = (size_t)nb * nt * n_q;
    void *d, *e, *l;
    if (ctx_take(c, 8 * (size_t)nb, &xy[ts]) || ctx_take(c, 8 * n, &d) || ctx_take(c, 4 * n, &e) ||
        ctx_take(c, 4 * n, &l)) {
      c->used = mark;
      std::memset(out, 0, sizeof(*out));
      return ZAV1_ENOMEM;
    }
    out->dist[ts] = (uint64_t*)d;
    out->eob[ts] = (int32_t*)e;
    out->lsum[ts] = (uint32_t*)l;
  }
  for (int ts = 0; ts < 19; ts++) {
    if (out->n_blocks[ts] == 0) continue;
    int rc = zav1_grid_xy((int32_t*)xy[ts], w, h, zav1::tx_w(ts), zav1::tx_h(ts), (zav1_stream_t)c->stream);
    if (rc >= 0)
      rc = zav1_txsearch_batch_q(d_plane, w, bd, ts, (const int32_t*)xy[ts], out->n_blocks[ts], out->types[ts],
                                 out->n_types[ts], qindices, n_q, out->dist[ts], out->eob[ts], out->lsum[ts],
                                 (zav1_stream_t)c->stream);
    if (rc < 0) {
      // all or nothing: let the enqueued work drain, then release the
      // reservation and the partly written outputs
      (void)hipStreamSynchronize(c->stream);
      c->used = mark;
      std::memset(out, 0, sizeof(*out));
      return rc;
    }
  }
  return ZAV1_OK;
}

int zav1_ctx_alloc(zav1_ctx* c, size_t bytes, void** p) {
  if (!c || !p) return ZAV1_EINVAL;
  return ctx_take(c, bytes, p);
}

int zav1_ctx_usage(const zav1_ctx* c, size_t* used, size_t* capacity) {
  if (!c) return ZAV1_EINVAL;
  if (used) *used = c->used;
  if (capacity) *capacity = c->cap;
  return ZAV1_OK;
}

int zav1_ctx_reset(zav1_ctx* c) {
  if (!c) return ZAV1_EINVAL;
  ZCHECK(hipStreamSynchronize(c->stream));
  c->used = 0;
  return ZAV1_OK;
}

zav1_stream_t zav1_ctx_stream(zav1_ctx* c) { return c ? (zav1_stream_t)c->stream : nullptr; }

int zav1_image_upload(zav1_ctx* c, const void* host, size_t host_stride, int w, int h, int bd, void** d_plane) {
  if (!c || !host || !d_plane || w <= 0 || h <= 0 || host_stride < (size_t)w || (bd != 8 && bd != 10))
    return ZAV1_EINVAL;
  const size_t es = bd > 8 ? 2 : 1;
  void* p = nullptr;
  const int rc = ctx_take(c, (size_t)w * h * es, &p);
  if (rc) return rc;
  ZCHECK(hipSetDevice(c->device));
  ZCHECK(hipMemcpy2DAsync(p, (size_t)w * es, host, host_stride * es, (size_t)w * es, (size_t)h,
                          hipMemcpyHostToDevice, c->stream));
  *d_plane = p;
  return ZAV1_OK;
}

int zav1_md_wave_submit(zav1_ctx* c, const zav1_md_wave* v) {
  if (!c || !v) return ZAV1_EINVAL;
  const int ts = tx_size_of(v->w, v->h);
  if (ts < 0 || v->n_units < 0 || v->k < 1 || v->k > v->n_cands || v->n_types < 1 || v->n_types > 16 || !v->types ||
      !v->d_sel || !v->d_dist || !v->d_eob || !v->d_cul || (v->txu_flags & ~7u) ||
      ((v->txu_flags & (ZAV1_TXU_RECON | ZAV1_TXU_SPATIAL)) && !v->d_recon_out) || v->qindex < 0 || v->qindex > 255)
    return ZAV1_EINVAL;
  if (v->n_units == 0) return ZAV1_OK;
  const long n_jobs = (long)v->n_units * v->k * v->n_types;
  if (n_jobs > INT32_MAX) return ZAV1_EINVAL;
  // intermediates from the arena; on ENOMEM nothing is enqueued and the arena is left as it was
  const size_t mark = c->used;
  void *satd = nullptr, *jobs = nullptr, *pred = nullptr;
  int rc = ctx_take(c, sizeof(uint32_t) * (size_t)v->n_units * v->n_cands, &satd);
  if (!rc) rc = ctx_take(c, sizeof(zav1_txunit_job) * (size_t)n_jobs, &jobs);
  if (!rc) rc = ctx_take(c, sizeof(uint16_t) * (size_t)v->n_units * v->k * v->w * v->h, &pred);
  if (rc) {
    c->used = mark;
    return rc;
  }
  ZCHECK(hipSetDevice(c->device));
  zav1_stream_t s = (zav1_stream_t)c->stream;
  rc = zav1_md_fast_satd(v->d_src, v->d_recon, v->stride, v->plane_w, v->plane_h, v->bd, v->w, v->h, v->d_units,
                         v->n_units, v->cands, v->n_cands, (uint32_t*)satd, s);
  if (!rc)
    rc = zav1_md_fast_select((const uint32_t*)satd, v->d_rate, v->lambda, v->n_units, v->n_cands, v->k,
                             (const int32_t*)v->d_units, (int)(sizeof(zav1_md_unit) / 4), v->w, v->h, v->types,
                             v->n_types, 1, v->d_sel, v->d_sel_cost, (zav1_txunit_job*)jobs, s);
  if (!rc)
    rc = zav1_md_predict_survivors(v->d_recon, v->stride, v->plane_w, v->plane_h, v->bd, v->w, v->h, v->d_units,
                                   v->n_units, v->cands, v->n_cands, v->d_sel, v->k, (uint16_t*)pred, s);
  if (!rc)
    rc = zav1_txunit_batch(v->d_src, v->stride, v->plane_w, v->plane_h, v->bd, ts, (const uint16_t*)pred,
                           (size_t)v->n_units * v->k * v->w * v->h, (const zav1_txunit_job*)jobs, (int)n_jobs,
                           v->qindex, v->txu_flags, v->d_dist, v->d_eob, v->d_cul, v->d_qcoeff, v->d_recon_out, s);
  if (rc && c->err == ZAV1_OK) c->err = rc;
  return rc;
}

int zav1_wave_collect(zav1_ctx* c) {
  if (!c) return ZAV1_EINVAL;
  ZCHECK(hipSetDevice(c->device));
  ZCHECK(hipStreamSynchronize(c->stream));
  const int rc = c->err;
  c->err = ZAV1_OK;
  return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Partition pre-pass (pd0.hip)
// ---------------------------------------------------------------------------
extern "C" {

int zav1_pd0_nodes_per_unit(const zav1_pd0_cfg* c) {
  if (!c) return ZAV1_EINVAL;
  return c->min_sq <= 4 ? 341 : 85;
}

int zav1_pd0_run(const zav1_pd0_cfg* c, const void* d_plane, const zav1_coeff_costs* d_costs, zav1_pd0_sb* d_sb,
                 zav1_pd0_node* d_nodes, uint8_t* d_leaf_map, zav1_stream_t s) {
  static_assert(sizeof(zav1_pd0_sb) == sizeof(zav1::Pd0SbInfo) && sizeof(zav1_pd0_sb) == 176, "pd0 sb layout");
  static_assert(sizeof(zav1_pd0_node) == sizeof(zav1::Pd0Node) && sizeof(zav1_pd0_node) == 16, "pd0 node layout");
  if (!c || c->width <= 0 || c->height <= 0 || c->width % 64 || c->height % 64 || c->n_images < 1 ||
      (c->bd != 8 && c->bd != 10) || (c->level != 0 && c->level != 1 && c->level != 5) || c->qp < 0 || c->qp > 63 ||
      c->qindex < 0 || c->qindex > 255 || c->ires < 0 || c->ires > 4 || (c->max_tx != 32 && c->max_tx != 64) ||
      (c->min_sq != 4 && c->min_sq != 8) || (c->min_sq == 4 && c->level != 1) || c->coeff_rate_est_lvl < 0 ||
      c->coeff_rate_est_lvl > 2 || c->tile_w < 0 || c->tile_h < 0 || c->tile_w % 64 || c->tile_h % 64 ||
      c->lambda == 0 || (long)c->width * c->height * c->n_images > (1L << 31) || !d_plane || !d_sb || !d_nodes ||
      !d_leaf_map || (c->level == 1 && !d_costs))
    return ZAV1_EINVAL;
  for (int i = 0; i < 4; i++)
    if (c->split_bits[i] < 0) return ZAV1_EINVAL;
  const zav1::QuantRow* rows = nullptr;
  const uint16_t *iscan = nullptr, *fscan = nullptr;
  int rc = quant_rows(&rows);
  if (!rc) rc = iscan_table(&iscan);
  if (!rc) rc = fscan_table(0, &fscan);
  if (rc) return rc;
  zav1::Pd0Params p{};
  p.width = c->width, p.height = c->height, p.bd = c->bd, p.level = c->level;
  p.n_sb = (c->width / 64) * (c->height / 64) * c->n_images;
  p.min_sq = c->min_sq, p.max_tx = c->max_tx, p.cap_max_block = c->cap_max_block;
  p.coeff_rate_est_lvl = c->coeff_rate_est_lvl, p.tile_w = c->tile_w, p.tile_h = c->tile_h, p.ires = c->ires;
  p.nps = c->min_sq <= 4 ? 341 : 85;
  // qp-based threshold scaling (pd0.rs:185)
  if (c->qp >= 46) {
    const double ex = -((double)(c->qp > 40 ? c->qp : 40) - 35.0) / 10.0;
    p.qw = (uint32_t)((1.05 - std::exp(ex)) * 10000.0);
    p.qwd = 10000;
  } else {
    p.qw = (uint32_t)(c->qp > 10 ? c->qp : 10);
    p.qwd = 63;
  }
  for (int i = 0; i < 4; i++) p.split_bits[i] = c->split_bits[i];
  p.none_bits = c->none_bits, p.skip0_bits = c->skip0_bits;
  for (int i = 0; i < 3; i++) p.txt_dc[i] = c->txt_dc[i];
  p.lambda = c->lambda;
  // the pass quantizes with the 8-bit rows: qindex at LVL_1, qindex + 8 at
  // LVL_0 / LVL_5 (lpd0_qp_offset)
  const int qi = c->level == 1 ? c->qindex : std::min(c->qindex + 8, 255);
  ZCHECK(zav1::pd0_run(d_plane, p, rows + qi, iscan, fscan, reinterpret_cast<const int32_t*>(d_costs),
                       reinterpret_cast<zav1::Pd0SbInfo*>(d_sb), reinterpret_cast<zav1::Pd0Node*>(d_nodes),
                       d_leaf_map, S(s)));
  return ZAV1_OK;
}

}  // extern "C"
