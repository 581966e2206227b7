// AV1 bitstream of a still image's luma decisions, on the host: a
// monochrome key frame (reduced still-picture sequence header, one frame
// OBU) whose tiles code the partition, the intra mode info and the quantized
// coefficients of every leaf with the AV1 arithmetic coder and adaptive
// CDFs.  Tiles are independent (fresh CDFs and contexts), so they are coded
// in parallel on host threads.
//
// Reference: the entropy walk of rust/crates/svtav1-encoder/src/pipeline.rs
// :2542-2811 (fresh FrameContext and coefficient CDFs per tile, tiles in
// raster order, size-prefixed), the range coder
// rust/crates/svtav1-entropy/src/range_coder.rs (OdEcEnc), CDF adaptation
// cdf.rs:64 update_cdf, the coefficient writer coeff_c.rs:924
// write_coeffs_txb_1d and the OBU writer obu.rs (sequence / frame header,
// tile info, tile group).  The syntax is the AV1 specification's
// (sections 5.5 sequence header, 5.9 frame header, 5.11 tile group).
//
// Scope: monochrome (luma only; NumPlanes == 1), square leaves 8..64 whose
// transform is the block (TX_MODE_LARGEST), no segmentation, delta q,
// palette, intra block copy, CDEF, restoration or loop filter (levels 0),
// image sides multiples of the superblock, uniform tiles whose counts are
// powers of two.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/zav1_dsp.h"
#include "av1_coef_cdfs.h"
#include "av1_part_cdfs.h"
#include "zav1_kernels.h"

namespace {

// ---------------------------------------------------------------------------
// Arithmetic coder (the AV1 multi-symbol range coder, 15-bit probabilities;
// bytes collected with their carries and resolved at the end)
// ---------------------------------------------------------------------------
constexpr unsigned kProbShift = 6, kMinProb = 4;

struct RangeEnc {
  std::vector<uint16_t> pre;  // output bytes before carry propagation
  uint32_t low = 0;
  uint32_t rng = 0x8000;
  int cnt = -9;

  void normalize(uint32_t l, uint32_t r) {
    const int d = 15 - (31 - __builtin_clz(r));  // leading zeros of r in 16 bits
    int c = cnt, s = c + d;
    if (s >= 0) {
      c += 16;
      uint32_t m = (1u << c) - 1;
      if (s >= 8) {
        pre.push_back((uint16_t)(l >> c));
        l &= m;
        c -= 8;
        m >>= 8;
      }
      pre.push_back((uint16_t)(l >> c));
      s = c + d - 24;
      l &= m;
    }
    low = l << d;
    rng = r << d;
    cnt = s;
  }
  // symbol s of an n-symbol inverse CDF (icdf[i] = 32768 - P(X <= i))
  void symbol(int s, const uint16_t* icdf, int n) {
    const uint32_t fl = s > 0 ? icdf[s - 1] : 32768u, fh = icdf[s];
    uint32_t l = low, r = rng;
    const int N = n - 1;
    if (fl < 32768u) {
      const uint32_t u = (((r >> 8) * (fl >> kProbShift)) >> (7 - kProbShift)) + kMinProb * (uint32_t)(N - (s - 1));
      const uint32_t v = (((r >> 8) * (fh >> kProbShift)) >> (7 - kProbShift)) + kMinProb * (uint32_t)(N - s);
      l += r - u;
      r = u - v;
    } else {
      r -= (((r >> 8) * (fh >> kProbShift)) >> (7 - kProbShift)) + kMinProb * (uint32_t)(N - s);
    }
    normalize(l, r);
  }
  void bit(int b) {  // equiprobable bool (the specification's L(1))
    const uint32_t f = 16384;
    uint32_t l = low, r = rng;
    const uint32_t v = (((r >> 8) * (f >> kProbShift)) >> (7 - kProbShift)) + kMinProb;
    if (b) l += r - v;
    r = b ? v : r - v;
    normalize(l, r);
  }
  void literal(uint32_t v, int n) {
    for (int i = n - 1; i >= 0; i--) bit((v >> i) & 1);
  }
  std::vector<uint8_t> done() {
    const uint64_t l = low;
    int c = cnt, s = 10 + c;
    const uint64_t m = 0x3FFF;
    uint64_t e = ((l + m) & ~m) | (m + 1);
    std::vector<uint16_t> b = pre;
    if (s > 0) {
      uint64_t n = (1ull << (c + 16)) - 1;
      do {
        b.push_back((uint16_t)(e >> (c + 16)));
        e &= n;
        s -= 8;
        c -= 8;
        n >>= 8;
      } while (s > 0);
    }
    std::vector<uint8_t> out(b.size());
    uint32_t carry = 0;
    for (size_t i = b.size(); i-- > 0;) {
      carry += b[i];
      out[i] = (uint8_t)carry;
      carry >>= 8;
    }
    return out;
  }
};

// CDF adaptation after coding symbol `val` of an n-symbol inverse CDF whose
// adaptation counter sits at icdf[n] (cdf.rs:64)
inline void adapt(uint16_t* icdf, int val, int n) {
  const uint16_t count = icdf[n];
  const int rate = 4 + (count >> 4) + (n > 3);
  for (int i = 0; i < n - 1; i++) {
    if (i < val) icdf[i] += (uint16_t)((32768 - icdf[i]) >> rate);
    else icdf[i] -= (uint16_t)(icdf[i] >> rate);
  }
  icdf[n] += count < 32;
}

// The adaptive CDFs of one tile: the default tables (coefficient CDFs of the
// frame qindex's q context)
struct Cdfs {
  uint16_t partition[20][11];
  uint16_t skip[3][3];
  uint16_t kf_y[5][5][14];
  uint16_t angle[8][8];
  uint16_t fi[22][3];
  uint16_t fi_mode[6];
  uint16_t ext_tx[3][4][13][17];
  uint16_t txb_skip[5][13][3];
  uint16_t eob_extra[5][2][9][3];
  uint16_t dc_sign[2][3][3];
  uint16_t eob16[2][2][6], eob32[2][2][7], eob64[2][2][8], eob128[2][2][9], eob256[2][2][10], eob512[2][2][11],
      eob1024[2][2][12];
  uint16_t base_eob[5][2][4][4];
  uint16_t base[5][2][42][5];
  uint16_t br[4][2][21][5];

  void init(int qindex) {
    const int q = qindex <= 20 ? 0 : (qindex <= 60 ? 1 : (qindex <= 120 ? 2 : 3));
    std::memcpy(partition, AV1_PARTITION_CDF, sizeof(partition));
    std::memcpy(skip, AV1_SKIP_CDF, sizeof(skip));
    std::memcpy(kf_y, AV1_KF_Y_CDF, sizeof(kf_y));
    std::memcpy(angle, AV1_ANGLE_DELTA_CDF, sizeof(angle));
    std::memcpy(fi, AV1_FILTER_INTRA_CDF, sizeof(fi));
    std::memcpy(fi_mode, AV1_FILTER_INTRA_MODE_CDF, sizeof(fi_mode));
    std::memcpy(ext_tx, AV1_INTRA_EXT_TX_CDF, sizeof(ext_tx));
    std::memcpy(txb_skip, AV1_TXB_SKIP_CDF[q], sizeof(txb_skip));
    std::memcpy(eob_extra, AV1_EOB_EXTRA_CDF[q], sizeof(eob_extra));
    std::memcpy(dc_sign, AV1_DC_SIGN_CDF[q], sizeof(dc_sign));
    std::memcpy(eob16, AV1_EOB_FLAG_CDF16[q], sizeof(eob16));
    std::memcpy(eob32, AV1_EOB_FLAG_CDF32[q], sizeof(eob32));
    std::memcpy(eob64, AV1_EOB_FLAG_CDF64[q], sizeof(eob64));
    std::memcpy(eob128, AV1_EOB_FLAG_CDF128[q], sizeof(eob128));
    std::memcpy(eob256, AV1_EOB_FLAG_CDF256[q], sizeof(eob256));
    std::memcpy(eob512, AV1_EOB_FLAG_CDF512[q], sizeof(eob512));
    std::memcpy(eob1024, AV1_EOB_FLAG_CDF1024[q], sizeof(eob1024));
    std::memcpy(base_eob, AV1_COEFF_BASE_EOB_CDF[q], sizeof(base_eob));
    std::memcpy(base, AV1_COEFF_BASE_CDF[q], sizeof(base));
    std::memcpy(br, AV1_COEFF_BR_CDF[q], sizeof(br));
  }
};

struct SymWriter {
  RangeEnc ec;
  void sym(int s, uint16_t* cdf, int n) {
    ec.symbol(s, cdf, n);
    adapt(cdf, s, n);
  }
};

// ---------------------------------------------------------------------------
// Coefficient syntax (spec 5.11.39 coeffs; coeff_c.rs:924)
// ---------------------------------------------------------------------------
int lg2i(int v) { return 31 - __builtin_clz((unsigned)v); }
int tx_of(int s) { return s == 4 ? 0 : (s == 8 ? 1 : (s == 16 ? 2 : (s == 32 ? 3 : 4))); }

// intra ext-tx set of a square transform: 0 DCT only (32 and 64 points), 1
// the 7-type set (DTT4_IDTX_1DDCT: 4x4, 8x8), 2 the 5-type set (16x16); the
// symbol of each type within its set (the specification's
// Tx_Type_Intra_Inv_Set1 / Set2 inverted)
int intra_eset(int size) { return size >= 32 ? 0 : (size == 16 ? 2 : 1); }
constexpr int8_t kSym7[16] = {1, 5, 6, 4, -1, -1, -1, -1, -1, 0, 2, 3, -1, -1, -1, -1};
constexpr int8_t kSym5[16] = {1, 3, 4, 2, -1, -1, -1, -1, -1, 0, -1, -1, -1, -1, -1, -1};
constexpr uint8_t kFiDir[5] = {0, 1, 2, 6, 0};  // filter-intra mode -> intra direction
constexpr uint8_t kModeCtx[13] = {0, 1, 2, 3, 4, 4, 4, 4, 3, 0, 1, 2, 0};

struct Leaf {
  int x, y, size, mode, delta, fi, tx_type, eob;
  const int32_t* q;  // packed min(size, 32)^2 levels
};

// scan of a transform class (2D 0, vertical 1 -> scan class 1, horizontal 2)
struct ScanSet {
  std::vector<uint16_t> s[5][3];
  ScanSet() {
    for (int k = 0; k < 5; k++) {
      const int n = std::min(4 << k, 32);
      for (int c = 0; c < 3; c++) {
        s[k][c].resize(n * n);
        zav1::build_scan(tx_of(4 << k), c, s[k][c].data());
      }
    }
  }
};

int scan_class(int t) { return t < 10 ? 0 : ((t & 1) ? 2 : 1); }  // V_* vertical (1), H_* horizontal (2)
int tx_class(int t) { return t < 10 ? 0 : ((t & 1) ? 1 : 2); }    // TX_CLASS 2D 0, HORIZ 1, VERT 2

// nz-map context of a non-final position for any class (coeff_c.rs:406)
int base_ctx_cls(const uint8_t* lv, int n, int pos, int cls) {
  const int r = pos / n, c = pos % n;
  auto L = [&](int rr, int cc) -> int { return (rr < n && cc < n) ? std::min<int>(lv[rr * n + cc], 3) : 0; };
  int mag = L(r, c + 1) + L(r + 1, c);
  if (cls == 0) mag += L(r + 1, c + 1) + L(r, c + 2) + L(r + 2, c);
  else if (cls == 2) mag += L(r + 2, c) + L(r + 3, c) + L(r + 4, c);
  else mag += L(r, c + 2) + L(r, c + 3) + L(r, c + 4);
  if ((cls | pos) == 0) return 0;
  const int ctx = std::min((mag + 1) >> 1, 4);
  if (cls == 0) return ctx + (pos == 0 ? 0 : (r + c < 2 ? 1 : (r + c < 4 ? 6 : 21)));
  const int p = cls == 1 ? c : r;
  return ctx + (p == 0 ? 26 : (p == 1 ? 31 : 36));
}
int br_ctx_cls(const uint8_t* lv, int n, int pos, int cls) {
  const int r = pos / n, c = pos % n;
  auto L = [&](int rr, int cc) -> int { return (rr < n && cc < n) ? lv[rr * n + cc] : 0; };
  int mag = L(r, c + 1) + L(r + 1, c);
  mag += cls == 0 ? L(r + 1, c + 1) : (cls == 1 ? L(r, c + 2) : L(r + 2, c));
  mag = std::min((mag + 1) >> 1, 6);
  if (pos == 0) return mag;
  if ((cls == 0 && r < 2 && c < 2) || (cls == 1 && c == 0) || (cls == 2 && r == 0)) return mag + 7;
  return mag + 14;
}

void write_golomb(RangeEnc& ec, int level) {
  const int x = level + 1;
  const int length = 32 - __builtin_clz((unsigned)x);
  for (int i = 0; i < length - 1; i++) ec.bit(0);
  for (int i = length - 1; i >= 0; i--) ec.bit((x >> i) & 1);
}

// all_zero .. signs of one luma transform block equal to its block (context
// 0 for all_zero); returns the dc category for the neighbours' dc_sign
// contexts (0 none, 1 negative, 2 positive)
int write_coeffs(SymWriter& w, Cdfs& cf, const ScanSet& scans, const Leaf& lf, int qindex, int dc_ctx) {
  const int size = lf.size, txs = tx_of(size), txs_ctx = std::min(txs, 4);  // square: (sqr + sqr_up + 1) >> 1
  w.sym(lf.eob == 0, cf.txb_skip[txs_ctx][0], 2);
  if (lf.eob == 0) return 0;
  // tx type (intra; the set has several types and qindex > 0)
  const int eset = intra_eset(size);
  if (eset != 0 && qindex > 0) {
    const int dir = lf.fi < 5 ? kFiDir[lf.fi] : lf.mode;
    const int sym = eset == 1 ? kSym7[lf.tx_type] : kSym5[lf.tx_type];
    w.sym(sym, cf.ext_tx[eset][lg2i(size) - 2][dir], eset == 1 ? 7 : 5);
  }
  const int n = std::min(size, 32), N = n * n;
  const int cls = tx_class(lf.tx_type);
  const uint16_t* scan = scans.s[lg2i(size) - 2][scan_class(lf.tx_type)].data();
  // eob position token and extra bits (coeff_c.rs:606)
  const int eob = lf.eob;
  int pt, start;
  if (eob <= 2) pt = eob, start = eob;
  else {
    pt = 3 + lg2i(eob - 1) - 1;  // eob in (2^(pt-2), 2^(pt-1)]
    start = (1 << (pt - 2)) + 1;
  }
  const int ems = lg2i(n) * 2 - 4, ectx = cls == 0 ? 0 : 1;
  switch (ems) {
    case 0: w.sym(pt - 1, cf.eob16[0][ectx], 5); break;
    case 1: w.sym(pt - 1, cf.eob32[0][ectx], 6); break;
    case 2: w.sym(pt - 1, cf.eob64[0][ectx], 7); break;
    case 3: w.sym(pt - 1, cf.eob128[0][ectx], 8); break;
    case 4: w.sym(pt - 1, cf.eob256[0][ectx], 9); break;
    case 5: w.sym(pt - 1, cf.eob512[0][ectx], 10); break;
    default: w.sym(pt - 1, cf.eob1024[0][ectx], 11); break;
  }
  const int offset_bits = pt >= 3 ? pt - 2 : 0;
  if (offset_bits > 0) {
    const int extra = eob - start;
    w.sym((extra >> (offset_bits - 1)) & 1, cf.eob_extra[txs_ctx][0][pt - 3], 2);
    for (int i = 1; i < offset_bits; i++) w.ec.bit((extra >> (offset_bits - 1 - i)) & 1);
  }
  // levels, last scan position first
  uint8_t lv[1024];
  for (int i = 0; i < N; i++) lv[i] = (uint8_t)std::min(std::abs(lf.q[i]), 127);
  const int br_txs = std::min(txs_ctx, 3);
  for (int c = eob - 1; c >= 0; c--) {
    const int pos = scan[c], level = std::abs(lf.q[pos]);
    if (c == eob - 1) {
      const int ctx = c == 0 ? 0 : (c <= N / 8 ? 1 : (c <= N / 4 ? 2 : 3));
      w.sym(std::min(level, 3) - 1, cf.base_eob[txs_ctx][0][ctx], 3);
    } else {
      w.sym(std::min(level, 3), cf.base[txs_ctx][0][base_ctx_cls(lv, n, pos, cls)], 4);
    }
    if (level > 2) {
      const int range = level - 3;
      uint16_t* cdf = cf.br[br_txs][0][br_ctx_cls(lv, n, pos, cls)];
      for (int idx = 0; idx < 12; idx += 3) {
        const int k = std::min(range - idx, 3);
        w.sym(k, cdf, 4);
        if (k < 3) break;
      }
    }
  }
  // signs, DC first, and the Golomb remainders
  for (int c = 0; c < eob; c++) {
    const int32_t v = lf.q[scan[c]];
    const int level = std::abs(v);
    if (!level) continue;
    if (c == 0) w.sym(v < 0, cf.dc_sign[0][dc_ctx], 2);
    else w.ec.bit(v < 0);
    if (level > 14) write_golomb(w.ec, level - 15);
  }
  const int32_t dc = lf.q[0];
  return dc < 0 ? 1 : (dc > 0 ? 2 : 0);
}

// ---------------------------------------------------------------------------
// Tile walk
// ---------------------------------------------------------------------------
struct Frame {
  int W, H, bd, sb, qindex, mi_cols, mi_rows;
  int tile_cols_log2, tile_rows_log2;
  std::vector<int> col_starts, row_starts;  // in mi (4x4) units, tile edges
  std::vector<int32_t> leaf_at;             // per 8x8 cell: leaf index
  const std::vector<Leaf>* leaves;
  const ScanSet* scans;
};

struct TileCtx {
  int mc0, mc1, mr0, mr1;  // tile bounds in mi
  std::vector<uint8_t> mode, skip, wlog;  // per mi of the frame (only the tile's used)
  std::vector<uint8_t> above_dc, left_dc;  // per mi column / row: dc category
};

void write_leaf(const Frame& f, TileCtx& t, Cdfs& cf, SymWriter& w, const Leaf& lf) {
  const int r = lf.y / 4, c = lf.x / 4, n4 = lf.size / 4;
  const bool up = r > t.mr0, left = c > t.mc0;
  const size_t mw = (size_t)f.mi_cols;
  const int skip = lf.eob == 0;
  // skip (context: the above and left blocks' skips)
  const int sctx = (up ? t.skip[(r - 1) * mw + c] : 0) + (left ? t.skip[r * mw + c - 1] : 0);
  w.sym(skip, cf.skip[sctx], 2);
  // y mode under the neighbours' mode contexts
  const int am = up ? t.mode[(r - 1) * mw + c] : 0, lm = left ? t.mode[r * mw + c - 1] : 0;
  w.sym(lf.mode, cf.kf_y[kModeCtx[am]][kModeCtx[lm]], 13);
  if (lf.mode >= 1 && lf.mode <= 8) w.sym(lf.delta + 3, cf.angle[lf.mode - 1], 7);
  if (lf.mode == 0 && lf.size <= 32) {
    const int bs = lf.size == 8 ? 3 : (lf.size == 16 ? 6 : 9);  // BLOCK_8X8 / 16X16 / 32X32
    w.sym(lf.fi < 5, cf.fi[bs], 2);
    if (lf.fi < 5) w.sym(lf.fi, cf.fi_mode, 5);
  }
  int dcat = 0;
  if (!skip) {
    int dsum = 0;
    for (int k = 0; k < n4; k++) {
      const int a = t.above_dc[c + k], l = t.left_dc[r + k];
      dsum += (a == 1 ? -1 : (a == 2 ? 1 : 0)) + (l == 1 ? -1 : (l == 2 ? 1 : 0));
    }
    const int dc_ctx = dsum < 0 ? 1 : (dsum > 0 ? 2 : 0);
    dcat = write_coeffs(w, cf, *f.scans, lf, f.qindex, dc_ctx);
  }
  for (int k = 0; k < n4; k++) {
    t.above_dc[c + k] = (uint8_t)dcat;
    t.left_dc[r + k] = (uint8_t)dcat;
  }
  for (int y = r; y < r + n4; y++)
    for (int x = c; x < c + n4; x++) {
      t.mode[y * mw + x] = (uint8_t)lf.mode;
      t.skip[y * mw + x] = (uint8_t)skip;
      t.wlog[y * mw + x] = (uint8_t)lg2i(n4);
    }
}

void write_partition(const Frame& f, TileCtx& t, Cdfs& cf, SymWriter& w, int r, int c, int bs4) {
  if (r >= f.mi_rows || c >= f.mi_cols) return;
  const int li = f.leaf_at[(size_t)(r / 2) * (f.mi_cols / 2) + c / 2];
  const Leaf& lf = (*f.leaves)[li];
  const bool is_leaf = lf.size / 4 == bs4;
  const bool up = r > t.mr0, left = c > t.mc0;
  const size_t mw = (size_t)f.mi_cols;
  const int bsl = lg2i(bs4);  // 1 (8x8) .. 5 (128x128)
  const int above = up && t.wlog[(r - 1) * mw + c] < bsl;
  const int lft = left && t.wlog[r * mw + c - 1] < bsl;
  const int ctx = lft * 2 + above;
  const int nsym = bsl == 1 ? 4 : (bsl == 5 ? 8 : 10);
  w.sym(is_leaf ? 0 : 3, cf.partition[(bsl - 1) * 4 + ctx], nsym);  // PARTITION_NONE / PARTITION_SPLIT
  if (is_leaf) {
    write_leaf(f, t, cf, w, lf);
    return;
  }
  const int h = bs4 / 2;
  write_partition(f, t, cf, w, r, c, h);
  write_partition(f, t, cf, w, r, c + h, h);
  write_partition(f, t, cf, w, r + h, c, h);
  write_partition(f, t, cf, w, r + h, c + h, h);
}

std::vector<uint8_t> write_tile(const Frame& f, int tr, int tc) {
  TileCtx t;
  t.mc0 = f.col_starts[tc], t.mc1 = f.col_starts[tc + 1];
  t.mr0 = f.row_starts[tr], t.mr1 = f.row_starts[tr + 1];
  const size_t n = (size_t)f.mi_cols * f.mi_rows;
  t.mode.assign(n, 0);
  t.skip.assign(n, 0);
  t.wlog.assign(n, 0);
  t.above_dc.assign(f.mi_cols, 0);
  t.left_dc.assign(f.mi_rows, 0);
  Cdfs cf;
  cf.init(f.qindex);
  SymWriter w;
  const int sb4 = f.sb / 4;
  for (int r = t.mr0; r < t.mr1; r += sb4) {
    std::fill(t.left_dc.begin(), t.left_dc.end(), 0);  // left context cleared per superblock row
    for (int c = t.mc0; c < t.mc1; c += sb4) write_partition(f, t, cf, w, r, c, sb4);
  }
  return w.ec.done();
}

// ---------------------------------------------------------------------------
// Headers
// ---------------------------------------------------------------------------
struct BitWriter {
  std::vector<uint8_t> b;
  int nbits = 0;
  void put(uint32_t v, int n) {
    for (int i = n - 1; i >= 0; i--) {
      if ((nbits & 7) == 0) b.push_back(0);
      if ((v >> i) & 1) b.back() |= (uint8_t)(0x80 >> (nbits & 7));
      nbits++;
    }
  }
  void trailing() {  // trailing_bits: a 1 then zeros to the byte boundary
    put(1, 1);
    while (nbits & 7) put(0, 1);
  }
  void align() {
    while (nbits & 7) put(0, 1);
  }
};

void leb128(std::vector<uint8_t>& o, size_t v) {
  do {
    uint8_t byte = v & 0x7f;
    v >>= 7;
    if (v) byte |= 0x80;
    o.push_back(byte);
  } while (v);
}
void obu(std::vector<uint8_t>& o, int type, const std::vector<uint8_t>& payload) {
  o.push_back((uint8_t)((type << 3) | 2));  // obu_has_size_field
  leb128(o, payload.size());
  o.insert(o.end(), payload.begin(), payload.end());
}
int tile_log2(int blk, int target) {
  int k = 0;
  while ((blk << k) < target) k++;
  return k;
}

std::vector<uint8_t> sequence_header(const Frame& f) {
  BitWriter w;
  w.put(0, 3);   // seq_profile (Main: 8 / 10 bits, monochrome allowed)
  w.put(1, 1);   // still_picture
  w.put(1, 1);   // reduced_still_picture_header
  w.put(31, 5);  // seq_level_idx[0]: no level constraints
  const int wb = std::max(1, 32 - __builtin_clz((unsigned)(f.W - 1))), hb = std::max(1, 32 - __builtin_clz((unsigned)(f.H - 1)));
  w.put(wb - 1, 4);
  w.put(hb - 1, 4);
  w.put(f.W - 1, wb);
  w.put(f.H - 1, hb);
  w.put(f.sb == 128, 1);  // use_128x128_superblock
  w.put(1, 1);            // enable_filter_intra
  w.put(1, 1);            // enable_intra_edge_filter
  w.put(0, 1);            // enable_superres
  w.put(0, 1);            // enable_cdef
  w.put(0, 1);            // enable_restoration
  w.put(f.bd == 10, 1);   // high_bitdepth
  w.put(1, 1);            // mono_chrome
  w.put(0, 1);            // color_description_present_flag
  w.put(0, 1);            // color_range
  w.put(0, 1);            // film_grain_params_present
  w.trailing();
  return w.b;
}

std::vector<uint8_t> frame_obu(const Frame& f, const std::vector<std::vector<uint8_t>>& tiles) {
  BitWriter w;
  // uncompressed_header of a shown key frame under the reduced still-picture
  // header (frame size, refresh flags, primary_ref_frame, error resilience
  // and the end-of-frame CDF update are implied)
  w.put(0, 1);  // disable_cdf_update
  w.put(0, 1);  // allow_screen_content_tools
  w.put(0, 1);  // render_and_frame_size_different
  // tile_info: uniform spacing
  const int sb_shift = f.sb == 128 ? 5 : 4, sb_size = sb_shift + 2;
  const int sb_cols = (f.mi_cols + (1 << sb_shift) - 1) >> sb_shift, sb_rows = (f.mi_rows + (1 << sb_shift) - 1) >> sb_shift;
  const int min_c = tile_log2(4096 >> sb_size, sb_cols);
  const int max_c = tile_log2(1, std::min(sb_cols, 64)), max_r = tile_log2(1, std::min(sb_rows, 64));
  const int min_tiles = std::max(min_c, tile_log2((4096 * 2304) >> (2 * sb_size), sb_rows * sb_cols));
  w.put(1, 1);  // uniform_tile_spacing_flag
  for (int k = min_c; k < max_c; k++) {
    w.put(k < f.tile_cols_log2, 1);  // increment_tile_cols_log2
    if (k >= f.tile_cols_log2) break;
  }
  const int min_r = std::max(min_tiles - f.tile_cols_log2, 0);
  for (int k = min_r; k < max_r; k++) {
    w.put(k < f.tile_rows_log2, 1);  // increment_tile_rows_log2
    if (k >= f.tile_rows_log2) break;
  }
  const int n_tiles = (int)(f.col_starts.size() - 1) * (int)(f.row_starts.size() - 1);
  if (f.tile_cols_log2 > 0 || f.tile_rows_log2 > 0) {
    w.put(0, f.tile_cols_log2 + f.tile_rows_log2);  // context_update_tile_id
    w.put(3, 2);                                    // tile_size_bytes_minus_1: 4-byte sizes
  }
  // quantization_params
  w.put(f.qindex, 8);  // base_q_idx
  w.put(0, 1);         // DeltaQYDc: delta_coded
  w.put(0, 1);         // using_qmatrix
  w.put(0, 1);         // segmentation_enabled
  w.put(0, 1);         // delta_q_present (base_q_idx > 0)
  // loop_filter_params: levels 0
  w.put(0, 6);
  w.put(0, 6);
  w.put(0, 3);  // loop_filter_sharpness
  w.put(0, 1);  // loop_filter_delta_enabled
  w.put(0, 1);  // tx_mode_select = 0: TX_MODE_LARGEST
  w.put(0, 1);  // reduced_tx_set
  w.align();    // frame OBU: byte_alignment before the tile group
  // tile_group_obu
  if (n_tiles > 1) {
    w.put(0, 1);  // tile_start_and_end_present_flag
    w.align();
  }
  std::vector<uint8_t> out = w.b;
  for (int i = 0; i < n_tiles; i++) {
    const std::vector<uint8_t>& t = tiles[i];
    if (i + 1 < n_tiles) {
      const uint32_t s = (uint32_t)t.size() - 1;
      for (int k = 0; k < 4; k++) out.push_back((uint8_t)(s >> (8 * k)));  // le(4)
    }
    out.insert(out.end(), t.begin(), t.end());
  }
  return out;
}

}  // namespace

namespace zav1 {

// One still into `stream` (ZAV1_OK) or an error code; shared by
// zav1_write_still and the wavefront's batch writer.
int write_still_vec(const zav1_still_desc* d, std::vector<uint8_t>& stream, std::string* why) {
  auto fail = [why](int rc, const char* what) {
    if (why) *why = what;
    return rc;
  };
  static_assert(sizeof(zav1_still_leaf) == 32, "still leaf layout");
  if (!d || (d->bd != 8 && d->bd != 10) || (d->sb != 64 && d->sb != 128) || d->width <= 0 ||
      d->height <= 0 || d->width % d->sb || d->height % d->sb || d->qindex < 1 || d->qindex > 255 ||
      d->n_leaves < 1 || !d->leaves || !d->levels || d->tile_cols < 1 || d->tile_rows < 1 ||
      (d->tile_cols & (d->tile_cols - 1)) || (d->tile_rows & (d->tile_rows - 1)) || d->width > 65536 ||
      d->height > 65536 || d->threads < 0)
    return fail(ZAV1_EINVAL, "frame description");
  Frame f;
  f.W = d->width, f.H = d->height, f.bd = d->bd, f.sb = d->sb, f.qindex = d->qindex;
  f.mi_cols = f.W / 4, f.mi_rows = f.H / 4;
  f.tile_cols_log2 = __builtin_ctz((unsigned)d->tile_cols), f.tile_rows_log2 = __builtin_ctz((unsigned)d->tile_rows);
  // tile layout as the frame header signals it (uniform spacing)
  const int sb_shift = f.sb == 128 ? 5 : 4, sb_size = sb_shift + 2;
  const int sb_cols = f.mi_cols >> sb_shift, sb_rows = f.mi_rows >> sb_shift;
  const int min_c = tile_log2(4096 >> sb_size, sb_cols), max_c = tile_log2(1, std::min(sb_cols, 64));
  const int max_r = tile_log2(1, std::min(sb_rows, 64));
  const int min_tiles = std::max(min_c, tile_log2((4096 * 2304) >> (2 * sb_size), sb_rows * sb_cols));
  const int min_r = std::max(min_tiles - f.tile_cols_log2, 0);
  if (f.tile_cols_log2 < min_c || f.tile_cols_log2 > max_c || f.tile_rows_log2 < min_r || f.tile_rows_log2 > max_r)
    return fail(ZAV1_ENOTSUP, "tile layout outside the frame's tile limits");
  const int tws = (sb_cols + (1 << f.tile_cols_log2) - 1) >> f.tile_cols_log2;
  const int ths = (sb_rows + (1 << f.tile_rows_log2) - 1) >> f.tile_rows_log2;
  for (int s = 0; s < sb_cols; s += tws) f.col_starts.push_back(s << sb_shift);
  f.col_starts.push_back(f.mi_cols);
  for (int s = 0; s < sb_rows; s += ths) f.row_starts.push_back(s << sb_shift);
  f.row_starts.push_back(f.mi_rows);
  // leaves and their 8x8 cells: every cell covered exactly once
  std::vector<Leaf> leaves(d->n_leaves);
  const int cw = f.W / 8, ch = f.H / 8;
  f.leaf_at.assign((size_t)cw * ch, -1);
  for (int i = 0; i < d->n_leaves; i++) {
    const zav1_still_leaf& s = d->leaves[i];
    Leaf& l = leaves[i];
    l.x = s.x, l.y = s.y, l.size = s.size, l.mode = s.mode, l.delta = s.delta, l.fi = s.fi_mode;
    l.tx_type = s.tx_type, l.eob = s.eob, l.q = d->levels + s.level_off;
    const int n = std::min<int>(l.size, 32);
    if ((l.size != 8 && l.size != 16 && l.size != 32 && l.size != 64) || l.x < 0 || l.y < 0 || l.x % l.size ||
        l.y % l.size || l.x + l.size > f.W || l.y + l.size > f.H || l.mode > 12 || l.delta < -3 || l.delta > 3 ||
        (l.delta != 0 && (l.mode < 1 || l.mode > 8)) || l.fi > 5 || (l.fi < 5 && (l.mode != 0 || l.size > 32)) ||
        l.tx_type > 15 || l.eob < 0 || l.eob > n * n)
      return fail(ZAV1_EINVAL, "leaf position, size, mode or eob");
    const int eset = intra_eset(l.size);
    if (l.eob > 0 && (eset == 0 ? l.tx_type != 0 : (eset == 1 ? kSym7[l.tx_type] : kSym5[l.tx_type]) < 0))
      return fail(ZAV1_EINVAL, "transform type outside the leaf size's intra set");
    for (int y = l.y / 8; y < (l.y + l.size) / 8; y++)
      for (int x = l.x / 8; x < (l.x + l.size) / 8; x++) {
        int32_t& a = f.leaf_at[(size_t)y * cw + x];
        if (a >= 0) return fail(ZAV1_EINVAL, "overlapping leaves");
        a = i;
      }
  }
  for (int32_t a : f.leaf_at)
    if (a < 0) return fail(ZAV1_EINVAL, "cells no leaf covers");
  // each leaf's levels must end at its eob (a tight eob: the last coded
  // position nonzero)
  static const ScanSet scans;
  for (const Leaf& l : leaves)
    if (l.eob > 0) {
      const uint16_t* sc = scans.s[lg2i(l.size) - 2][scan_class(l.tx_type)].data();
      if (l.q[sc[l.eob - 1]] == 0) return fail(ZAV1_EINVAL, "levels whose last nonzero is not at the eob");
    }
  f.leaves = &leaves;
  f.scans = &scans;
  const int tcn = (int)f.col_starts.size() - 1, trn = (int)f.row_starts.size() - 1, nt = tcn * trn;
  std::vector<std::vector<uint8_t>> tiles(nt);
  int threads = d->threads > 0 ? d->threads : (int)std::max(1u, std::thread::hardware_concurrency());
  threads = std::min(threads, nt);
  if (threads <= 1) {
    for (int i = 0; i < nt; i++) tiles[i] = write_tile(f, i / tcn, i % tcn);
  } else {
    std::vector<std::thread> pool;
    for (int k = 0; k < threads; k++)
      pool.emplace_back([&, k] {
        for (int i = k; i < nt; i += threads) tiles[i] = write_tile(f, i / tcn, i % tcn);
      });
    for (auto& th : pool) th.join();
  }
  stream.clear();
  obu(stream, 2, {});  // temporal delimiter
  obu(stream, 1, sequence_header(f));
  obu(stream, 6, frame_obu(f, tiles));
  return ZAV1_OK;
}

}  // namespace zav1

namespace zav1 {
void set_last_error(const std::string& what);  // zav1_last_error()'s text on this thread
}

extern "C" int zav1_write_still(const zav1_still_desc* d, uint8_t* out, size_t cap, size_t* size) {
  if (!size) return ZAV1_EINVAL;
  std::vector<uint8_t> stream;
  std::string why;
  const int rc = zav1::write_still_vec(d, stream, &why);
  if (rc) {
    zav1::set_last_error("zav1_write_still: " + why);
    return rc;
  }
  *size = stream.size();
  if (!out || cap < stream.size()) return ZAV1_ENOMEM;
  std::memcpy(out, stream.data(), stream.size());
  return ZAV1_OK;
}
