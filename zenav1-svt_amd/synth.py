"""Seeded synthetic still-image content for the benchmark workloads.

All content is generated once on the host (integer / deterministic double
math) and uploaded, so every backend sees identical samples.  Seeds are
config_index * 1000 + image_index.  Small parameter draws use
splitmix64 -> xorshift64* (multiplier 0x2545F4914F6CDD1D, the generator the
reference's differential tests use); bulk noise uses numpy's PCG64 seeded
from the same splitmix64 stream.
"""
import numpy as np

M64 = (1 << 64) - 1


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & M64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return x, z ^ (z >> 31)


class XorShift64Star:
    def __init__(self, seed):
        _, s = splitmix64(seed)
        self.s = s or 1

    def next(self):
        x = self.s
        x ^= x >> 12
        x ^= (x << 25) & M64
        x ^= x >> 27
        self.s = x
        return (x * 0x2545F4914F6CDD1D) & M64

    def uniform(self):
        return (self.next() >> 11) / float(1 << 53)

    def randint(self, lo, hi):
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1))


def _np_rng(seed):
    _, s = splitmix64(seed ^ 0xA5A5A5A5)
    return np.random.Generator(np.random.PCG64(s))


def spectral_noise(h, w, alpha, seed):
    """Real field with power spectrum ~ 1/f^alpha (seeded random phases)."""
    rng = _np_rng(seed)
    fy = np.fft.fftfreq(h)[:, None]
    fx = np.fft.rfftfreq(w)[None, :]
    f = np.sqrt(fx * fx + fy * fy)
    f[0, 0] = 1.0
    amp = f ** (-alpha / 2.0)
    amp[0, 0] = 0.0
    phase = rng.uniform(0, 2 * np.pi, amp.shape)
    spec = amp * np.exp(1j * phase)
    field = np.fft.irfft2(spec, s=(h, w))
    field -= field.min()
    mx = field.max()
    return field / (mx if mx > 0 else 1.0)


def _draw_lines(img, n, wmin, wmax, lo, hi, rnd):
    h, w = img.shape
    yy, xx = np.mgrid[0:h, 0:w]
    for _ in range(n):
        cx, cy = rnd.uniform() * w, rnd.uniform() * h
        th = rnd.uniform() * np.pi
        width = rnd.randint(wmin, wmax)
        val = rnd.randint(lo, hi)
        d = np.abs((xx - cx) * np.sin(th) - (yy - cy) * np.cos(th))
        img[d < width / 2.0] = val


def config2_luma(h=2048, w=2048, seed=2000):
    """1/f^2 spectral noise scaled to [64, 940] plus 30 random-orientation
    hard lines of width 1-4 px (10-bit)."""
    f = spectral_noise(h, w, 2.0, seed)
    img = np.round(64 + f * (940 - 64)).astype(np.int32)
    _draw_lines(img, 30, 1, 4, 64, 940, XorShift64Star(seed))
    return np.clip(img, 0, 1023).astype(np.uint16)


def config1_luma(h=512, w=512, seed=1000):
    """Sum of 3 seeded linear gradients + 20 rectangles U[16,235] + N(0, 4^2)."""
    rnd = XorShift64Star(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    img = np.zeros((h, w))
    for _ in range(3):
        gx, gy = rnd.uniform() - 0.5, rnd.uniform() - 0.5
        img += (gx * xx / w + gy * yy / h) * 120 + 40
    img /= 3.0
    img += 60
    for _ in range(20):
        x0, y0 = rnd.randint(0, w - 8), rnd.randint(0, h - 8)
        x1, y1 = min(w, x0 + rnd.randint(8, w // 3)), min(h, y0 + rnd.randint(8, h // 3))
        img[y0:y1, x0:x1] = rnd.randint(16, 235)
    img += _np_rng(seed).normal(0, 4, (h, w))
    return np.clip(np.round(img), 0, 255).astype(np.uint8)


def config3_recon(h=4096, w=4096, seed=3000):
    """Loop-filter input: 8-bit 1/f noise source, and a reconstruction with
    8x8 mean-flattening on 30% of the 8x8 blocks (blocky artifacts)."""
    f = spectral_noise(h, w, 1.0, seed)
    src = np.clip(np.round(16 + f * 219), 0, 255).astype(np.uint8)
    rng = _np_rng(seed + 1)
    rec = src.astype(np.int32).copy()
    bh, bw = h // 8, w // 8
    mask = rng.uniform(size=(bh, bw)) < 0.30
    blocks = rec.reshape(bh, 8, bw, 8)
    means = np.round(blocks.mean(axis=(1, 3))).astype(np.int32)
    flat = np.repeat(np.repeat(means, 8, 0), 8, 1)
    m = np.repeat(np.repeat(mask, 8, 0), 8, 1)
    rec[m] = flat[m]
    rec = rec + rng.integers(-2, 3, rec.shape)
    return src, np.clip(rec, 0, 255).astype(np.uint8)


def config4_tile_column(h=4096, w=2048, seed=4000):
    """One tile column (2048 wide, 4096 high) of the 8192x4096 10-bit frame:
    1/f noise scaled to [64, 940] with 25% of the area as screen regions
    (256x256 cells: flat fills from a palette of at most 8 colours, with
    1-px glyph strokes in palette colours)."""
    f = spectral_noise(h, w, 1.0, seed)
    img = np.round(64 + f * (940 - 64)).astype(np.int32)
    rnd = XorShift64Star(seed + 1)
    cs = 256
    for cy in range(0, h, cs):
        for cx in range(0, w, cs):
            if rnd.uniform() >= 0.25:
                continue
            pal = [rnd.randint(64, 940) for _ in range(rnd.randint(2, 8))]
            cell = img[cy:cy + cs, cx:cx + cs]
            cell[:] = pal[0]
            for _ in range(rnd.randint(4, 12)):  # flat fills
                y0, x0 = rnd.randint(0, cs - 16), rnd.randint(0, cs - 16)
                y1, x1 = min(cs, y0 + rnd.randint(8, 96)), min(cs, x0 + rnd.randint(8, 128))
                cell[y0:y1, x0:x1] = pal[rnd.randint(0, len(pal) - 1)]
            for _ in range(rnd.randint(20, 60)):  # glyph strokes, 1 px wide
                y0, x0 = rnd.randint(0, cs - 12), rnd.randint(0, cs - 12)
                v = pal[rnd.randint(0, len(pal) - 1)]
                if rnd.uniform() < 0.5:
                    cell[y0, x0:x0 + rnd.randint(3, 12)] = v
                else:
                    cell[y0:y0 + rnd.randint(3, 12), x0] = v
    return np.clip(img, 0, 1023).astype(np.uint16)


def serving_mix(n, seed=5000):
    """Requests of the serving mix: (height, width, bd, qp, preset), area
    log-uniform in 256^2 .. 4096x3072, aspect 1:1 .. 16:9 (either
    orientation), 50/50 8 / 10 bits, qp uniform in 20..50, preset uniform in
    {4, 6, 8}; sides rounded to multiples of 64."""
    rnd = XorShift64Star(seed)
    out = []
    lo, hi = np.log(256.0 * 256.0), np.log(4096.0 * 3072.0)
    for _ in range(n):
        area = np.exp(lo + (hi - lo) * rnd.uniform())
        aspect = 1.0 + (16.0 / 9.0 - 1.0) * rnd.uniform()
        w = np.sqrt(area * aspect)
        h = area / w
        if rnd.uniform() < 0.5:
            w, h = h, w
        w = int(max(64, min(4096, round(w / 64) * 64)))
        h = int(max(64, min(4096, round(h / 64) * 64)))
        bd = 8 if rnd.uniform() < 0.5 else 10
        qp = rnd.randint(20, 50)
        preset = (4, 6, 8)[rnd.randint(0, 2)]
        out.append((h, w, bd, qp, preset))
    return out


def serving_images(requests, seed=5000):
    """Content of the serving requests: crops, at seeded offsets, of one
    4096x4096 config-2 style field (1/f^2 noise + lines), at each request's
    size and bit depth (8-bit content as the 10-bit field >> 2)."""
    pool = config2_luma(4096, 4096, seed=seed)
    rnd = XorShift64Star(seed + 7)
    out = []
    for req in requests:
        h, w, bd = req[:3]
        y0, x0 = rnd.randint(0, 4096 - h), rnd.randint(0, 4096 - w)
        img = np.ascontiguousarray(pool[y0:y0 + h, x0:x0 + w])
        out.append((img >> 2).astype(np.uint16) if bd == 8 else img)
    return out


def qp_to_qindex(qp):
    """CQP qp -> qindex (rust/crates/svtav1-encoder/src/rate_control.rs:129)."""
    if qp < 62:
        return 4 * qp
    return 249 if qp == 62 else 255


def config3_chroma(h=2048, w=2048, seed=3000):
    """4:2:0 chroma planes of the loop-filter input: 1/f noise sources with
    4x4 mean-flattening on 30% of the 4x4 blocks in the reconstructions."""
    out = []
    for k in range(2):
        s = seed + 10 + 2 * k
        f = spectral_noise(h, w, 1.0, s)
        src = np.clip(np.round(64 + f * 128), 0, 255).astype(np.uint8)
        rng = _np_rng(s + 1)
        rec = src.astype(np.int32).copy()
        bh, bw = h // 4, w // 4
        mask = rng.uniform(size=(bh, bw)) < 0.30
        means = np.round(rec[:bh * 4, :bw * 4].reshape(bh, 4, bw, 4).mean(axis=(1, 3))).astype(np.int32)
        flat = np.repeat(np.repeat(means, 4, 0), 4, 1)
        m = np.repeat(np.repeat(mask, 4, 0), 4, 1)
        sub = rec[:bh * 4, :bw * 4]
        sub[m] = flat[m]
        rec = rec + rng.integers(-1, 2, rec.shape)
        out.append((src, np.clip(rec, 0, 255).astype(np.uint8)))
    return out


def partition_geometry(mi_rows, mi_cols, seed, sb=64, skip_frac=0.0):
    """Synthetic per-4x4 (mi) block / transform geometry of a frame: each
    SB is split recursively (none / quad / horizontal / vertical halves) into
    blocks aligned to their own size, and each block carries transforms of
    its own size or half of it (capped at 64).  Returns a dict of row-major
    mi_rows x mi_cols arrays: block_id u32, bw, bh, tw, th (px) and, when
    skip_frac > 0, skip u8 (a block's skip flag)."""
    rnd = XorShift64Star(seed)
    g = {k: np.zeros((mi_rows, mi_cols), np.uint8) for k in ('bw', 'bh', 'tw', 'th', 'skip')}
    g['block_id'] = np.zeros((mi_rows, mi_cols), np.uint32)
    nid = [0]

    def leaf(r, c, h, w):
        tdep = rnd.randint(0, 1) if min(w, h) >= 8 else 0
        tw, th = min(w >> tdep, 64), min(h >> tdep, 64)
        sl = (slice(r // 4, (r + h) // 4), slice(c // 4, (c + w) // 4))
        g['block_id'][sl] = nid[0]
        g['bw'][sl] = min(w, 255)
        g['bh'][sl] = min(h, 255)
        g['tw'][sl] = tw
        g['th'][sl] = th
        g['skip'][sl] = 1 if rnd.uniform() < skip_frac else 0
        nid[0] += 1

    def split(r, c, s):
        if r >= mi_rows * 4 or c >= mi_cols * 4:
            return
        u = rnd.uniform()
        if s == 4:
            leaf(r, c, 4, 4)
        elif u < 0.30 or (s == 64 and u < 0.05):
            leaf(r, c, s, s)
        elif u < 0.70:
            h = s // 2
            for dr in (0, h):
                for dc in (0, h):
                    split(r + dr, c + dc, h)
        elif u < 0.85:
            leaf(r, c, s // 2, s)
            if r + s // 2 < mi_rows * 4:
                leaf(r + s // 2, c, s // 2, s)
        else:
            leaf(r, c, s, s // 2)
            if c + s // 2 < mi_cols * 4:
                leaf(r, c + s // 2, s, s // 2)

    for r in range(0, mi_rows * 4, sb):
        for c in range(0, mi_cols * 4, sb):
            split(r, c, sb)
    if skip_frac <= 0:
        del g['skip']
    return g


def key_frame_lf_levels(ac_q, bd):
    """Deblocking levels of a key frame from the AC quantizer step
    (rust/crates/svtav1-encoder/src/deblock.rs:55): [luma V, luma H, U, V]."""
    if bd == 8:
        fg = (ac_q * 17563 - 421574 + (1 << 17)) >> 18
    else:
        fg = ((ac_q * 20723 + 4060632 + (1 << 19)) >> 20) - 4
    fc = int(fg / 2) if fg >= 0 else -((-fg) // 2)
    y = min(max(fg, 0), 63)
    uv = min(max(fc, 0), 63)
    return [y, y, uv, uv]


# ---------------------------------------------------------------------------
# Synthetic coefficient rate tables for the RDOQ trellis.  The trellis prices
# with tables its caller builds from the frame's CDFs
# (rust/crates/svtav1-encoder/src/quant.rs:514-:632); the default CDFs are not
# part of this repository, so per-symbol costs are drawn at random and the
# tables' secondary entries derived exactly as that builder does: the base
# cost deltas (entries 4..7 of base_cost), the cumulative base-range costs
# lps[0..12] from four-symbol rates, and their deltas lps[13..25].
# ---------------------------------------------------------------------------
TXB_WORDS, EOB_WORDS = 944, 22
# word offsets of the TxbCosts fields (quant.rs:459)
SKIP, BASE_EOB, BASE, EOB_EXTRA, DC_SIGN, LPS = 0, 26, 38, 374, 392, 398


def symbol_costs(rng, n, scale):
    """Costs of an n-symbol alphabet in 1/512 bit: -log2 of a random
    distribution with every probability >= 4/32768."""
    p = rng.dirichlet(np.full(n, 0.7))
    p = np.maximum(p, 4 / 32768)
    p /= p.sum()
    return np.round(-np.log2(p) * 512 * scale).astype(np.int64)


def txb_costs(rng, scale=1.0):
    t = np.zeros(TXB_WORDS, np.int64)
    for ctx in range(13):
        t[SKIP + 2 * ctx:SKIP + 2 * ctx + 2] = symbol_costs(rng, 2, scale)
    for ctx in range(4):
        t[BASE_EOB + 3 * ctx:BASE_EOB + 3 * ctx + 3] = symbol_costs(rng, 3, scale)
    for ctx in range(42):
        b = symbol_costs(rng, 4, scale)
        o = BASE + 8 * ctx
        t[o:o + 4] = b
        t[o + 4] = 0
        t[o + 5] = b[1] + 512 - b[0]
        t[o + 6] = b[2] - b[1]
        t[o + 7] = b[3] - b[2]
    for ctx in range(9):
        t[EOB_EXTRA + 2 * ctx:EOB_EXTRA + 2 * ctx + 2] = symbol_costs(rng, 2, scale)
    for ctx in range(3):
        t[DC_SIGN + 2 * ctx:DC_SIGN + 2 * ctx + 2] = symbol_costs(rng, 2, scale)
    for ctx in range(21):
        br = symbol_costs(rng, 4, scale)
        lps = np.zeros(26, np.int64)
        prev, i = 0, 0
        while i < 12:
            for j in range(3):
                lps[i + j] = prev + br[j]
            prev += br[3]
            i += 3
        lps[12] = prev
        lps[13] = lps[0]
        for i in range(1, 13):
            lps[i + 13] = lps[i] - lps[i - 1]
        t[LPS + 26 * ctx:LPS + 26 * ctx + 26] = lps
    return t.astype(np.int32)


def eob_costs(rng, n_sym, scale=1.0):
    e = np.zeros(EOB_WORDS, np.int64)
    for ctx in range(2):
        e[11 * ctx:11 * ctx + n_sym] = symbol_costs(rng, n_sym, scale)
    return e.astype(np.int32)


def coeff_costs(seed, scale=1.0):
    """A frame's tables (zav1_coeff_costs layout): txb[10] then eob[7][2];
    eob_multi_size m codes 5 + m symbols."""
    rng = np.random.default_rng(seed)
    txb = [txb_costs(rng, scale) for _ in range(10)]
    eob = [eob_costs(rng, min(5 + m, 11), scale) for m in range(7) for _ in range(2)]
    return np.concatenate(txb + eob)


def tables_for(costs, txs_ctx, eob_ms, plane):
    t = costs[(txs_ctx * 2 + plane) * TXB_WORDS:(txs_ctx * 2 + plane + 1) * TXB_WORDS]
    o = 10 * TXB_WORDS + (eob_ms * 2 + plane) * EOB_WORDS
    return t, costs[o:o + EOB_WORDS]

