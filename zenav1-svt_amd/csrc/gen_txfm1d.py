"""Generate txfm1d_gen.h: straight-line HIP device code for every AV1 1D
forward and inverse transform used by the encoder (DCT 4..64, ADST 4/8/16,
identity 4..64), specialised per cos_bit so every cosine is an immediate.

The DCT networks are produced by their recursive structure (mirror add/sub,
half-size DCT on the sums, "odd part" rotation ladder on the differences);
this yields exactly the stage-by-stage rounding points of the reference
butterflies (rust/crates/svtav1-dsp/src/fwd_txfm.rs:184-1460,
inv_txfm.rs:79-1486).  The ADST networks follow fwd_txfm.rs:1497-1760 and
inv_txfm.rs:177,1497,1581.

Every value lives in a named SSA temporary, so after inlining the compiler
keeps a whole 64-point transform in VGPRs with no local-memory array.

Run:  python3 gen_txfm1d.py   (writes txfm1d_gen.h beside this file)
"""
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))

# AV1 sinpi rows for cos_bit 10..16 (libaom data; spec "sinpi" constants).
SINPI = [
    [0, 330, 621, 836, 951], [0, 660, 1241, 1672, 1901], [0, 1321, 2482, 3344, 3803],
    [0, 2642, 4964, 6689, 7606], [0, 5283, 9929, 13377, 15212], [0, 10566, 19858, 26755, 30424],
    [0, 21133, 39716, 53510, 60849],
]


def cospi(cb, j):
    return int(math.floor(math.cos(j * math.pi / 128.0) * (1 << cb) + 0.5))


def bitrev(x, bits):
    r = 0
    for _ in range(bits):
        r = (r << 1) | (x & 1)
        x >>= 1
    return r


class Emitter:
    """SSA emitter: every op creates a fresh temporary.

    Alongside the code it runs an interval analysis over magnitudes: every
    temporary gets a bound g * B + c for inputs bounded by B in magnitude,
    and every half butterfly records the linear constraints under which its
    24-bit-operand, 32-bit-sum form (zbtf24) is exact: both operands below
    2^23 and |w0 * a| + |w1 * b| + 2^(cb - 1) below 2^31.  `bmax()` is the
    largest input bound meeting all of them (cancellations are ignored, so
    the bound is conservative)."""

    def __init__(self, cb, narrow=False):
        self.cb = cb
        self.narrow = narrow
        self.lines = []
        self.n = 0
        self.bnd = {}
        self.cons = []  # (alpha, beta, limit): alpha * B + beta <= limit

    def _b(self, name):
        return self.bnd[name[1:]] if name.startswith('-') else self.bnd[name]

    def tmp(self, expr, bound=None):
        name = 't%d' % self.n
        self.n += 1
        self.lines.append('  const int32_t %s = %s;' % (name, expr))
        if bound is None:
            if expr.startswith('x['):
                bound = (1.0, 0.0)
            elif expr.startswith('-'):
                bound = self._b(expr[1:])
            else:
                raise ValueError('no bound for ' + expr)
        self.bnd[name] = bound
        return name

    def c(self, k):
        v = cospi(self.cb, abs(k))
        return -v if k < 0 else v

    def btf(self, w0, a, w1, b):
        (ga, ca), (gb, cb_) = self._b(a), self._b(b)
        r = 1 << (self.cb - 1)
        self.cons.append((ga, ca, (1 << 23) - 1))
        self.cons.append((gb, cb_, (1 << 23) - 1))
        self.cons.append((abs(w0) * ga + abs(w1) * gb, abs(w0) * ca + abs(w1) * cb_ + r, (1 << 31) - 1))
        g = (abs(w0) * ga + abs(w1) * gb) / float(1 << self.cb)
        c = (abs(w0) * ca + abs(w1) * cb_ + r) / float(1 << self.cb) + 1.0
        if self.narrow and abs(w0) == abs(w1):
            # equal weights: w0 * (a +- b) with one 24-bit multiply; the
            # operand a +- b is bounded by the sum of the operand bounds
            self.cons.append((ga + gb, ca + cb_, (1 << 23) - 1))
            op = '+' if w0 == w1 else '-'
            return self.tmp('zbtf24m(%d, %s %s %s, %d)' % (w0, a, op, b, self.cb), (g, c))
        f = 'zbtf24' if self.narrow else 'zbtf'
        return self.tmp('%s(%d, %s, %d, %s, %d)' % (f, w0, a, w1, b, self.cb), (g, c))

    def add(self, a, b, clamp):
        e = '%s + %s' % (a, b)
        (g1, c1), (g2, c2) = self._b(a), self._b(b)
        return self.tmp('zclamp(%s, rng)' % e if clamp else e, (g1 + g2, c1 + c2))

    def sub(self, a, b, clamp):
        e = '%s - %s' % (a, b)
        (g1, c1), (g2, c2) = self._b(a), self._b(b)
        return self.tmp('zclamp(%s, rng)' % e if clamp else e, (g1 + g2, c1 + c2))

    def bmax(self):
        return min(int((lim - beta) // alpha) for alpha, beta, lim in self.cons if alpha > 0)


# ----------------------------------------------------------------------------
# forward DCT
# ----------------------------------------------------------------------------
def gen_fdct(em, x):
    n = len(x)
    out = [None] * n

    def dct(v, oidx):
        N = len(v)
        if N == 2:
            out[oidx[0]] = em.btf(em.c(32), v[0], em.c(32), v[1])
            out[oidx[1]] = em.btf(-em.c(32), v[1], em.c(32), v[0])
            return
        M = N // 2
        a = [em.add(v[i], v[N - 1 - i], False) for i in range(M)]
        b = [None] * M
        for i in range(M):
            b[M - 1 - i] = em.sub(v[i], v[N - 1 - i], False)
        dct(a, [oidx[2 * k] for k in range(M)])
        odd(b, [oidx[2 * k + 1] for k in range(M)])

    def odd(u, oidx):
        M = len(u)
        u = list(u)
        lg = M.bit_length() - 1
        c = em.c
        if M >= 4:
            for j in range(M // 4, M // 2):
                m = M - 1 - j
                nj = em.btf(-c(32), u[j], c(32), u[m])
                nm = em.btf(c(32), u[m], c(32), u[j])
                u[j], u[m] = nj, nm
        S = M // 2
        while S >= 2:
            for k in range(M // S):
                p = k * S
                for i in range(S // 2):
                    lo, hi = p + i, p + S - 1 - i
                    if k % 2 == 0:
                        nl, nh = em.add(u[lo], u[hi], False), em.sub(u[lo], u[hi], False)
                    else:
                        nl, nh = em.sub(u[hi], u[lo], False), em.add(u[hi], u[lo], False)
                    u[lo], u[hi] = nl, nh
            if S >= 4:
                nbh = (M // 2) // S
                bits = nbh.bit_length() - 1
                for k in range(nbh):
                    p = k * S
                    A = (32 * S // M) * (1 + 4 * bitrev(k, bits))
                    B = 64 - A
                    for j in range(p + S // 4, p + S // 2):
                        m = M - 1 - j
                        nj = em.btf(-c(A), u[j], c(B), u[m])
                        nm = em.btf(c(A), u[m], c(B), u[j])
                        u[j], u[m] = nj, nm
                    for j in range(p + S // 2, p + 3 * S // 4):
                        m = M - 1 - j
                        nj = em.btf(-c(B), u[j], -c(A), u[m])
                        nm = em.btf(c(B), u[m], -c(A), u[j])
                        u[j], u[m] = nj, nm
            S //= 2
        for j in range(M // 2):
            m = M - 1 - j
            Q = (32 // M) * (1 + 4 * bitrev(j, lg - 1))
            nj = em.btf(c(64 - Q), u[j], c(Q), u[m])
            nm = em.btf(c(64 - Q), u[m], -c(Q), u[j])
            u[j], u[m] = nj, nm
        for j in range(M):
            out[oidx[bitrev(j, lg)]] = u[j]

    dct(list(x), list(range(n)))
    return out


# ----------------------------------------------------------------------------
# inverse DCT (cos_bit 12, clamped adds)
# ----------------------------------------------------------------------------
def gen_idct(em, coef):
    c = em.c

    def idct(cf):
        N = len(cf)
        if N == 2:
            return [em.btf(c(32), cf[0], c(32), cf[1]), em.btf(c(32), cf[0], -c(32), cf[1])]
        M = N // 2
        E = idct([cf[2 * k] for k in range(M)])
        u = iodd([cf[2 * k + 1] for k in range(M)])
        y = [None] * N
        for i in range(M):
            y[i] = em.add(E[i], u[M - 1 - i], True)
            y[N - 1 - i] = em.sub(E[i], u[M - 1 - i], True)
        return y

    def iodd(cf):
        M = len(cf)
        lg = M.bit_length() - 1
        u = [cf[bitrev(j, lg)] for j in range(M)]
        for j in range(M // 2):
            m = M - 1 - j
            Q = (32 // M) * (1 + 4 * bitrev(j, lg - 1))
            nj = em.btf(c(64 - Q), u[j], -c(Q), u[m])
            nm = em.btf(c(Q), u[j], c(64 - Q), u[m])
            u[j], u[m] = nj, nm
        S = 2
        while S <= M // 2:
            if S >= 4:
                nbh = (M // 2) // S
                bits = nbh.bit_length() - 1
                for k in range(nbh):
                    p = k * S
                    A = (32 * S // M) * (1 + 4 * bitrev(k, bits))
                    B = 64 - A
                    for j in range(p + S // 4, p + S // 2):
                        m = M - 1 - j
                        nj = em.btf(-c(A), u[j], c(B), u[m])
                        nm = em.btf(c(B), u[j], c(A), u[m])
                        u[j], u[m] = nj, nm
                    for j in range(p + S // 2, p + 3 * S // 4):
                        m = M - 1 - j
                        nj = em.btf(-c(B), u[j], -c(A), u[m])
                        nm = em.btf(-c(A), u[j], c(B), u[m])
                        u[j], u[m] = nj, nm
            for k in range(M // S):
                p = k * S
                for i in range(S // 2):
                    lo, hi = p + i, p + S - 1 - i
                    if k % 2 == 0:
                        nl, nh = em.add(u[lo], u[hi], True), em.sub(u[lo], u[hi], True)
                    else:
                        nl, nh = em.sub(u[hi], u[lo], True), em.add(u[hi], u[lo], True)
                    u[lo], u[hi] = nl, nh
            S *= 2
        if M >= 4:
            for j in range(M // 4, M // 2):
                m = M - 1 - j
                nj = em.btf(-c(32), u[j], c(32), u[m])
                nm = em.btf(c(32), u[j], c(32), u[m])
                u[j], u[m] = nj, nm
        return u

    return idct(list(coef))


# ----------------------------------------------------------------------------
# ADST 8/16 (forward and inverse) — shared stage vocabulary
# ----------------------------------------------------------------------------
def _rot_f1(em, v, p, X):
    a, b = v[p], v[p + 1]
    v[p] = em.btf(em.c(X), a, em.c(64 - X), b)
    v[p + 1] = em.btf(em.c(64 - X), a, -em.c(X), b)


def _rot_f2(em, v, p, X):
    a, b = v[p], v[p + 1]
    v[p] = em.btf(-em.c(64 - X), a, em.c(X), b)
    v[p + 1] = em.btf(em.c(X), a, em.c(64 - X), b)


def _rot32(em, v, p):
    a, b = v[p], v[p + 1]
    v[p] = em.btf(em.c(32), a, em.c(32), b)
    v[p + 1] = em.btf(em.c(32), a, -em.c(32), b)


def _addsub(em, v, d, clamp):
    for b0 in range(0, len(v), 2 * d):
        for i in range(d):
            x, y = v[b0 + i], v[b0 + i + d]
            v[b0 + i] = em.add(x, y, clamp)
            v[b0 + i + d] = em.sub(x, y, clamp)


def gen_fadst8(em, x):
    v = [x[0], '-' + x[7], '-' + x[3], x[4], '-' + x[1], x[6], x[2], '-' + x[5]]
    v = [em.tmp(e) if e.startswith('-') else e for e in v]
    _rot32(em, v, 2); _rot32(em, v, 6)
    _addsub(em, v, 2, False)
    _rot_f1(em, v, 4, 16); _rot_f2(em, v, 6, 16)
    _addsub(em, v, 4, False)
    for k in range(4):
        _rot_f1(em, v, 2 * k, 4 + 16 * k)
    return [v[i] for i in (1, 6, 3, 4, 5, 2, 7, 0)]


def gen_fadst16(em, x):
    order = [(0, 1), (15, -1), (7, -1), (8, 1), (3, -1), (12, 1), (4, 1), (11, -1),
             (1, -1), (14, 1), (6, 1), (9, -1), (2, 1), (13, -1), (5, -1), (10, 1)]
    v = [x[i] if s > 0 else em.tmp('-' + x[i]) for i, s in order]
    for p in range(2, 16, 4):
        _rot32(em, v, p)
    _addsub(em, v, 2, False)
    for b0 in (0, 8):
        _rot_f1(em, v, b0 + 4, 16); _rot_f2(em, v, b0 + 6, 16)
    _addsub(em, v, 4, False)
    _rot_f1(em, v, 8, 8); _rot_f1(em, v, 10, 40); _rot_f2(em, v, 12, 8); _rot_f2(em, v, 14, 40)
    _addsub(em, v, 8, False)
    for k in range(8):
        _rot_f1(em, v, 2 * k, 2 + 8 * k)
    return [v[i] for i in (1, 14, 3, 12, 5, 10, 7, 8, 9, 6, 11, 4, 13, 2, 15, 0)]


def gen_iadst8(em, x):
    v = [x[i] for i in (7, 0, 5, 2, 3, 4, 1, 6)]
    for k in range(4):
        _rot_f1(em, v, 2 * k, 4 + 16 * k)
    _addsub(em, v, 4, True)
    _rot_f1(em, v, 4, 16); _rot_f2(em, v, 6, 16)
    _addsub(em, v, 2, True)
    _rot32(em, v, 2); _rot32(em, v, 6)
    src = (0, 4, 6, 2, 3, 7, 5, 1)
    return [v[s] if i % 2 == 0 else em.tmp('-' + v[s]) for i, s in enumerate(src)]


def gen_iadst16(em, x):
    v = [x[i] for i in (15, 0, 13, 2, 11, 4, 9, 6, 7, 8, 5, 10, 3, 12, 1, 14)]
    for k in range(8):
        _rot_f1(em, v, 2 * k, 2 + 8 * k)
    _addsub(em, v, 8, True)
    _rot_f1(em, v, 8, 8); _rot_f1(em, v, 10, 40); _rot_f2(em, v, 12, 8); _rot_f2(em, v, 14, 40)
    _addsub(em, v, 4, True)
    for b0 in (0, 8):
        _rot_f1(em, v, b0 + 4, 16); _rot_f2(em, v, b0 + 6, 16)
    _addsub(em, v, 2, True)
    for p in range(2, 16, 4):
        _rot32(em, v, p)
    src = (0, 8, 12, 4, 6, 14, 10, 2, 3, 11, 15, 7, 5, 13, 9, 1)
    return [v[s] if i % 2 == 0 else em.tmp('-' + v[s]) for i, s in enumerate(src)]


def emit_fn(name, n, body_fn, cb, needs_rng=False, narrow=False):
    """Returns (code, bmax, (g, c)): the input bound under which the narrow
    form is exact and the output bound g * B + c."""
    em = Emitter(cb, narrow)
    ins = ['x[%d]' % i for i in range(n)]
    em.lines.append('  // load')
    loaded = [em.tmp(v) for v in ins]
    outs = body_fn(em, loaded)
    sig = 'static __device__ __forceinline__ void %s(int32_t* __restrict__ x%s)' % (
        name, ', int rng' if needs_rng else '')
    store = ['  x[%d] = %s;' % (i, o) for i, o in enumerate(outs)]
    g = max(em._b(o)[0] for o in outs)
    c = max(em._b(o)[1] for o in outs)
    return '\n'.join([sig + ' {'] + em.lines + store + ['}', '']), em.bmax(), (g, c)


def main():
    parts = [
        '// GENERATED by gen_txfm1d.py — do not edit.  AV1 1D transforms as',
        '// straight-line SSA device code; see the generator for provenance.',
        '#pragma once',
        '#include <stdint.h>',
        '',
        '// i64 half butterfly, rounded right shift by cos_bit.',
        'static __device__ __forceinline__ int32_t zbtf(int32_t w0, int32_t a, int32_t w1, int32_t b, int cb) {',
        '  const int64_t r = (int64_t)w0 * a + (int64_t)w1 * b;',
        '  return (int32_t)((r + ((int64_t)1 << (cb - 1))) >> cb);',
        '}',
        '// The same butterfly with 24-bit-operand multiplies and a 32-bit sum:',
        '// exact when |a|, |b| < 2^23 and |w0 a| + |w1 b| + 2^(cb-1) < 2^31, which',
        '// the generator proves for the input bound of each *_n variant.',
        'static __device__ __forceinline__ int32_t zbtf24(int32_t w0, int32_t a, int32_t w1, int32_t b, int cb) {',
        '  return (__mul24(w0, a) + __mul24(w1, b) + (1 << (cb - 1))) >> cb;',
        '}',
        '// Equal-magnitude weights: w * (a +- b), the operand a +- b bounded like a and b.',
        'static __device__ __forceinline__ int32_t zbtf24m(int32_t w, int32_t ab, int cb) {',
        '  return (__mul24(w, ab) + (1 << (cb - 1))) >> cb;',
        '}',
        '// signed clamp to `bit` bits (no-op when bit <= 0).',
        'static __device__ __forceinline__ int32_t zclamp(int32_t v, int bit) {',
        '  if (bit <= 0) return v;',
        '  const int32_t mx = (int32_t)((1u << (bit - 1)) - 1u), mn = -mx - 1;',
        '  return v < mn ? mn : (v > mx ? mx : v);',
        '}',
        '',
    ]
    bounds = []  # (kind, n, cb, bmax, g, c)
    for cb in (10, 11, 12, 13):
        for kind, fn, sizes in (('fdct', gen_fdct, (4, 8, 16, 32, 64)), ('fadst', gen_fadst8, (8,)),
                                ('fadst', gen_fadst16, (16,))):
            for n in sizes:
                code, bmax, (g, c) = emit_fn('%s%d_cb%d' % (kind, n, cb), n, fn, cb)
                parts.append(code)
                parts.append(emit_fn('%s%d_cb%d_n' % (kind, n, cb), n, fn, cb, narrow=True)[0])
                bounds.append((kind, n, cb, bmax, g, c))
    for n in (4, 8, 16, 32, 64):
        parts.append(emit_fn('idct%d' % n, n, gen_idct, 12, needs_rng=True)[0])
    parts.append(emit_fn('iadst8', 8, gen_iadst8, 12, needs_rng=True)[0])
    parts.append(emit_fn('iadst16', 16, gen_iadst16, 12, needs_rng=True)[0])
    # input bound of each narrow forward variant and the output bound g * B + c
    parts.append('// Interval analysis of the forward transforms: the *_n form is exact for')
    parts.append('// inputs |x| <= bmax; outputs satisfy |y| <= g * B + c for inputs |x| <= B.')
    parts.append('struct ZTxBound {')
    parts.append('  long long bmax;')
    parts.append('  double g, c;')
    parts.append('};')
    parts.append('// kind 0 DCT, 1 ADST (8 and 16 points)')
    parts.append('constexpr ZTxBound ztx_bound(int kind, int n, int cb) {')
    for kind, n, cb, bmax, g, c in bounds:
        k = 0 if kind == 'fdct' else 1
        parts.append('  if (kind == %d && n == %d && cb == %d) return ZTxBound{%dLL, %.17g, %.17g};' % (
            k, n, cb, bmax, g, c))
    parts.append('  return ZTxBound{-1LL, 0.0, 0.0};')
    parts.append('}')
    parts.append('')
    # sinpi rows for the 4-point ADST, indexed by cos_bit - 10
    rows = ',\n  '.join('{%s}' % ', '.join(str(v) for v in r) for r in SINPI)
    parts.append('static __device__ __constant__ const int32_t kSinpi[7][5] = {\n  %s};\n' % rows)
    with open(os.path.join(HERE, 'txfm1d_gen.h'), 'w') as f:
        f.write('\n'.join(parts))
    # Algorithmic integer-op counts per 1D transform (the roofline numerator):
    # add/sub/neg = 1, clamp = 2, half butterfly = 2 mul + add + round add +
    # shift = 5.  Identity: 1 (x2, x4) or 3 (x sqrt2 scalings) per sample;
    # 4-point ADST: 7 mul + 10 add + 4 round-shift pairs.
    ops = {}
    for name, n, fn in (('dct', 4, gen_fdct), ('dct', 8, gen_fdct), ('dct', 16, gen_fdct),
                        ('dct', 32, gen_fdct), ('dct', 64, gen_fdct), ('adst', 8, gen_fadst8),
                        ('adst', 16, gen_fadst16)):
        em = Emitter(12)
        fn(em, [em.tmp('x[%d]' % i) for i in range(n)])
        cnt = 0
        for line in em.lines[n:]:
            cnt += 5 if 'zbtf' in line else (2 if 'zclamp' in line else 1)
        ops['%s%d' % (name, n)] = cnt
    ops['adst4'] = 7 + 10 + 8
    for n in (4, 8, 16, 32, 64):
        ops['identity%d' % n] = n * (3 if n in (4, 16, 64) else 1)
    import json
    with open(os.path.join(HERE, 'txfm1d_ops.json'), 'w') as f:
        json.dump(ops, f, indent=1, sort_keys=True)
    print('wrote txfm1d_gen.h')


if __name__ == '__main__':
    main()
