"""CPU check of the forward transforms' narrow (24-bit operand, 32-bit sum)
butterflies against the generator's interval analysis.

txfm1d_gen.h carries, for every forward DCT / ADST and cos_bit, a wide form
(64-bit butterflies, as the reference computes) and a narrow form, with the
largest input bound `bmax` under which the narrow form is claimed exact.
This test interprets both generated SSA bodies in Python integers on inputs
at that bound (random values, random sign patterns, constant extremes) and
checks every narrow butterfly's operands (< 2^23) and sum (< 2^31) and that
the outputs equal the wide form's.  A wrong bound in the analysis (a
dropped term, a gain off by a factor) shows up as a violated limit."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEN = os.path.join(ROOT, 'zenav1-svt_amd', 'csrc', 'txfm1d_gen.h')


def _functions():
    src = open(GEN).read()
    fns = {}
    for m in re.finditer(r'static __device__ __forceinline__ void (\w+)\(int32_t\* __restrict__ x(?:, int rng)?\) \{\n(.*?)\n\}',
                         src, re.S):
        fns[m.group(1)] = m.group(2).splitlines()
    bounds = {}
    for m in re.finditer(r'if \(kind == (\d) && n == (\d+) && cb == (\d+)\) return ZTxBound\{(-?\d+)LL', src):
        bounds[(int(m.group(1)), int(m.group(2)), int(m.group(3)))] = int(m.group(4))
    return fns, bounds


def _run(lines, x, check=False):
    env = {}
    for ln in lines:
        ln = ln.strip()
        m = re.match(r'const int32_t (t\d+) = (.*);$', ln)
        if m:
            name, e = m.groups()
            mm = re.match(r'x\[(\d+)\]$', e)
            if mm:
                v = x[int(mm.group(1))]
            elif re.match(r'-(t\d+)$', e):
                v = -env[e[1:]]
            elif re.match(r'(t\d+) ([+-]) (t\d+)$', e):
                a, op, b = re.match(r'(t\d+) ([+-]) (t\d+)$', e).groups()
                v = env[a] + env[b] if op == '+' else env[a] - env[b]
            elif e.startswith('zbtf24m('):
                mm = re.match(r'zbtf24m\((-?\d+), (t\d+) ([+-]) (t\d+), (\d+)\)$', e)
                assert mm, e
                w, a, op, b, cb = mm.groups()
                w, cb = int(w), int(cb)
                ab = env[a] + env[b] if op == '+' else env[a] - env[b]
                s = w * ab + (1 << (cb - 1))
                if check:
                    assert abs(ab) < (1 << 23), (ln, ab)
                    assert -(1 << 31) <= s < (1 << 31), (ln, s)
                v = s >> cb
            else:
                mm = re.match(r'(zbtf24|zbtf)\((-?\d+), (t\d+), (-?\d+), (t\d+), (\d+)\)$', e)
                assert mm, e
                f, w0, a, w1, b, cb = mm.groups()
                w0, w1, cb = int(w0), int(w1), int(cb)
                av, bv = env[a], env[b]
                s = w0 * av + w1 * bv + (1 << (cb - 1))
                if check and f == 'zbtf24':
                    assert abs(av) < (1 << 23) and abs(bv) < (1 << 23), (ln, av, bv)
                    assert -(1 << 31) <= s < (1 << 31), (ln, s)
                v = s >> cb
            assert -(1 << 31) <= v < (1 << 31), (ln, v)
            env[name] = v
            continue
        m = re.match(r'x\[(\d+)\] = (t\d+);$', ln)
        if m:
            x[int(m.group(1))] = env[m.group(2)]
    return x


@pytest.mark.parametrize('cb', [10, 11, 12, 13])
def test_narrow_forward_transforms_exact_at_bound(cb):
    fns, bounds = _functions()
    rng = np.random.default_rng(cb)
    for kind, name, sizes in ((0, 'fdct', (4, 8, 16, 32, 64)), (1, 'fadst', (8, 16))):
        for n in sizes:
            wide, narrow = fns['%s%d_cb%d' % (name, n, cb)], fns['%s%d_cb%d_n' % (name, n, cb)]
            bmax = bounds[(kind, n, cb)]
            assert bmax > 0
            trials = [rng.integers(-bmax, bmax + 1, n) for _ in range(40)]
            trials += [rng.choice([-bmax, bmax], n) for _ in range(40)]
            trials += [np.full(n, bmax), np.full(n, -bmax), np.array([bmax, -bmax] * (n // 2))]
            for x in trials:
                xs = [int(v) for v in x]
                assert _run(narrow, list(xs), check=True) == _run(wide, list(xs)), (name, n, cb)


def test_bounds_are_tight_enough_to_matter():
    """The analysis admits the fused kernels' column inputs (10-bit residuals
    << 2) for every DCT of up to 32 points and every ADST."""
    _, bounds = _functions()
    for cb in (12, 13):
        for n in (4, 8, 16, 32):
            assert bounds[(0, n, cb)] >= 1023 << 2, (n, cb)
        for n in (8, 16):
            assert bounds[(1, n, cb)] >= 1023 << 2, (n, cb)
