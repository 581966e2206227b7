"""GPU parity of the Wiener unit filter (zav1_wiener_apply), the finer tap
search (zav1_wiener_refine) and the unit SSE's 10-bit decomposition
(zav1_unit_sse) against the oracle's restatements of
rust/crates/svtav1-dsp/src/restoration.rs:1288 / :1921 / :1991 and
rust/crates/svtav1-encoder/src/restoration.rs:513: every pixel, tap and
SSE bit-exact, at 8 and 10 bits, unit sizes 32..256, luma and 4:2:0 chroma
geometry, and on the 4096x4096 8-bit loop-filter frame."""
import importlib.util
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_spec = importlib.util.spec_from_file_location(
    'zav1_synth', os.path.join(os.path.dirname(os.path.abspath(__file__)), '..', 'zenav1-svt_amd', 'synth.py'))
synth = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(synth)

TAP_MIN = (3 - 8, -7 - 16, 15 - 32)
TAP_MAX = (3 - 1 + 8, -7 - 1 + 16, 15 - 1 + 32)


def random_filter(rng, win5):
    t0 = 0 if win5 else int(rng.integers(TAP_MIN[0], TAP_MAX[0] + 1))
    t1 = int(rng.integers(TAP_MIN[1], TAP_MAX[1] + 1))
    t2 = int(rng.integers(TAP_MIN[2], TAP_MAX[2] + 1))
    return np.array([t0, t1, t2, -2 * (t0 + t1 + t2), t2, t1, t0, 0], np.int16)


def units_for(z, rng, n, win5, p_wiener=0.85):
    u = np.zeros(n, z.WIENER_UNIT)
    for i in range(n):
        u[i]['hfilter'] = random_filter(rng, win5)
        u[i]['vfilter'] = random_filter(rng, win5)
        u[i]['type'] = int(rng.random() < p_wiener)
    return u


def noisy(rng, h, w, bd):
    hi = (1 << bd) - 1
    p = np.clip(rng.normal(hi / 2, hi / 5, (h, w)), 0, hi)
    for _ in range(10):
        r, c = rng.integers(0, h - 4), rng.integers(0, w - 4)
        p[r:r + rng.integers(2, 40), c:c + rng.integers(2, 40)] = rng.integers(0, hi + 1)
    return p.astype(np.uint16 if bd > 8 else np.uint8)


def apply_case(z, oracle, d, bd, unit, ss_y, u):
    D = z.DeviceArray
    h, w = d.shape
    dd, do = D.from_host(d), D(d.shape, d.dtype)
    z.wiener_apply(dd, do, w, h, bd, unit, ss_y, D.from_host(u))
    filt = np.stack([u['hfilter'], u['vfilter']], axis=1)
    want = oracle.wiener_filter_plane(d, bd, unit, ss_y, u['type'], filt)
    got = do.to_host()
    assert np.array_equal(got, want), np.argwhere(got != want)[:6]


@pytest.mark.parametrize('bd,h,w,unit,ss_y,win5,seed', [(8, 200, 232, 64, 0, False, 1), (10, 150, 136, 64, 1, True, 2),
                                                         (8, 130, 300, 128, 0, False, 3), (10, 96, 96, 32, 1, True, 4),
                                                         (10, 520, 260, 256, 0, False, 5), (8, 72, 40, 32, 0, False, 6)])
def test_wiener_apply_bitexact(gpu, oracle, bd, h, w, unit, ss_y, win5, seed):
    rng = np.random.default_rng(seed)
    d = noisy(rng, h, w, bd)
    nh, nv = gpu.rest_unit_grid(unit, w, h)
    u = units_for(gpu, rng, nh * nv, win5)
    u[0]['hfilter'] = [7, -23, -17, 66, -17, -23, 7, 0]
    u[0]['vfilter'] = [-5, 8, 46, -98, 46, 8, -5, 0]
    u[0]['type'] = 1
    apply_case(gpu, oracle, d, bd, unit, ss_y, u)


def test_wiener_apply_loopfilter_frame(gpu, oracle):
    """The 4096x4096 8-bit reconstruction (units of 128) and one 2048x2048
    chroma plane (units of 64, ss_y 1)."""
    rng = np.random.default_rng(3100)
    _, y = synth.config3_recon(4096, 4096, seed=3000)
    nh, nv = gpu.rest_unit_grid(128, 4096, 4096)
    apply_case(gpu, oracle, y, 8, 128, 0, units_for(gpu, rng, nh * nv, False))
    (_, u), _ = synth.config3_chroma(2048, 2048, seed=3000)
    nh, nv = gpu.rest_unit_grid(64, 2048, 2048)
    apply_case(gpu, oracle, u, 8, 64, 1, units_for(gpu, rng, nh * nv, True))


@pytest.mark.parametrize('bd,h,w,unit,ss_y,win,max_one,seed', [(8, 192, 256, 64, 0, 7, False, 11),
                                                                (10, 200, 136, 64, 1, 5, False, 12),
                                                                (10, 256, 256, 128, 0, 7, True, 13),
                                                                (8, 136, 72, 32, 1, 5, False, 14)])
def test_wiener_refine_bitexact(gpu, oracle, bd, h, w, unit, ss_y, win, max_one, seed):
    z = gpu
    rng = np.random.default_rng(seed)
    src = noisy(rng, h, w, bd)
    hi = (1 << bd) - 1
    dgd = np.clip(src.astype(np.int64) + rng.integers(-hi // 16, hi // 16 + 1, src.shape), 0, hi).astype(src.dtype)
    nh, nv = z.rest_unit_grid(unit, w, h)
    n = nh * nv
    u = units_for(z, rng, n, win == 5, p_wiener=1.0)
    D = z.DeviceArray
    du, de = D.from_host(u), D((n,), np.int64)
    z.wiener_refine(D.from_host(dgd), D.from_host(src), w, h, bd, win, unit, ss_y, du, de, max_one_step=max_one)
    gu, ge = du.to_host(), de.to_host()
    moved = 0
    for i in range(n):
        lim = oracle.unit_limits(unit, w, h, ss_y, i % nh, i // nh)
        e, hf, vf = oracle.wiener_refine_unit(dgd, src, bd, win, lim, u[i]['hfilter'], u[i]['vfilter'],
                                              max_one_step=max_one)
        assert int(ge[i]) == e, (i, int(ge[i]), e)
        assert list(gu[i]['hfilter']) == list(hf) and list(gu[i]['vfilter']) == list(vf), i
        moved += int(list(hf) != list(u[i]['hfilter']) or list(vf) != list(u[i]['vfilter']))
    assert moved > 0


def test_unit_sse_hbd_truncates_the_right_strip(gpu, oracle):
    """A 31 x 600 10-bit plane in units of 256: the second unit's right strip
    (15 columns x 352 rows at error 1023) sums past 2^32 and is truncated."""
    z = gpu
    h, w = 600, 31
    a, b = np.zeros((h, w), np.uint16), np.full((h, w), 1023, np.uint16)
    D = z.DeviceArray
    out = D((2,), np.uint64)
    z.unit_sse(D.from_host(a), D.from_host(b), w, h, 10, 256, 0, out)
    got = out.to_host()
    L = oracle.lrlib()
    for i in range(2):
        lim = oracle.unit_limits(256, w, h, 0, 0, i)
        assert int(got[i]) == L.or_sse_region(a.ctypes.data, b.ctypes.data, w, 10, *lim)
    assert int(got[1]) != (lim[3] - lim[2]) * w * 1023 ** 2
