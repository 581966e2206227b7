"""GPU parity of the CDEF frame pass (zav1_cdef_apply) against the oracle's
restatement of apply_cdef_frame / apply_cdef_frame_hbd
(rust/crates/svtav1-encoder/src/cdef.rs:343 / :556): every output pixel and
the evidence counters bit-exact, at 8 and 10 bits, partial filter blocks,
with and without chroma and skip, and on the 4096x4096 8-bit 4:2:0
loop-filter reconstruction (1/f noise with 8x8 / 4x4 mean-flattened
blocks)."""
import importlib.util
import os

import numpy as np
import pytest

_spec = importlib.util.spec_from_file_location(
    'zav1_synth', os.path.join(os.path.dirname(os.path.abspath(__file__)), '..', 'zenav1-svt_amd', 'synth.py'))
synth = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(synth)

pytestmark = pytest.mark.gpu


def run(z, oracle, y, u, v, bd, damping, fs, skip):
    D = z.DeviceArray
    h, w = y.shape
    din = [D.from_host(y), D.from_host(u) if u is not None else None, D.from_host(v) if v is not None else None]
    dout = [D(y.shape, y.dtype), D(u.shape, u.dtype) if u is not None else None,
            D(v.shape, v.dtype) if v is not None else None]
    dst = D((2,), np.uint64)
    dsk = D.from_host(skip) if skip is not None else None
    mr, mc = skip.shape if skip is not None else (0, 0)
    z.cdef_apply(din, dout, w, h, bd, damping, D.from_host(np.ascontiguousarray(fs, np.uint8)), dsk, mc, mr, dst)
    oy, ou, ov, st = oracle.cdef_apply_frame(y, u, v, bd, damping, fs, skip)
    gy = dout[0].to_host()
    assert np.array_equal(gy, oy), np.argwhere(gy != oy)[:6]
    if u is not None:
        assert np.array_equal(dout[1].to_host(), ou) and np.array_equal(dout[2].to_host(), ov)
    assert dst.to_host().tolist() == st.tolist()
    return st


@pytest.mark.parametrize('bd,h,w,chroma,skipf,seed', [(8, 136, 200, True, 0.3, 1), (10, 136, 200, True, 0.0, 2),
                                                       (10, 96, 320, False, 0.5, 3), (8, 64, 64, True, 0.9, 4),
                                                       (10, 200, 136, True, 0.2, 5)])
def test_cdef_apply_bitexact(gpu, oracle, bd, h, w, chroma, skipf, seed):
    rng = np.random.default_rng(seed)
    hi = (1 << bd) - 1
    dt = np.uint16 if bd > 8 else np.uint8
    y = np.clip(rng.normal(hi / 2, hi / 5, (h, w)), 0, hi).astype(dt)
    for _ in range(8):
        r, c = rng.integers(0, h - 8), rng.integers(0, w - 8)
        y[r:r + rng.integers(4, 30), c:c + rng.integers(4, 30)] = rng.integers(0, hi + 1)
    u = np.clip(rng.normal(hi / 2, hi / 7, (h // 2, w // 2)), 0, hi).astype(dt) if chroma else None
    v = np.clip(rng.normal(hi / 2, hi / 7, (h // 2, w // 2)), 0, hi).astype(dt) if chroma else None
    nfb = ((w + 63) // 64) * ((h + 63) // 64)
    fs = rng.integers(0, 64, (nfb, 2)).astype(np.uint8)
    fs[0] = (0, 3)
    skip = (rng.random(((h + 3) // 4, (w + 3) // 4)) < skipf).astype(np.uint8) if skipf else None
    st = run(gpu, oracle, y, u, v, bd, int(rng.integers(3, 7)), fs, skip)
    assert st[0] > 0


def test_cdef_apply_loopfilter_frame(gpu, oracle):
    """The 4096x4096 8-bit 4:2:0 loop-filter reconstruction, every filter
    block at its own strength pair, 20% of the mi signaled skip."""
    _, y = synth.config3_recon(4096, 4096, seed=3000)
    (_, u), (_, v) = synth.config3_chroma(2048, 2048, seed=3000)
    rng = np.random.default_rng(3001)
    fs = rng.integers(0, 64, (64 * 64, 2)).astype(np.uint8)
    skip = synth.partition_geometry(1024, 1024, 3002, skip_frac=0.2)['skip']
    st = run(gpu, oracle, y, u, v, 8, 5, fs, skip)
    assert st[1] > 0
