"""GPU parity of the fused tx-type search (DC prediction -> residual -> fwd ->
fold -> fp quantize -> distortion / eob / level sum) against the oracle."""
import importlib.util
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _mod(name):
    path = os.path.join(ROOT, 'zenav1-svt_amd', name + '.py')
    spec = importlib.util.spec_from_file_location('zav1_' + name, path)
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


@pytest.mark.parametrize('bd,qp', [(10, 20), (10, 40), (8, 32)])
def test_txsearch_all_sizes_bitexact(gpu, oracle, bd, qp):
    zav1 = gpu
    synth, ts = _mod('synth'), _mod('txsearch')
    plane = synth.config2_luma(256, 256, seed=2000 + qp)
    if bd == 8:
        plane = (plane >> 2).astype(np.uint16)
    qi = synth.qp_to_qindex(qp)
    sweep = ts.TxSearchSweep(zav1, plane, bd, [qi])
    sweep.run()
    zav1.sync()
    for tx_size in range(19):
        xy, types, dist, eob, lsum = sweep.results(tx_size)
        rd, re, rl = oracle.txsearch_plane(plane, bd, tx_size, xy, types, qi)
        assert np.array_equal(dist, rd), (tx_size, np.argwhere(dist != rd)[:3])
        assert np.array_equal(eob, re), tx_size
        assert np.array_equal(lsum, rl), tx_size


def test_txsearch_ragged_and_edges(gpu, oracle):
    """Blocks on the frame's first row/column (no neighbours), a ragged pair
    count, extreme samples."""
    zav1 = gpu
    rng = np.random.default_rng(9)
    plane = rng.integers(0, 1024, (96, 160)).astype(np.uint16)
    plane[:8, :8] = 1023
    plane[8:16, :8] = 0
    xy = np.array([[0, 0], [0, 8], [8, 0], [24, 40], [152, 88], [100, 60], [4, 4]], np.int32)
    for tx_size in (0, 1, 5, 6, 13):
        w, h = zav1.TX_DIMS[tx_size]
        ok = (xy[:, 0] + w <= 160) & (xy[:, 1] + h <= 96)
        sub = xy[ok]
        types = np.array([0, 3, 5, 9, 15], np.uint8)
        d_plane = zav1.DeviceArray.from_host(plane)
        d_xy = zav1.DeviceArray.from_host(sub)
        n = sub.shape[0] * types.size
        dd, de, dl = zav1.DeviceArray(n, np.uint64), zav1.DeviceArray(n, np.int32), zav1.DeviceArray(n, np.uint32)
        zav1.txsearch_batch(d_plane, 160, 10, tx_size, d_xy, sub.shape[0], types, 255, dd, de, dl)
        rd, re, rl = oracle.txsearch_plane(plane, 10, tx_size, sub, types, 255)
        assert np.array_equal(dd.to_host(), rd) and np.array_equal(de.to_host(), re)
        assert np.array_equal(dl.to_host(), rl)
