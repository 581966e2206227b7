"""Host mirror of the AV1 still writer (zav1_write_still): the leaf record
and descriptor layouts of include/zav1_dsp.h and helpers that build them.

The writer replaces the reference's entropy walk and OBU packing for the
luma plane (rust/crates/svtav1-encoder/src/pipeline.rs:2542-2811 over
rust/crates/svtav1-entropy/src/range_coder.rs and coeff_c.rs:924); it runs
on host threads, serial per tile as the range coder is."""
import ctypes

import numpy as np

STILL_LEAF = np.dtype([('x', '<i4'), ('y', '<i4'), ('size', 'u1'), ('mode', 'u1'), ('delta', 'i1'), ('fi_mode', 'u1'),
                       ('tx_type', 'u1'), ('pad', 'u1', 3), ('eob', '<i4'), ('pad2', '<i4'), ('level_off', '<u8')])
assert STILL_LEAF.itemsize == 32


class StillDesc(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ('width', 'height', 'bd', 'sb', 'tile_cols', 'tile_rows', 'qindex',
                                              'n_leaves', 'threads', 'pad')] + [('leaves', ctypes.c_void_p),
                                                                                ('levels', ctypes.c_void_p)]


def write_still(zav1, W, H, bd, sb, tiles, qindex, leaves, levels, threads=0):
    """(status, stream bytes or None): the still of `leaves` (STILL_LEAF,
    any order) with their packed levels (int32, each leaf's min(size, 32)^2
    at its level_off)."""
    leaves = np.ascontiguousarray(leaves, STILL_LEAF)
    levels = np.ascontiguousarray(levels, np.int32)
    d = StillDesc()
    d.width, d.height, d.bd, d.sb = W, H, bd, sb
    d.tile_cols, d.tile_rows, d.qindex, d.n_leaves, d.threads = tiles[0], tiles[1], qindex, len(leaves), threads
    d.leaves, d.levels = leaves.ctypes.data, levels.ctypes.data
    size = ctypes.c_size_t()
    L = zav1.lib()
    out = np.empty(max(1 << 16, W * H // 2), np.uint8)
    rc = L.zav1_write_still(ctypes.byref(d), out.ctypes.data, out.size, ctypes.byref(size))
    if rc == zav1.ZAV1_ENOMEM:
        out = np.empty(size.value, np.uint8)
        rc = L.zav1_write_still(ctypes.byref(d), out.ctypes.data, out.size, ctypes.byref(size))
    if rc:
        return rc, None
    return rc, out[:size.value].tobytes()


def leaves_uniform(dec, lev, leaf, cands):
    """Still leaves of one image decided on uniform leaves: dec the per-leaf
    decision records ((H / leaf) x (W / leaf), fields cand / mode / fi_mode /
    tx_type / eob), lev the image's level plane (H x W int32, each leaf's
    min(leaf, 32)^2 levels in raster at its top-left corner), cands the
    configuration's candidate records (for the angle delta)."""
    gh, gw = dec.shape
    P = min(leaf, 32)
    n = gh * gw
    leaves = np.zeros(n, STILL_LEAF)
    ys, xs = np.meshgrid(np.arange(gh) * leaf, np.arange(gw) * leaf, indexing='ij')
    d = dec.reshape(-1)
    leaves['x'], leaves['y'], leaves['size'] = xs.reshape(-1), ys.reshape(-1), leaf
    leaves['mode'], leaves['fi_mode'], leaves['tx_type'] = d['mode'], d['fi_mode'], d['tx_type']
    leaves['delta'] = np.asarray(cands['delta'])[d['cand']]
    leaves['eob'] = d['eob']
    leaves['level_off'] = np.arange(n, dtype=np.uint64) * (P * P)
    blocks = lev.reshape(gh, leaf, gw, leaf)[:, :P, :, :P].transpose(0, 2, 1, 3).reshape(n, P * P)
    return leaves, np.ascontiguousarray(blocks, np.int32)


def legal_tile_cols(width, height, sb):
    """The fewest uniform tile columns (a power of two) that keep every tile
    within AV1's MAX_TILE_AREA (4096 x 2304 samples): 1 up to that area, 2
    for a 4096 x 4096 frame at SB64."""
    sb_cols = -(-width // sb)
    tc = 1
    while tc * 2 <= sb_cols and -(-sb_cols // tc) * sb * height > 4096 * 2304:
        tc *= 2
    return tc
