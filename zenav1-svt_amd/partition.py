"""Host driver of the partition pre-pass (zav1_pd0_*): the light PD0 pass of
the allintra presets over every 64x64 unit of a batch of stills, in the
reference's per-preset configuration
(rust/crates/svtav1-encoder/src/pipeline.rs:7939-8420 choosing among
pd0_pick_sb_partition / _lvl0 / _m6_eval, rust/crates/svtav1-encoder/src/
pd0.rs).

`pd0_config(preset, bd, ...)` picks the pass level as the reference does:
presets >= 9 run the effective-M9 pass (LVL_6 / LVL_5 by the per-unit
detector) at 8 bits and the full-RD LVL_0 at 10 bits; presets 2..8 run
LVL_1 (exact coefficient rate at zero contexts; the small-eob approximation
at presets 7 and 8; the 64x64 variance cap from preset 8; 4x4 blocks up to
preset 3).  `PartitionPass` runs it on a device plane and returns the
per-unit records, node records and leaf maps.
"""
import numpy as np

PD0_CFG = np.dtype([('width', '<i4'), ('height', '<i4'), ('n_images', '<i4'), ('bd', '<i4'), ('level', '<i4'),
                    ('qp', '<i4'), ('qindex', '<i4'), ('ires', '<i4'), ('cap_max_block', '<i4'), ('max_tx', '<i4'),
                    ('min_sq', '<i4'), ('coeff_rate_est_lvl', '<i4'), ('tile_w', '<i4'), ('tile_h', '<i4'),
                    ('split_bits', '<i4', 4), ('none_bits', '<i4'), ('skip0_bits', '<i4'), ('txt_dc', '<i4', 3),
                    ('pad', '<i4'), ('lambda', '<u8')])
PD0_SB = np.dtype([('var', '<u2', 85), ('mode', 'u1'), ('max_sq', 'u1'), ('subres', 'u1'), ('pad', 'u1', 3)])
PD0_NODE = np.dtype([('cost', '<u8'), ('flags', 'u1'), ('pad', 'u1', 7)])
VISITED, TESTED, SPLIT = 1, 2, 8
assert PD0_CFG.itemsize == 104 and PD0_SB.itemsize == 176 and PD0_NODE.itemsize == 16


def pd0_level(preset, bd):
    """The PD0 level of an allintra preset: eff-M9 (5) at 8 bits and the
    forced full-RD level 0 at 10 bits from preset 9, level 1 below."""
    if preset >= 9:
        return 0 if bd == 10 else 5
    return 1


def pd0_config(zav1, width, height, bd, qp, qindex, preset, *, n_images=1, tiles=(1, 1), max_tx=64):
    """A zav1_pd0_cfg for a batch of equal-size images at a preset; tiles are
    (columns, rows) of uniform tiles, given in 64x64 units per tile."""
    c = np.zeros(1, PD0_CFG)
    r = c[0]
    r['width'], r['height'], r['n_images'], r['bd'] = width, height, n_images, bd
    r['level'] = pd0_level(preset, bd)
    r['qp'], r['qindex'], r['max_tx'] = qp, qindex, max_tx
    r['min_sq'] = 4 if (preset <= 3 and r['level'] == 1) else 8
    r['cap_max_block'] = 1 if (r['level'] == 1 and preset >= 8) else 0
    r['coeff_rate_est_lvl'] = 2 if preset in (7, 8) else 1
    if r['level'] == 1:
        r['tile_w'], r['tile_h'] = tiles[0] * 64 if tiles[0] else 0, tiles[1] * 64 if tiles[1] else 0
    zav1.check(zav1.lib().zav1_pd0_rates(c.ctypes.data), 'zav1_pd0_rates')
    return c


class PartitionPass:
    """The pass over a stacked device plane (n_images x height rows)."""

    def __init__(self, zav1, cfg):
        self.z = zav1
        self.cfg = np.ascontiguousarray(cfg, PD0_CFG)
        c = self.cfg[0]
        self.n_units = (int(c['width']) // 64) * (int(c['height']) // 64) * int(c['n_images'])
        self.nps = int(zav1.lib().zav1_pd0_nodes_per_unit(self.cfg.ctypes.data))
        self.d_sb = zav1.DeviceArray(self.n_units, PD0_SB)
        self.d_nodes = zav1.DeviceArray(self.n_units * self.nps, PD0_NODE)
        self.d_leaf = zav1.DeviceArray(self.n_units * 256, np.uint8)
        self.d_costs = None
        if int(c['level']) == 1:
            self.d_costs = zav1.DeviceArray.from_host(zav1.build_coeff_costs(int(c['qindex'])))

    def run(self, d_plane, stream=None):
        L = self.z.lib()
        rc = L.zav1_pd0_run(self.cfg.ctypes.data, d_plane.ptr,
                            self.d_costs.ptr if self.d_costs is not None else None,
                            self.d_sb.ptr, self.d_nodes.ptr, self.d_leaf.ptr, stream)
        self.z.check(rc, 'zav1_pd0_run')

    def results(self):
        return (self.d_sb.to_host(), self.d_nodes.to_host().reshape(self.n_units, self.nps),
                self.d_leaf.to_host().reshape(self.n_units, 16, 16))

    def close(self):
        for d in (self.d_sb, self.d_nodes, self.d_leaf, self.d_costs):
            if d is not None:
                d.free()


def leaf_sizes(leaf_map):
    """Leaf sizes of one unit in coding (Z) order from its 16 x 16 map."""
    out = []

    def walk(sq, x, y):
        if (1 << int(leaf_map[y // 4, x // 4])) == sq:
            out.append(sq)
            return
        h = sq // 2
        for i in range(4):
            walk(h, x + (i & 1) * h, y + (i >> 1) * h)

    walk(64, 0, 0)
    return out
