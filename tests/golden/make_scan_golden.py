"""Digest the AV1 scan orders that the reference holds as C-extracted data
(rust/crates/svtav1-entropy/src/scan_tables.rs, eb_av1_scan_orders) into
tests/golden/scan_golden.json: per (tx_size, scan class) the length, the first
16 entries and the sha256 of the whole table as little-endian u16."""
import hashlib
import json
import os
import re

import numpy as np

SRC = '/root/reference/rust/crates/svtav1-entropy/src/scan_tables.rs'
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    s = open(SRC).read()
    out = []
    for m in re.finditer(r'static SCAN_TS(\d+)_C(\d+): \[u16; (\d+)\] = \[(.*?)\];', s, re.S):
        v = np.array([int(x) for x in re.findall(r'\d+', m.group(4))], dtype='<u2')
        out.append({'tx_size': int(m.group(1)), 'scan_class': int(m.group(2)), 'len': int(v.size),
                    'head': v[:16].tolist(), 'sha256': hashlib.sha256(v.tobytes()).hexdigest()})
    tt = re.search(r'TX_TYPE_TO_SCAN_INDEX: \[u8; 16\] = \[(.*?)\]', s).group(1)
    with open(os.path.join(HERE, 'scan_golden.json'), 'w') as f:
        json.dump({'source': 'rust/crates/svtav1-entropy/src/scan_tables.rs',
                   'tx_type_to_scan_index': [int(x) for x in tt.split(',')], 'tables': out}, f)
    print('wrote %d scan digests' % len(out))


if __name__ == '__main__':
    main()
