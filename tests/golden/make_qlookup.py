"""Extract the AV1 Dc_Qlookup / Ac_Qlookup tables (8- and 10-bit) that the
reference holds as data and write them as a JSON fixture plus the two C data
headers (oracle copy and product copy).

Sources (read as text in the build container only):
  rust/crates/svtav1-dsp/src/quant_tables.rs:9,30         DC/AC_QLOOKUP_8
  rust/crates/svtav1-encoder/src/bd10_qlookup_tables.rs:5 DC/AC_QLOOKUP_10
These are the normative AV1 spec tables (section 7.12.2, Dc_Qlookup/Ac_Qlookup),
so every implementation carries identical numbers.
"""
import json
import os
import re
import sys

REF = '/root/reference/rust/crates/'
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def grab(path, name):
    s = open(path).read()
    m = re.search(r'static %s: \[i16; 256\] = \[(.*?)\];' % name, s, re.S)
    v = [int(x) for x in re.findall(r'-?\d+', m.group(1))]
    assert len(v) == 256, (name, len(v))
    return v


def c_header(tables, guard):
    out = ['/* AV1 Dc_Qlookup / Ac_Qlookup (spec 7.12.2) for bit depths 8 and 10.',
           ' * Generated by tests/golden/make_qlookup.py; data only. */',
           '#ifndef %s' % guard, '#define %s' % guard, '#include <stdint.h>']
    for key, cname in (('dc8', 'DC_QLOOKUP_8'), ('ac8', 'AC_QLOOKUP_8'),
                       ('dc10', 'DC_QLOOKUP_10'), ('ac10', 'AC_QLOOKUP_10')):
        vals = tables[key]
        out.append('static const int16_t %s[256] = {' % cname)
        for i in range(0, 256, 16):
            out.append('  ' + ', '.join(str(v) for v in vals[i:i + 16]) + ',')
        out.append('};')
    out.append('#endif')
    return '\n'.join(out) + '\n'


def main():
    t = {
        'dc8': grab(REF + 'svtav1-dsp/src/quant_tables.rs', 'DC_QLOOKUP_8'),
        'ac8': grab(REF + 'svtav1-dsp/src/quant_tables.rs', 'AC_QLOOKUP_8'),
        'dc10': grab(REF + 'svtav1-encoder/src/bd10_qlookup_tables.rs', 'DC_QLOOKUP_10'),
        'ac10': grab(REF + 'svtav1-encoder/src/bd10_qlookup_tables.rs', 'AC_QLOOKUP_10'),
    }
    with open(os.path.join(HERE, 'av1_qlookup.json'), 'w') as f:
        json.dump(t, f)
    with open(os.path.join(ROOT, 'oracle', 'av1_qlookup_tables.h'), 'w') as f:
        f.write(c_header(t, 'ORACLE_AV1_QLOOKUP_TABLES_H'))
    with open(os.path.join(ROOT, 'zenav1-svt_amd', 'csrc', 'av1_qlookup_tables.h'), 'w') as f:
        f.write(c_header(t, 'ZAV1_AV1_QLOOKUP_TABLES_H'))
    print('wrote qlookup fixture + headers')


if __name__ == '__main__':
    sys.exit(main())
