"""Write the AV1 default partition and skip CDFs (the frame-initial tables the
light partition pre-pass prices its PARTITION_NONE / PARTITION_SPLIT and
skip symbols from) as a data fixture (partition_cdfs.json) and as C headers
for the oracle and the library.

Source (read as text in the build container only):
  rust/crates/svtav1-entropy/src/context.rs:341  DEFAULT_PARTITION_CDF
  rust/crates/svtav1-entropy/src/context.rs:435  DEFAULT_SKIP_CDF

Both are AV1 specification data (Default_Partition_W8..W128_Cdf and
Default_Skip_Cdf), stored as inverse CDFs (32768 - cumulative), zero padded
to 11 entries per partition row: rows 0-3 (8x8) have 4 symbols, rows 4-15
(16x16 .. 64x64) 10 symbols, rows 16-19 (128x128) 8 symbols.
"""
import ast
import json
import os
import re

REF = '/root/reference/rust/crates/svtav1-entropy/src/context.rs'
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def rust_static(text, name):
    m = re.search(r'(?:pub\s+)?(?:static|const)\s+%s\s*:[^=]*=\s*' % re.escape(name), text)
    if not m:
        raise KeyError(name)
    i, depth, j = m.end(), 0, m.end()
    while True:
        ch = text[j]
        if ch == '[':
            depth += 1
        elif ch == ']':
            depth -= 1
            if depth == 0:
                break
        j += 1
    return ast.literal_eval(re.sub(r'//[^\n]*', '', text[i:j + 1]))


def header(data, guard):
    out = ['/* Generated by tests/golden/make_partition_cdfs.py: AV1 default partition',
           ' * and skip CDFs (inverse CDFs, zero padded).  Specification data; do not',
           ' * edit. */',
           '#ifndef %s' % guard, '#define %s' % guard, '#include <stdint.h>', '']
    p = data['DEFAULT_PARTITION_CDF']
    out.append('static const uint16_t AV1_PARTITION_CDF[%d][11] = {' % len(p))
    for row in p:
        out.append('  {' + ','.join(str(v) for v in row) + '},')
    out.append('};')
    s = data['DEFAULT_SKIP_CDF']
    out.append('static const uint16_t AV1_SKIP_CDF[%d][3] = {' % len(s))
    for row in s:
        out.append('  {' + ','.join(str(v) for v in row) + '},')
    out.append('};')
    out += ['', '#endif', '']
    return '\n'.join(out)


def main():
    text = open(REF).read()
    data = {n: rust_static(text, n) for n in ('DEFAULT_PARTITION_CDF', 'DEFAULT_SKIP_CDF')}
    assert len(data['DEFAULT_PARTITION_CDF']) == 20 and all(len(r) == 11 for r in data['DEFAULT_PARTITION_CDF'])
    assert len(data['DEFAULT_SKIP_CDF']) == 3
    fixture = {'source': 'rust/crates/svtav1-entropy/src/context.rs:341,:435 (AV1 spec data)'}
    fixture.update(data)
    with open(os.path.join(HERE, 'partition_cdfs.json'), 'w') as f:
        json.dump(fixture, f, separators=(',', ':'))
        f.write('\n')
    for path, guard in ((os.path.join(ROOT, 'oracle', 'av1_part_cdfs.h'), 'OR_AV1_PART_CDFS_H'),
                        (os.path.join(ROOT, 'zenav1-svt_amd', 'csrc', 'av1_part_cdfs.h'), 'ZAV1_AV1_PART_CDFS_H')):
        with open(path, 'w') as f:
            f.write(header(data, guard))


if __name__ == '__main__':
    main()
