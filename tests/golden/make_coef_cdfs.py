"""Write the AV1 default coefficient CDFs and the tx-type CDFs as a data
fixture (coef_cdfs.json) and as C headers for the oracle and the library,
with the key-frame luma mode, angle-delta and filter-intra CDFs the intra
mode rates come from.

Source (read as text in the build container only):
  rust/crates/svtav1-entropy/src/default_cdfs.rs   (C-layout default CDFs,
      the AV1 specification's Default_*_Cdf tables, 4 q contexts)
  rust/crates/svtav1-entropy/src/context.rs:372     (AV1_PROB_COST)

The values are AV1 specification data.  The headers carry them as plain
arrays in the reference's C layout:
  TXB_SKIP     [q][txs_ctx 5][ctx 13][3]
  EOB_EXTRA    [q][txs_ctx 5][plane 2][ctx 9][3]
  DC_SIGN      [q][plane 2][ctx 3][3]
  EOB_FLAGn    [q][plane 2][ctx 2][k]       n = 16 .. 1024, k = log2(n) + 2
  COEFF_BASE_EOB [q][txs_ctx 5][plane 2][ctx 4][4]
  COEFF_BASE   [q][txs_ctx 5][plane 2][ctx 42][5]
  COEFF_BR     [q][txs_ctx 4][plane 2][ctx 21][5]
  INTRA_EXT_TX [eset 3][sq_tx 4][intra_dir 13][17]
  KF_Y         [above ctx 5][left ctx 5][14]      (key-frame luma mode)
  ANGLE_DELTA  [directional mode 8][8]
  FILTER_INTRA [block size 22][3]                 (use_filter_intra flag)
  FILTER_INTRA_MODE [6]
Every CDF row is an inverse CDF (32768 - cumulative), 0 marking the last
symbol, the trailing entry the adaptation counter.
"""
import ast
import json
import os
import re

REF = '/root/reference/rust/crates/svtav1-entropy/src/'
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

TABLES = ['TXB_SKIP_CDF', 'EOB_EXTRA_CDF', 'DC_SIGN_CDF', 'EOB_FLAG_CDF16', 'EOB_FLAG_CDF32', 'EOB_FLAG_CDF64',
          'EOB_FLAG_CDF128', 'EOB_FLAG_CDF256', 'EOB_FLAG_CDF512', 'EOB_FLAG_CDF1024', 'COEFF_BASE_EOB_CDF',
          'COEFF_BASE_CDF', 'COEFF_BR_CDF', 'INTRA_EXT_TX_CDF', 'KF_Y_CDF', 'ANGLE_DELTA_CDF', 'FILTER_INTRA_CDF',
          'FILTER_INTRA_MODE_CDF']


def rust_array(text, name):
    m = re.search(r'pub (?:static|const) %s\s*:[^=]*=\s*' % re.escape(name), text)
    if not m:
        raise KeyError(name)
    i, depth = m.end(), 0
    j = i
    while True:
        ch = text[j]
        if ch == '[':
            depth += 1
        elif ch == ']':
            depth -= 1
            if depth == 0:
                break
        j += 1
    body = re.sub(r'//[^\n]*', '', text[i:j + 1])
    return ast.literal_eval(body)


def dims(a):
    d = []
    while isinstance(a, list):
        d.append(len(a))
        a = a[0]
    return d


def nested(a):
    if isinstance(a[0], list):
        return '{' + ','.join(nested(x) for x in a) + '}' + ('\n' if not isinstance(a[0][0], list) else '')
    return '{' + ','.join(str(v) for v in a) + '}'


def header(data, guard):
    out = ['/* Generated by tests/golden/make_coef_cdfs.py: AV1 default coefficient and',
           ' * intra tx-type CDFs (inverse CDFs, C layout, 4 q contexts) and the',
           ' * probability cost table.  Specification data; do not edit. */',
           '#ifndef %s' % guard, '#define %s' % guard, '#include <stdint.h>', '']
    for name in TABLES:
        a = data[name]
        dd = ''.join('[%d]' % x for x in dims(a))
        out.append('static const uint16_t AV1_%s%s = %s;' % (name, dd, nested(a)))
    out.append('static const uint16_t AV1_PROB_COST[128] = {')
    pc = data['AV1_PROB_COST']
    for k in range(0, 128, 16):
        out.append('  ' + ','.join(str(v) for v in pc[k:k + 16]) + ',')
    out.append('};')
    out += ['', '#endif', '']
    return '\n'.join(out)


def main():
    text = open(REF + 'default_cdfs.rs').read()
    data = {name: rust_array(text, name) for name in TABLES}
    data['AV1_PROB_COST'] = rust_array(open(REF + 'context.rs').read(), 'AV1_PROB_COST')
    fixture = {'source': 'rust/crates/svtav1-entropy/src/default_cdfs.rs, context.rs:372 (AV1 spec data)',
               'dims': {k: dims(v) for k, v in data.items()}}
    fixture.update(data)
    with open(os.path.join(HERE, 'coef_cdfs.json'), 'w') as f:
        json.dump(fixture, f, separators=(',', ':'))
        f.write('\n')
    for path, guard in ((os.path.join(ROOT, 'oracle', 'av1_coef_cdfs.h'), 'OR_AV1_COEF_CDFS_H'),
                        (os.path.join(ROOT, 'zenav1-svt_amd', 'csrc', 'av1_coef_cdfs.h'), 'ZAV1_AV1_COEF_CDFS_H')):
        with open(path, 'w') as f:
            f.write(header(data, guard))
    print({k: dims(v) for k, v in data.items()})


if __name__ == '__main__':
    main()
