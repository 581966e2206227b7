"""Extract the 1D transform known-answer vectors that the reference's own
test file holds (values measured from C SVT-AV1 by the reference's
tools/extract_golden.c) into tests/golden/txfm1d_golden.json.

Source, read as text in the build container only:
  rust/svtav1/tests/golden_parity.rs:36-720
Each record: {"test", "line", "func", "arg", "input", "golden"}.
"""
import json
import os
import re

SRC = '/root/reference/rust/svtav1/tests/golden_parity.rs'
HERE = os.path.dirname(os.path.abspath(__file__))


def parse_array(expr):
    expr = expr.strip()
    m = re.match(r'\[\s*(-?\d+)(?:i32)?\s*;\s*(\d+)\s*\]$', expr)
    if m:
        return [int(m.group(1))] * int(m.group(2))
    m = re.match(r'\[(.*)\]$', expr, re.S)
    return [int(x) for x in re.findall(r'-?\d+(?=i32|\b)', m.group(1).replace('i32', ''))]


def main():
    text = open(SRC).read()
    lines = text.split('\n')
    out = []
    for m in re.finditer(r'#\[test\]\nfn (\w+)\(\) \{\n(.*?)\n\}\n', text, re.S):
        name, body = m.group(1), m.group(2)
        call = re.search(r'(?:svtav1_dsp::inv_txfm::)?(fdct\d+|fadst\d+|fidentity\d+|idct\d+|iadst\d+|iidentity\d+)'
                         r'\(&(\w+), &mut output, (\d+)\)', body)
        if not call or name.startswith('roundtrip'):
            continue
        func, arg = call.group(1), int(call.group(3))
        n = int(re.search(r'\d+$', func).group(0))
        mi = re.search(r'let (?:mut )?input = (\[.*?\]);', body, re.S)
        inp = parse_array(mi.group(1))
        if len(inp) != n:
            inp = (inp * n)[:n]
        for fm in re.finditer(r'for i in 0\.\.(\d+) \{\s*input\[i\] = i as i32 \* (\d+) - (\d+);\s*\}', body):
            inp = [i * int(fm.group(2)) - int(fm.group(3)) for i in range(int(fm.group(1)))]
        for am in re.finditer(r'\n\s*input\[(\d+)\] = (-?\d+);', body):
            inp[int(am.group(1))] = int(am.group(2))
        mg = re.search(r'let golden = (\[.*?\]);', body, re.S)
        if mg:
            gold = parse_array(mg.group(1))
        else:
            dc = re.search(r'assert_eq!\(output\[0\], (-?\d+)', body)
            gold = ([int(dc.group(1))] + [0] * (n - 1)) if dc else [0] * n
        assert len(gold) == n, (name, len(gold))
        line = text[:m.start()].count('\n') + 2
        out.append({'test': name, 'line': line, 'func': func, 'arg': arg, 'input': inp, 'golden': gold})
    with open(os.path.join(HERE, 'txfm1d_golden.json'), 'w') as f:
        json.dump({'source': 'rust/svtav1/tests/golden_parity.rs', 'vectors': out}, f, indent=0)
    print('wrote %d vectors' % len(out))


if __name__ == '__main__':
    main()
