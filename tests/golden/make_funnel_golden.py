"""Extract the MDS0 fast-loop records of the instrumented C encoder that the
reference holds as text captures, and write them as a JSON fixture.

Source (read as text in the build container only):
  rust/docs/captures/gradient_{64x64,128x128}_q{20,40,55}_p6.m6fnl.txt
  (8-bit gradient content y[r][c] = (r*255/h) ^ ((c*3) & 0x3f),
  rust/svtav1/examples/perf_encode.rs:29-42, preset 6)

Per block: origin, bsize, square side, the candidates in buffer order
(mode, filter-intra mode, MDS0 `dist` = Hadamard SATD, luma / chroma mode
rates, fast cost `fc`), the fast lambda (the full-loop lambda printed on the
block's FLC lines, which the fast cost also uses) and the POST0 survivor
order (buffer ids).  Blocks at the frame origin have no coded neighbours, so
their predictions (and SATDs) follow from the source alone.

A second fixture, funnel_full.json, keeps every block's full-loop records:
  mds1: the FLC st=1 lines with their MDS1 line (mode, filter-intra mode,
        luma coefficient rate `ybits`, frequency-domain distortion `ydist`,
        `eob0`): DCT_DCT, quantize_b, tx depth 0
  txt:  per MDS3 candidate (POST2 order, closed by its FLC st=3 line) the
        TXT records (depth, txb index, tx type, eob, rate `bits`, spatial
        distortion `dist` (SSE << 4), txb-skip and dc-sign contexts) and the
        TXSDEP lines
  win:  the block's winner (WIN line)
"""
import glob
import json
import os
import re

CAPS = '/root/reference/rust/docs/captures/'
HERE = os.path.dirname(os.path.abspath(__file__))


def parse(path):
    blocks = []
    cur = None
    for line in open(path):
        if not line.startswith('M6FNL '):
            continue
        kv = dict(re.findall(r'(\w+)=(-?\d+)', line))
        tag = line.split()[1]
        if tag == 'BLK':
            m = re.search(r'org=\((\d+),(\d+)\)', line)
            cur = {'org': [int(m.group(1)), int(m.group(2))], 'bsize': int(kv['bsize']), 'sq': int(kv['sq']),
                   'qidx': int(kv['qidx']), 'mds0': [], 'order': None, 'lambda': None}
            blocks.append(cur)
        elif tag == 'MDS0':
            cur['mds0'].append([int(kv[k]) for k in ('buf', 'mode', 'fi', 'dist', 'flr', 'fcr', 'fc')])
        elif tag == 'POST0' and cur['order'] is None:
            cur['order'] = [int(b) for b in re.findall(r'=buf(\d+)', line)]
        elif tag == 'FLC' and cur['lambda'] is None:
            cur['lambda'] = int(kv['lambda'])
    return [b for b in blocks if b['mds0'] and b['order'] is not None and b['lambda'] is not None]


def parse_full(path):
    blocks = []
    cur = None
    pend = []
    for line in open(path):
        if not line.startswith('M6FNL '):
            continue
        kv = dict(re.findall(r'(\w+)=(-?\d+)', line))
        tag = line.split()[1]
        if tag == 'BLK':
            m = re.search(r'org=\((\d+),(\d+)\)', line)
            cur = {'org': [int(m.group(1)), int(m.group(2))], 'bsize': int(kv['bsize']), 'sq': int(kv['sq']),
                   'qidx': int(kv['qidx']), 'lambda': None, 'mds1': [], 'post2': [], 'txt': [], 'win': None}
            blocks.append(cur)
            pend = []
        elif tag == 'FLC':
            if cur['lambda'] is None:
                cur['lambda'] = int(kv['lambda'])
            rec = {k: int(kv[k]) for k in ('mode', 'fi', 'ybits', 'ydist', 'flr', 'fcr', 'full')}
            if kv['st'] == '1':
                cur['mds1'].append(rec)
            elif kv['st'] == '3':
                cur['txt'].append({'mode': rec['mode'], 'fi': rec['fi'], 'ybits': rec['ybits'],
                                   'ydist': rec['ydist'], 'records': pend})
                pend = []
        elif tag == 'MDS1':
            last = cur['mds1'][-1]
            assert int(kv['mode']) == last['mode'] and int(kv['fi']) == last['fi']
            last['eob0'] = int(kv['eob0'])
        elif tag == 'POST2':
            cur['post2'] = [[int(a), int(b)] for a, b in re.findall(r'buf\d+:m(\d+):fi(\d+)', line)]
        elif tag == 'TXT':
            pend.append({'kind': 'txt', **{k: int(kv[k]) for k in ('dep', 'txb', 'tt', 'eob', 'bits', 'dist',
                                                                   'tsc', 'dsc')}})
        elif tag == 'TXSDEP':
            pend.append({'kind': 'txsdep', **{k: int(kv[k]) for k in ('dep', 'ybits', 'tszbit', 'dist')}})
        elif tag == 'WIN':
            cur['win'] = {k: int(kv[k]) for k in ('mode', 'fi', 'txdep', 'tt0', 'eoby0', 'full', 'rate', 'dist')}
    return blocks


def main():
    out = {'source': 'rust/docs/captures/gradient_*_p6.m6fnl.txt (MDS0 / POST0 / FLC lines)',
           'gradient_source': 'y[r][c] = (r*255/h) ^ ((c*3) & 0x3f)  (rust/svtav1/examples/perf_encode.rs:29-42)',
           'mds0_fields': ['buf', 'mode', 'fi', 'dist', 'flr', 'fcr', 'fc'], 'files': []}
    for p in sorted(glob.glob(CAPS + 'gradient_*_p6.m6fnl.txt')):
        m = re.search(r'gradient_(\d+)x(\d+)_q(\d+)_p6', p)
        out['files'].append({'name': os.path.basename(p), 'w': int(m.group(1)), 'h': int(m.group(2)),
                             'qp': int(m.group(3)), 'blocks': parse(p)})
    with open(os.path.join(HERE, 'funnel_mds0.json'), 'w') as f:
        json.dump(out, f, separators=(',', ':'))
        f.write('\n')
    print(sum(len(x['blocks']) for x in out['files']), 'blocks')
    full = {'source': 'rust/docs/captures/gradient_*_p6.m6fnl.txt (FLC / MDS1 / POST2 / TXT / TXSDEP / WIN lines)',
            'gradient_source': out['gradient_source'], 'files': []}
    for p in sorted(glob.glob(CAPS + 'gradient_*_p6.m6fnl.txt')):
        m = re.search(r'gradient_(\d+)x(\d+)_q(\d+)_p6', p)
        full['files'].append({'name': os.path.basename(p), 'w': int(m.group(1)), 'h': int(m.group(2)),
                              'qp': int(m.group(3)), 'blocks': parse_full(p)})
    with open(os.path.join(HERE, 'funnel_full.json'), 'w') as f:
        json.dump(full, f, separators=(',', ':'))
        f.write('\n')


if __name__ == '__main__':
    main()
