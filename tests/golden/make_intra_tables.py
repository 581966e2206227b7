"""Extract the AV1 intra-prediction data tables the reference holds
(spec tables: Sm_Weights_Tx_*, Dr_Intra_Derivative, Intra_Filter_Taps,
Mode_To_Angle) into tests/golden/intra_tables.json plus C data headers for
the oracle and the product.

Sources (read as text in the build container only):
  rust/crates/svtav1-dsp/src/intra_pred.rs:319 DR_INTRA_DERIVATIVE,
    :883-896 SM_WEIGHTS_*, :913 FILTER_INTRA_TAPS
  rust/crates/svtav1-encoder/src/intra_edge.rs:896 MODE_TO_ANGLE_MAP
"""
import json
import os
import re

REF = '/root/reference/rust/crates/'
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def nums(s):
    return [int(x) for x in re.findall(r'-?\d+', s)]


def main():
    ip = open(REF + 'svtav1-dsp/src/intra_pred.rs').read()
    ie = open(REF + 'svtav1-encoder/src/intra_edge.rs').read()
    t = {}
    t['dr_intra_derivative'] = nums(re.search(r'static DR_INTRA_DERIVATIVE: \[u16; 90\] = \[(.*?)\];', ip, re.S).group(1))
    for n in (4, 8, 16, 32, 64):
        t['sm_weights_%d' % n] = nums(re.search(r'static SM_WEIGHTS_%d: \[u8; %d\] = \[(.*?)\];' % (n, n), ip, re.S).group(1))
    taps = nums(re.search(r'static FILTER_INTRA_TAPS: \[\[\[i8; 8\]; 8\]; 5\] = \[(.*?)\];', ip, re.S).group(1))
    t['filter_intra_taps'] = [taps[m * 64:(m + 1) * 64] for m in range(5)]
    t['mode_to_angle'] = nums(re.search(r'MODE_TO_ANGLE_MAP: \[i32; 13\] = \[(.*?)\];', ie, re.S).group(1))
    assert len(t['dr_intra_derivative']) == 90 and len(taps) == 320 and len(t['mode_to_angle']) == 13
    with open(os.path.join(HERE, 'intra_tables.json'), 'w') as f:
        json.dump(t, f)

    def header(guard, qual):
        L = ['/* AV1 intra-prediction spec tables (Sm_Weights, Dr_Intra_Derivative,',
             ' * Intra_Filter_Taps, Mode_To_Angle).  Generated by',
             ' * tests/golden/make_intra_tables.py; data only. */',
             '#ifndef %s' % guard, '#define %s' % guard, '#include <stdint.h>']
        L.append('%sstatic const uint16_t DR_INTRA_DERIVATIVE[90] = {%s};' % (qual, ', '.join(map(str, t['dr_intra_derivative']))))
        sm = []
        for n in (4, 8, 16, 32, 64):
            sm += t['sm_weights_%d' % n]
        L.append('/* concatenated weights for sizes 4, 8, 16, 32, 64: offsets 0, 4, 12, 28, 60 */')
        L.append('%sstatic const uint8_t SM_WEIGHTS_ALL[124] = {%s};' % (qual, ', '.join(map(str, sm))))
        L.append('%sstatic const int8_t FILTER_INTRA_TAPS[320] = {%s};  /* [mode][k][tap] */' % (qual, ', '.join(map(str, taps))))
        L.append('%sstatic const int16_t MODE_TO_ANGLE[13] = {%s};' % (qual, ', '.join(map(str, t['mode_to_angle']))))
        L.append('#endif')
        return '\n'.join(L) + '\n'
    open(os.path.join(ROOT, 'oracle', 'av1_intra_tables.h'), 'w').write(header('ORACLE_AV1_INTRA_TABLES_H', ''))
    open(os.path.join(ROOT, 'zenav1-svt_amd', 'csrc', 'av1_intra_tables.h'), 'w').write(
        header('ZAV1_AV1_INTRA_TABLES_H', '__device__ __constant__ '))
    print('wrote intra tables')


if __name__ == '__main__':
    main()
