"""Build the zav1 gfx950 HIP library (libzav1_hip.so) in-tree.

Every source under csrc/ is compiled with hipcc for gfx950 only and linked
into one shared library whose exported symbols are the C-ABI declared in
include/zav1_dsp.h.  Packed-FP32 instruction selection is disabled for all
device code (-packed-fp32-ops off); integer kernels never need it.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
ROOT = os.path.dirname(HERE)
OUT = os.path.join(HERE, 'libzav1_hip.so')
BUILD = os.path.join(HERE, '_build')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

FLAGS = ['-O3', '-std=c++20', '-fPIC', '-fwrapv', '--offload-arch=gfx950',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-Wno-unused-command-line-argument', '-I' + CSRC, '-I' + os.path.join(ROOT, 'include')]


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith('.hip') or f.endswith('.cpp'))


def compile_one(src):
    obj = os.path.join(BUILD, src + '.o')
    cmd = [HIPCC, '-x', 'hip', '-c', os.path.join(CSRC, src), '-o', obj] + FLAGS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('compile failed: %s\n%s%s' % (' '.join(cmd), r.stdout, r.stderr))
    return obj


def build(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    subprocess.check_call([sys.executable, os.path.join(CSRC, 'gen_txfm1d.py')], stdout=subprocess.DEVNULL)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = OUT + '.tmp'
    cmd = [HIPCC, '-shared', '-o', tmp] + objs + ['--offload-arch=gfx950', '-fPIC']
    subprocess.check_call(cmd)
    os.replace(tmp, OUT)
    return OUT


if __name__ == '__main__':
    print(build())
