"""Build libfrhip.so for gfx950 in-tree with hipcc.

Steps: generate the structured operator kernels (csrc/gen_ops.py) into
csrc/_build/, then compile every HIP source into one shared library next to
this file.  Run as ``python -m pyfr_amd.build``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
BUILD = os.path.join(CSRC, '_build')
LIB = os.path.join(HERE, 'libfrhip.so')

SOURCES = ['frhip_runtime.hip', 'frhip_ops.hip', 'frhip_dense.hip',
           'frhip_pointwise.hip', 'frhip_bcs.hip', 'frhip_blasext.hip']

HIPFLAGS = [
    '--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-command-line-argument', '-I' + CSRC,
]


def hipcc():
    for c in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', 'hipcc'):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError('hipcc not found')


def build(verbose=False, jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    for f in os.listdir(BUILD):
        if f.startswith('gen_ops'):
            os.remove(os.path.join(BUILD, f))
    subprocess.run([sys.executable, os.path.join(CSRC, 'gen_ops.py'), BUILD],
                   check=True)
    gens = sorted(os.path.join(BUILD, f) for f in os.listdir(BUILD)
                  if f.startswith('gen_ops') and f.endswith('.hip'))

    # Compile each translation unit to an object in parallel, then link.
    srcs = [os.path.join(CSRC, s) for s in SOURCES] + gens
    objs, procs = [], []
    for s in srcs:
        o = os.path.join(BUILD, os.path.basename(s) + '.o')
        objs.append(o)
        cmd = [hipcc(), *[f for f in HIPFLAGS if f != '-shared'], '-c', s,
               '-o', o]
        if verbose:
            print(' '.join(cmd))
        procs.append(subprocess.Popen(cmd))
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError('hipcc compilation failed')
    tmp = LIB + '.tmp'
    cmd = [hipcc(), '--offload-arch=gfx950', '-shared', '-fPIC', *objs,
           '-o', tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose=True))
