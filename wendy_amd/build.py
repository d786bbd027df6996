"""Compile libwendy_amd.so (HIP for gfx950) in-tree.

    python -m wendy_amd.build

Every HIP source under ``wendy_amd/csrc`` is compiled to an object in
``wendy_amd/_build`` (in parallel), then linked into
``wendy_amd/libwendy_amd.so``.  Flags:

* ``-ffp-contract=off``: no fused multiply-add contraction, so every product
  and sum rounds separately, as in the reference's x86-64 build;
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32
  instructions are generated.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
BUILD = os.path.join(HERE, '_build')
LIB = os.path.join(HERE, 'libwendy_amd.so')
ARCH = os.environ.get('WDY_OFFLOAD_ARCH', 'gfx950')

COMMON = ['-O3', '-std=c++17', '-fPIC', '-ffp-contract=off',
          '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
          '-Wall', '-Wno-unused-result',
          '-I', os.path.join(ROOT, 'include')]


def hipcc():
    return os.environ.get('HIPCC', shutil.which('hipcc') or '/opt/rocm/bin/hipcc')


HOST_FLAGS = ['-O2', '-std=c++17', '-fPIC', '-ffp-contract=off', '-Wall',
              '-I', os.path.join(ROOT, 'include')]


def sources():
    """HIP sources (device + host code, built by hipcc for gfx950) and plain
    C++ host sources (built by g++)."""
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith('.hip') or f.endswith('.cpp'))


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    procs = []
    objs = []
    for src in sources():
        base, ext = os.path.splitext(os.path.basename(src))
        obj = os.path.join(BUILD, base + ('.o' if ext == '.hip' else '.host.o'))
        objs.append(obj)
        if ext == '.hip':
            cmd = [cc, '--offload-arch=' + ARCH, *COMMON, '-c', src, '-o', obj]
        else:
            cmd = [os.environ.get('CXX', 'g++'), *HOST_FLAGS, '-c', src, '-o', obj]
        if verbose:
            print(' '.join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT)))
    failed = []
    for src, p in procs:
        out = p.communicate()[0].decode(errors='replace')
        lines = [ln for ln in out.splitlines()
                 if 'packed-fp32-ops' not in ln or 'error' in ln]
        if p.returncode != 0:
            failed.append((src, '\n'.join(lines)))
        elif verbose and lines:
            print('\n'.join(lines))
    if failed:
        for src, out in failed:
            sys.stderr.write('--- %s\n%s\n' % (src, out))
        raise RuntimeError('HIP compilation failed: %s'
                           % ', '.join(os.path.basename(s) for s, _ in failed))
    tmp = LIB + '.tmp'
    cmd = [cc, '--offload-arch=' + ARCH, '-shared', '-fPIC', *objs, '-o', tmp]
    if verbose:
        print(' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
