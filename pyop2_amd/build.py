"""Build ``libop2hip.so`` in-tree with hipcc for gfx950.

Usage: ``python -m pyop2_amd.build``.  The library goes to
``pyop2_amd/lib/libop2hip.so``; it needs no GPU to build.
"""

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIBDIR = os.path.join(HERE, 'lib')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = os.environ.get('OP2HIP_ARCH', 'gfx950')

SOURCES = ['runtime.hip', 'plan.cpp', 'plan_device.hip', 'plan_colour.hip', 'sparsity.hip',
           'linalg.hip',
           'cg_batch.hip', 'krylov.hip', 'staging.hip']

# Device code is built without packed FP32 instructions (some of their forms
# have given wrong results on this hardware) and without FP contraction, so
# device FP64 arithmetic follows the host's IEEE order.
DEVICE_FLAGS = ['--offload-arch=%s' % ARCH, '-O3', '-std=c++17', '-fPIC',
                '-ffp-contract=off',
                '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops']


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       universal_newlines=True)
    if p.returncode != 0:
        sys.stderr.write(p.stdout)
        raise RuntimeError("command failed: %s" % ' '.join(cmd))
    return p.stdout


def build(verbose=False):
    """Compile every source to an object and link the shared library."""
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, 'obj')
    os.makedirs(objdir, exist_ok=True)
    objs, cmds = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, os.path.splitext(src)[0] + '.o')
        if src.endswith('.hip'):
            cmd = [HIPCC, '-x', 'hip'] + DEVICE_FLAGS + ['-c', path, '-o', obj]
        else:
            cmd = [HIPCC, '-O3', '-std=c++17', '-fPIC', '-c', path, '-o', obj]
        if verbose:
            print(' '.join(cmd))
        cmds.append(cmd)
        objs.append(obj)
    # the sources compile independently: one hipcc process each
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(len(cmds), 8)) as pool:
        list(pool.map(_run, cmds))
    out = os.path.join(LIBDIR, 'libop2hip.so')
    tmp = out + '.tmp'
    cmd = [HIPCC, '--offload-arch=%s' % ARCH, '-shared', '-fPIC', '-o', tmp] + objs + \
        ['-lhiprtc', '-Wl,-rpath,/opt/rocm/lib']
    if verbose:
        print(' '.join(cmd))
    _run(cmd)
    os.replace(tmp, out)
    return out


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
