"""Build librgs_amd.so in-tree with hipcc for gfx950.

    python -m rgs70_classical_challenge_amd.build

Compiles every csrc/*.hip and csrc/*.cpp into one shared library next to this
file.  Always recompiles (the build check must compile every time).
"""
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
INCLUDE = os.path.join(ROOT, 'include')
LIB = os.path.join(HERE, 'librgs_amd.so')
ARCH = os.environ.get('RGS_OFFLOAD_ARCH', 'gfx950')

COMMON = ['-O3', '-std=c++17', '-fPIC', f'-I{INCLUDE}', f'-I{CSRC}',
          # packed FP32 instructions stay off (see the build notes)
          '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
          '-Wno-unused-command-line-argument']


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(('.hip', '.cpp')))


def build(verbose=True):
    hipcc = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
    objdir = os.path.join(HERE, '_objs')
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    t0 = time.time()
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src) + '.o')
        cmd = [hipcc, f'--offload-arch={ARCH}', '-c', src, '-o', obj] + COMMON
        if src.endswith('.cpp'):
            cmd = [hipcc, '-x', 'c++', '-c', src, '-o', obj, '-O3', '-std=c++17', '-fPIC', '-pthread',
                   f'-I{INCLUDE}', f'-I{CSRC}']
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = False
    for src, p in procs:
        out = p.communicate()[0].decode(errors='replace')
        out = '\n'.join(l for l in out.splitlines() if 'packed-fp32-ops' not in l and l.strip())
        if p.returncode != 0:
            failed = True
            sys.stderr.write(f'--- {src}\n{out}\n')
        elif verbose and out:
            sys.stderr.write(f'--- {src} (warnings)\n{out}\n')
    if failed:
        raise RuntimeError('hipcc failed')
    tmp = LIB + '.tmp'
    link = [hipcc, f'--offload-arch={ARCH}', '-shared', '-o', tmp] + objs + ['-fPIC', '-pthread']
    subprocess.run(link, check=True)
    os.replace(tmp, LIB)
    if verbose:
        print(f'built {LIB} in {time.time() - t0:.1f}s')
    return LIB


if __name__ == '__main__':
    build()
