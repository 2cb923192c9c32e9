"""Build libspk.so (HIP kernels + C ABI) for gfx950 with hipcc.

Usage:  python -m spinterps_amd.build        (or build() from Python)

The library is written next to this file so that it travels with the
source tree; nothing is installed.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / 'csrc'
OUT = HERE / 'libspk.so'
SOURCES = ['spk_core.hip', 'spk_factor.hip', 'spk_cells.hip', 'spk_neb.hip',
           'spk_vgfit.hip', 'spk_mlp.hip', 'spk_class.hip', 'spk_int8.hip',
           'spk_pinv.hip', 'spk_idw.hip', 'spk_nebk.hip', 'spk_bf16.hip']
# rocSOLVER (dsyevd) and rocBLAS (dgemm) for the pseudo-inverse fallback of
# steps whose covariance block does not factorise
LIBS = ['-L/opt/rocm/lib', '-Wl,-rpath,/opt/rocm/lib', '-lrocsolver',
        '-lrocblas']

HIPCC_FLAGS = [
    '--offload-arch=gfx950', '-O3', '-fPIC', '-shared', '-std=c++17',
    # packed FP32 instructions stay off: some of their forms have given
    # wrong results without an error on this hardware
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-result', '-Wno-unused-command-line-argument',
]


def hipcc_path():
    for cand in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc',
                 shutil.which('hipcc')):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError('hipcc not found: the HIP toolchain is required')


def build(verbose=False, jobs=None):
    """Compile every translation unit and link libspk.so."""
    hipcc = hipcc_path()
    objdir = HERE / '_build'
    objdir.mkdir(exist_ok=True)
    srcs = [SRC / s for s in SOURCES if (SRC / s).exists()]
    objs = []
    procs = []
    for s in srcs:
        o = objdir / (s.stem + '.o')
        objs.append(o)
        cmd = [hipcc] + [f for f in HIPCC_FLAGS if f != '-shared'] + [
            '-c', str(s), '-o', str(o)]
        if verbose:
            print(' '.join(cmd))
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                          stderr=subprocess.STDOUT)))
    errs = []
    for s, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            errs.append(f'--- {s.name} ---\n{out.decode(errors="replace")}')
        elif verbose and out:
            print(out.decode(errors='replace'))
    if errs:
        raise RuntimeError('hipcc failed:\n' + '\n'.join(errs))
    tmp = OUT.with_suffix('.so.tmp')
    cmd = [hipcc, '--offload-arch=gfx950', '-shared', '-fPIC', '-o', str(tmp)]
    cmd += [str(o) for o in objs] + LIBS
    if verbose:
        print(' '.join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed:\n' + r.stdout + r.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
