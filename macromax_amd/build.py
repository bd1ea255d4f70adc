"""Build libmacromax_hip.so in-tree with hipcc for gfx950 (no torch extension machinery involved).

Usage: ``python -m macromax_amd.build`` (or ``build()`` from Python). Every source under ``csrc/`` (the C ABI and
one translation unit per register-kernel axis length) is compiled to an object in parallel, then the objects are
linked into the library next to this file.
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB_NAME = 'libmacromax_hip.so'
LIB_PATH = os.path.join(HERE, LIB_NAME)
ARCH = os.environ.get('MMX_OFFLOAD_ARCH', 'gfx950')


def sources() -> list[str]:
    """The C ABI (mmx_capi.hip) and the per-length register-kernel units (mmx_reg_<N>.hip)."""
    return sorted(glob.glob(os.path.join(CSRC, '*.hip')))


def _hipcc() -> str:
    return shutil.which('hipcc') or '/opt/rocm/bin/hipcc'


def compile_command(src: str, obj: str) -> list[str]:
    return [
        _hipcc(), f'--offload-arch={ARCH}', '-O3', '-std=c++17', '-fPIC', '-c',
        # packed-FP32 VALU forms are kept out of every kernel
        '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
        '-Wall', '-Wno-unused-function', '-Wno-unused-variable',
        '-I', os.path.join(HERE, '..', 'include'),
        src, '-o', obj,
    ]


def link_command(objs: list[str], out_path: str) -> list[str]:
    return [_hipcc(), f'--offload-arch={ARCH}', '-shared', '-fPIC', *objs, '-o', out_path]


def _jobs() -> int:
    env = os.environ.get('MAX_JOBS')
    n = int(env) if env and env.isdigit() else (os.cpu_count() or 4)
    return max(1, min(n, 16))


def build(verbose: bool = True) -> str:
    """Compile the HIP library (always recompiles) and return its path."""
    srcs = sources()
    with tempfile.TemporaryDirectory(prefix='mmx_build_') as tmpdir:
        objs = [os.path.join(tmpdir, os.path.basename(s)[:-4] + '.o') for s in srcs]
        cmds = [compile_command(s, o) for s, o in zip(srcs, objs)]
        if verbose:
            print(f'compiling {len(cmds)} units with {_jobs()} jobs:', flush=True)
            for c in cmds:
                print(' '.join(c), flush=True)
        with ThreadPoolExecutor(max_workers=_jobs()) as pool:
            results = list(pool.map(lambda c: subprocess.run(c, capture_output=True, text=True), cmds))
        failed = [(c, r) for c, r in zip(cmds, results) if r.returncode != 0]
        for c, r in failed:
            sys.stderr.write(' '.join(c) + '\n' + r.stdout + r.stderr)
        if failed:
            raise RuntimeError(f'hipcc failed for {len(failed)} unit(s)')
        if verbose:
            # keep warnings visible, but the host pass's note about the packed-fp32 feature is expected
            lines = [ln for r in results for ln in r.stderr.splitlines() if 'packed-fp32-ops' not in ln]
            if lines:
                print('\n'.join(lines))
        tmp = LIB_PATH + '.tmp'
        cmd = link_command(objs, tmp)
        if verbose:
            print(' '.join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f'hipcc link failed with exit code {res.returncode}')
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == '__main__':
    build()
