"""Build libmacromax_hip.so in-tree with hipcc for gfx950 (no torch extension machinery involved).

Usage: ``python -m macromax_amd.build`` (or ``build()`` from Python). The library is written next to this file.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB_NAME = 'libmacromax_hip.so'
LIB_PATH = os.path.join(HERE, LIB_NAME)
SOURCES = ['mmx_capi.hip']
# the streaming-ceiling probe of bench.py (a measurement aid, not loaded by the solver)
PROBE_SOURCE = 'mmx_stream_probe.hip'
PROBE_LIB_PATH = os.path.join(HERE, 'libmmx_stream_probe.so')
HEADERS = ['mmx_plan.hpp', 'mmx_device.hpp', 'mmx_kernels.hpp', 'mmx_magnetic.hpp', 'mmx_magfused.hpp',
           'mmx_pointwise.hpp', 'mmx_regfft.hpp', 'mmx_regkernels.hpp', 'mmx_setup.hpp']
ARCH = os.environ.get('MMX_OFFLOAD_ARCH', 'gfx950')


def hipcc_command(out_path: str = LIB_PATH, sources=None) -> list[str]:
    hipcc = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    return [
        hipcc, f'--offload-arch={ARCH}', '-O3', '-std=c++17', '-fPIC', '-shared',
        # packed-FP32 VALU forms are kept out of every kernel
        '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
        '-Wall', '-Wno-unused-function', '-Wno-unused-variable',
        '-I', os.path.join(HERE, '..', 'include'),
        *[os.path.join(CSRC, s) for s in (SOURCES if sources is None else sources)],
        '-o', out_path,
    ]


def _compile(out_path: str, sources, verbose: bool):
    tmp = out_path + '.tmp'
    cmd = hipcc_command(tmp, sources)
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f'hipcc failed with exit code {res.returncode}')
    if verbose and res.stderr.strip():
        # keep warnings visible, but the host pass's note about the packed-fp32 feature is expected
        lines = [ln for ln in res.stderr.splitlines() if 'packed-fp32-ops' not in ln]
        if lines:
            print('\n'.join(lines))
    os.replace(tmp, out_path)


def build(verbose: bool = True) -> str:
    """Compile the HIP library and the streaming probe (always recompiles); return the library's path."""
    _compile(PROBE_LIB_PATH, [PROBE_SOURCE], verbose)
    _compile(LIB_PATH, SOURCES, verbose)
    return LIB_PATH


if __name__ == '__main__':
    build()
