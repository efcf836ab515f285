"""Build recipe for libsrf_hip.so (gfx950), written in-tree next to this file.

Run as ``python -m srf_hip.build`` (with simple-rf_amd/ on sys.path) or call ``build()``.
Each .hip source is compiled to an object in parallel, then linked into one shared library.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
CSRC_DIR = PKG_DIR.parent / 'csrc'
REPO_DIR = PKG_DIR.parent.parent
LIB_PATH = PKG_DIR / 'libsrf_hip.so'
OBJ_DIR = PKG_DIR.parent / '_build'

ARCH = os.environ.get('SRF_OFFLOAD_ARCH', 'gfx950')

HIPCC_FLAGS = [
    f'--offload-arch={ARCH}',
    '-O3',
    '-std=c++17',
    '-fPIC',
    # every rounding step of the reference ops is spelled out in the sources
    '-ffp-contract=off',
    # no packed-fp32 (v_pk_*_f32) code generation
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wall', '-Wno-unused-function', '-Wno-unused-variable', '-Wno-unused-but-set-variable',
    f'-I{REPO_DIR / "include"}',
]


def _hipcc() -> str:
    for cand in (shutil.which('hipcc'), '/opt/rocm/bin/hipcc'):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError('hipcc not found: the HIP toolchain (ROCm) is required to build libsrf_hip.so')


def sources() -> list[Path]:
    return sorted(CSRC_DIR.glob('*.hip'))


def _up_to_date(src: Path, obj: Path) -> bool:
    if not obj.exists():
        return False
    deps = [src, *CSRC_DIR.glob('*.h'), REPO_DIR / 'include' / 'srf_hip.h']
    return obj.stat().st_mtime > max(d.stat().st_mtime for d in deps if d.exists())


def _compile(src: Path, extra: list[str], incremental: bool = False) -> Path:
    obj = OBJ_DIR / (src.stem + '.o')
    if incremental and _up_to_date(src, obj):
        return obj
    cmd = [_hipcc(), *HIPCC_FLAGS, *extra, '-c', str(src), '-o', str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f'hipcc failed for {src.name}:\n{" ".join(cmd)}\n{res.stdout}\n{res.stderr}')
    return obj


def build(verbose: bool = False, extra_flags: list[str] | None = None, jobs: int | None = None,
          incremental: bool = False) -> Path:
    """Compile every csrc/*.hip for gfx950 and link libsrf_hip.so.

    Rebuilds every object unless ``incremental`` (development use: objects newer than their
    source and all headers are reused)."""
    OBJ_DIR.mkdir(exist_ok=True)
    srcs = sources()
    if not srcs:
        raise RuntimeError(f'no HIP sources under {CSRC_DIR}')
    extra = list(extra_flags or [])
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)), 16)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, extra, incremental), srcs))
    tmp = LIB_PATH.with_suffix('.so.tmp')
    cmd = [_hipcc(), f'--offload-arch={ARCH}', '-shared', '-fPIC', *map(str, objs), '-o', str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f'link failed:\n{" ".join(cmd)}\n{res.stdout}\n{res.stderr}')
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(f'built {LIB_PATH} from {len(srcs)} sources', file=sys.stderr)
    return LIB_PATH


if __name__ == '__main__':
    build(verbose=True, incremental='--incremental' in sys.argv)
