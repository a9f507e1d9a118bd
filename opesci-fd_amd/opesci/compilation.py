"""Compiler driver: builds the gfx950 library (replaces opesci/compilation.py).

The reference compiles each generated model with g++/clang++/icpc
(compilation.py:26-104).  Here the kernels are hand-written HIP compiled once
for gfx950 into ``opesci-fd_amd/lib/libopesci_hip.so``; compiling a model only
lowers it (opesci.plan) and makes sure that library exists.  Build errors
raise RuntimeError naming the log file, like the reference.
"""
import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)
CSRC = os.path.join(ROOT, 'csrc')
LIB_DIR = os.path.join(ROOT, 'lib')
BUILD_DIR = os.path.join(ROOT, '_build')
LIB_NAME = 'libopesci_hip.so'
INCLUDE = os.path.join(os.path.dirname(ROOT), 'include')

ARCH = os.environ.get('OPESCI_OFFLOAD_ARCH', 'gfx950')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
HIP_FLAGS = ['-O3', '-std=c++17', '-fPIC', '--offload-arch=%s' % ARCH,
             '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
             '-Wno-unused-result', '-I', INCLUDE, '-I', CSRC]

__all__ = ['HIPCompiler', 'get_compiler', 'library_path', 'build_library']


def library_path():
    return os.path.join(LIB_DIR, LIB_NAME)


def _run(cmd, log):
    with open(log, 'w') as f:
        f.write('Compiling: %s\n' % ' '.join(cmd))
        f.flush()
        try:
            subprocess.check_call(cmd, stdout=f, stderr=f)
        except OSError:
            raise RuntimeError('OSError during compilation\nPlease check if compiler exists: %s' % cmd[0])
        except subprocess.CalledProcessError:
            raise RuntimeError('Error during compilation:\nCompilation command: %s\nLog file: %s'
                               % (' '.join(cmd), log))


def build_library(jobs=None, verbose=False):
    """Compile every csrc/*.hip for gfx950 and link libopesci_hip.so (always rebuilds)."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, '*.hip')))
    if not srcs:
        raise RuntimeError('no HIP sources under %s' % CSRC)
    objs = []

    def one(src):
        base = os.path.splitext(os.path.basename(src))[0]
        obj = os.path.join(BUILD_DIR, base + '.o')
        _run([HIPCC] + HIP_FLAGS + ['-c', src, '-o', obj], os.path.join(BUILD_DIR, base + '.log'))
        if verbose:
            print('compiled', src, file=sys.stderr)
        return obj

    jobs = jobs or min(8, len(srcs))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(one, srcs))
    out = library_path()
    tmp = out + '.tmp'
    _run([HIPCC, '-shared', '-fPIC', '--offload-arch=%s' % ARCH, '-o', tmp] + objs,
         os.path.join(BUILD_DIR, 'link.log'))
    os.replace(tmp, out)
    return out


class HIPCompiler(object):
    """hipcc for gfx950; compile() returns the library a lowered model runs on."""
    name = 'hip'

    def compile(self, src, out=None, shared=True):
        lib = library_path()
        if not os.path.exists(lib):
            build_library()
        print('Compiled:', lib)
        return lib

    @property
    def _ivdep(self):
        return None


def get_compiler(name):
    if name in (None, 'hip', 'hipcc', 'amdclang++', 'gfx950'):
        return HIPCompiler()
    if name in ('g++', 'gnu', 'icpc', 'intel', 'clang', 'clang++'):
        print("note: compiler %r selects a host C++ toolchain in opesci-fd; this backend "
              "always builds HIP for %s" % (name, ARCH))
        return HIPCompiler()
    raise ValueError('Unknown compiler.')
