"""CPU emulation of generated pipeline libraries (test infrastructure).

A generated pipeline is compiled for the host instead of the GPU: the HIP
kernel source with ``emu/emu_prelude.h`` (workgroup = ``blockDim.x`` host
threads meeting at a barrier, shared arrays as statics), the unchanged host
stub, and ``emu/emu_rt.cpp``, a CPU implementation of the libpmx_rt calls
the stub makes.  With ``asan=True`` everything is built with
AddressSanitizer, so an out-of-bounds global read, LDS slot overflow or
stray store in the emitted index arithmetic stops the run with a report,
on the CPU, before the kernel ever runs on a GPU.

Values match the GPU's bit for bit wherever only IEEE +, -, *, / and
conversions are involved; device math builtins (exp, pow, ...) use the
host libm here, so pipelines using them are compared with a tolerance.
"""

import ctypes
import hashlib
import os
import re
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
EMU = os.path.join(HERE, 'emu')
CACHE = os.environ.get('PMX_EMU_CACHE', os.path.join(REPO, 'tests', '.emu_cache'))

_SIG = re.compile(r'extern "C" __global__ (?:__launch_bounds__\(\d+\) )?void (\w+)\(([^)]*)\)')


def _split_args(s):
    return [a.strip() for a in s.split(',') if a.strip()]


def _arg_type(decl):
    decl = decl.replace('__restrict__', ' ')
    toks = decl.replace('*', ' * ').split()
    return ' '.join(toks[:-1])


def trampolines(hip_src):
    """C++ that registers a ``void(void**)`` trampoline per kernel."""
    out = []
    regs = []
    for m in _SIG.finditer(hip_src):
        name, args = m.group(1), _split_args(m.group(2))
        casts = ', '.join('*(%s*)a[%d]' % (_arg_type(d), i) for i, d in enumerate(args))
        out.append('static void pmx_emu_t_%s(void** a) { %s(%s); }' % (name, name, casts))
        regs.append('    pmx_emu_register("%s", &pmx_emu_t_%s);' % (name, name))
    out.append('extern "C" void pmx_emu_register(const char*, void (*)(void**));')
    out.append('namespace { struct pmx_emu_reg { pmx_emu_reg() {')
    out += regs
    out.append('} } pmx_emu_reg_instance; }')
    return '\n'.join(out) + '\n'


def asan_runtime():
    return subprocess.check_output(['gcc', '-print-file-name=libasan.so']).decode().strip()


def build(pipe, asan=False):
    """Compile ``pipe`` for the CPU emulator; returns an EmulatedPipeline."""
    code = pipe.generate_code()
    flags = ['-O1', '-g', '-std=c++17', '-fPIC', '-shared', '-pthread', '-w',
             '-ffp-contract=off', '-fno-fast-math']
    if asan:
        flags += ['-fsanitize=address', '-fno-omit-frame-pointer']
    key = hashlib.sha256((code.hip_src + '\0' + code.host_src + '\0' + ' '.join(flags) +
                          open(os.path.join(EMU, 'emu_rt.cpp')).read() +
                          open(os.path.join(EMU, 'emu_prelude.h')).read()).encode()).hexdigest()[:20]
    d = os.path.join(CACHE, '%s_%s' % (pipe.name, key))
    so = os.path.join(d, 'libemu_%s.so' % pipe.name)
    if not os.path.exists(so):
        os.makedirs(d, exist_ok=True)
        kern = os.path.join(d, 'kernels_emu.cpp')
        with open(kern, 'w') as f:
            f.write('#include "emu_prelude.h"\n' + code.hip_src + '\n' + trampolines(code.hip_src))
        stub = os.path.join(d, 'stub_emu.cpp')
        lines = []
        for l in code.host_src.split('\n'):
            if '.incbin' in l:
                l = 'extern "C" const unsigned char pmx_co_image[1] = {0};'
            lines.append(l)
        with open(stub, 'w') as f:
            f.write('\n'.join(lines))
        tmp = so + '.tmp%d' % os.getpid()
        cmd = ['g++'] + flags + ['-I' + os.path.join(EMU, 'include'), '-I' + EMU,
                                 '-I' + os.path.join(REPO, 'include'),
                                 os.path.join(EMU, 'emu_rt.cpp'), kern, stub, '-o', tmp,
                                 '-Wl,-Bsymbolic']
        subprocess.check_call(cmd)
        os.replace(tmp, so)
    return EmulatedPipeline(pipe, code, so)


class EmulatedPipeline(object):
    """Same calling convention as runtime.BuiltPipeline, host arrays only."""

    def __init__(self, pipe, code, so_path):
        self.pipe = pipe
        self.code = code
        self.lib = ctypes.CDLL(so_path)
        self.lib.pmx_last_error.restype = ctypes.c_char_p
        self.name = pipe.name
        self.params = pipe.sorted_parameters()
        self.inputs = pipe.sorted_inputs()
        self.outputs = pipe.sorted_outputs()

    def _check(self, rc, what):
        bad = ctypes.c_int.in_dll(self.lib, 'pmx_emu_misaligned')
        if bad.value:
            n, bad.value = bad.value, 0
            raise RuntimeError('%s: %d misaligned 128-bit loads' % (what, n))
        if rc != 0:
            raise RuntimeError('%s failed (%d): %s' % (what, rc,
                                                       self.lib.pmx_last_error().decode()))

    def _param_args(self, values):
        from polymage_amd.utils import convert_to_ctype
        return [convert_to_ctype(p.typ.c_type_name(), values[p.name]) for p in self.params]

    def run_host(self, values):
        args = self._param_args(values)
        for f in self.inputs + self.outputs:
            a = values[f.name]
            assert isinstance(a, np.ndarray) and a.flags['C_CONTIGUOUS'], f.name
            assert a.dtype == f.typ.np_dtype(), (f.name, a.dtype)
            args.append(ctypes.c_void_p(a.ctypes.data))
        fn = self.lib['pipeline_%s_host' % self.name]
        fn.restype = ctypes.c_int
        self._check(fn(*args), 'pipeline_%s_host' % self.name)

    def batch_create(self, frames):
        """frames: list of value dicts with host arrays; returns the handle
        of pipeline_<name>_batch_create."""
        from polymage_amd.utils import _CTYPE
        n = len(frames)
        args = [ctypes.c_int(n)]
        keep = []
        for p in self.params:
            ct = _CTYPE[p.typ.c_type_name()]
            arr = (ct * n)(*[fr[p.name] for fr in frames])
            keep.append(arr)
            args.append(arr)
        for f in self.inputs + self.outputs:
            arr = (ctypes.c_void_p * n)(*[fr[f.name].ctypes.data for fr in frames])
            keep.append(arr)
            args.append(arr)
        h = ctypes.c_void_p()
        args.append(ctypes.byref(h))
        self._check(self.lib['pipeline_%s_batch_create' % self.name](*args), 'batch_create')
        return h

    def batch_run(self, h, mode='persistent'):
        fn = {'persistent': 'batch_run', 'frames': 'batch_run_frames'}[mode]
        self._check(self.lib['pipeline_%s_%s' % (self.name, fn)](h, ctypes.c_void_p(0)), fn)

    def batch_destroy(self, h):
        self._check(self.lib['pipeline_%s_batch_destroy' % self.name](h), 'batch_destroy')

    def run_batch(self, frames, mode='persistent'):
        """frames: list of value dicts with host arrays; runs the batch API."""
        h = self.batch_create(frames)
        self.batch_run(h, mode)
        self.batch_destroy(h)
