"""Python side of the runtime: build, cache, load and call pipelines.

Flow for a Pipeline (the MI355X analogue of the reference's
``generate_code`` -> ``g++`` -> ``ctypes.cdll`` path, tuner.py:322-350,
525-558 and apps/python/{cpp_compiler,loader}.py):

  1. ``pipe.generate_code()`` -> HIP kernels + C++ host stub;
  2. kernels -> code object with hiprtc (``pmx_compile_hip`` in
     libpmx_rt.so; works without a GPU), cached by source hash under
     ``polymage_amd/_kcache``;
  3. host stub (code object embedded with .incbin) -> ``lib<name>_<hash>.so``
     with g++, linked to libpmx_rt.so and libamdhip64;
  4. ``ctypes.CDLL`` -> ``pipeline_<name>`` (host pointers, reference ABI)
     and ``pipeline_<name>_dev`` (device pointers + stream).

Nothing here falls back to a CPU path: if libpmx_rt.so is missing the
import raises.
"""

import ctypes
import hashlib
import os
import subprocess
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
INCLUDE = os.path.join(REPO, 'include')
RT_LIB = os.path.join(HERE, 'libpmx_rt.so')
CACHE = os.environ.get('PMX_KCACHE', os.path.join(HERE, '_kcache'))
ROCM = os.environ.get('ROCM_PATH', '/opt/rocm')
ARCH = os.environ.get('PMX_ARCH', 'gfx950')

# Device code flags: no FP contraction (results must equal the reference
# C semantics bit for bit), no packed-FP32 instructions.
HIPRTC_OPTS = ['-O3', '-ffp-contract=off', '-std=c++17',
               '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops']

_lock = threading.Lock()
_rt = None


class RuntimeMissing(RuntimeError):
    pass


def rt():
    """Load libpmx_rt.so (raises if it was not built)."""
    global _rt
    if _rt is None:
        if not os.path.exists(RT_LIB):
            raise RuntimeMissing("libpmx_rt.so not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(RT_LIB, mode=ctypes.RTLD_GLOBAL)
        lib.pmx_last_error.restype = ctypes.c_char_p
        lib.pmx_compile_hip.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                        ctypes.POINTER(ctypes.c_char_p), ctypes.c_int,
                                        ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t]
        lib.pmx_arena_bytes.restype = ctypes.c_size_t
        _rt = lib
    return _rt


def last_error():
    return rt().pmx_last_error().decode(errors='replace')


def check(rc, what='pmx'):
    if rc != 0:
        raise RuntimeError('%s failed (%d): %s' % (what, rc, last_error()))


def compile_hip(src, out_path, name='kernels.hip', opts=None):
    opts = HIPRTC_OPTS if opts is None else opts
    arr = (ctypes.c_char_p * len(opts))(*[o.encode() for o in opts])
    log = ctypes.create_string_buffer(1 << 16)
    rc = rt().pmx_compile_hip(src.encode(), name.encode(), ARCH.encode(), arr, len(opts),
                              out_path.encode(), log, len(log))
    if rc != 0:
        raise RuntimeError('hiprtc compile failed: %s\n%s' % (last_error(), log.value.decode(errors='replace')[-4000:]))
    return log.value.decode(errors='replace')


def host_compile(cpp, so):
    cmd = ['g++', '-O2', '-std=c++17', '-fPIC', '-shared', '-D__HIP_PLATFORM_AMD__',
           '-I' + INCLUDE, '-I' + os.path.join(ROCM, 'include'), cpp, '-o', so,
           '-L' + HERE, '-lpmx_rt', '-L' + os.path.join(ROCM, 'lib'), '-lamdhip64',
           '-Wl,-rpath,$ORIGIN/..', '-Wl,-rpath,$ORIGIN/../..', '-Wl,-rpath,' + HERE,
           '-Wl,-rpath,' + os.path.join(ROCM, 'lib'),
           '-w']
    subprocess.check_call(cmd)


class BuiltPipeline(object):
    """A compiled pipeline: ctypes handles to both entry points."""

    def __init__(self, pipe, code, so_path):
        self.pipe = pipe
        self.code = code
        self.so_path = so_path
        rt()
        self.lib = ctypes.CDLL(so_path)
        self.name = pipe.name
        self.params = pipe.sorted_parameters()
        self.inputs = pipe.sorted_inputs()
        self.outputs = pipe.sorted_outputs()
        self.fn_host = self.lib['pipeline_' + self.name]
        self.fn_host.restype = None
        self.fn_host_rc = self.lib['pipeline_%s_host' % self.name]
        self.fn_host_rc.restype = ctypes.c_int
        self.fn_dev = self.lib['pipeline_%s_dev' % self.name]
        self.fn_dev.restype = ctypes.c_int

    def _param_args(self, values):
        from .utils import convert_to_ctype
        return [convert_to_ctype(p.typ.c_type_name(), values[p.name]) for p in self.params]

    def run_host(self, values):
        """Reference-style call on host ndarrays (synchronous)."""
        args = self._param_args(values)
        keep = []
        for f in self.inputs + self.outputs:
            a = values[f.name]
            assert isinstance(a, np.ndarray) and a.flags['C_CONTIGUOUS'], f.name
            assert a.dtype == f.typ.np_dtype(), (f.name, a.dtype)
            keep.append(a)
            args.append(ctypes.c_void_p(a.ctypes.data))
        check(self.fn_host_rc(*args), 'pipeline_%s' % self.name)

    # -- batches of frames (persistent work-queue launches) ----------------
    def batch_create(self, frames):
        """frames: list of dicts like run_dev's ``values`` (device
        pointers).  Returns an opaque handle for batch_run*."""
        from .utils import _CTYPE
        n = len(frames)
        args = [ctypes.c_int(n)]
        keep = []
        for p in self.params:
            ct = _CTYPE[p.typ.c_type_name()]
            arr = (ct * n)(*[fr[p.name] for fr in frames])
            keep.append(arr)
            args.append(arr)
        for f in self.inputs + self.outputs:
            ptrs = []
            for fr in frames:
                v = fr[f.name]
                ptrs.append(v.data_ptr() if hasattr(v, 'data_ptr') else int(v))
            arr = (ctypes.c_void_p * n)(*ptrs)
            keep.append(arr)
            args.append(arr)
        h = ctypes.c_void_p()
        args.append(ctypes.byref(h))
        check(self.lib['pipeline_%s_batch_create' % self.name](*args), 'batch_create')
        return h

    def batch_run(self, handle, stream=0, mode='persistent'):
        fn = {'persistent': 'batch_run', 'graph': 'batch_run_graph',
              'frames': 'batch_run_frames'}[mode]
        check(self.lib['pipeline_%s_%s' % (self.name, fn)](handle, ctypes.c_void_p(stream)), fn)

    def batch_destroy(self, handle):
        check(self.lib['pipeline_%s_batch_destroy' % self.name](handle), 'batch_destroy')

    def status(self):
        """pipeline_<name>_status(): 0, or the negative code of a tile that
        exceeded its planned LDS box since the last call (the word is then
        cleared).  Synchronises the null stream."""
        fn = self.lib['pipeline_%s_status' % self.name]
        fn.restype = ctypes.c_int
        return fn()

    def run_dev(self, values, stream=0):
        """Device call: values[name] are device pointers (ints) or objects
        with ``data_ptr()``; ``stream`` is a hipStream_t as int."""
        args = self._param_args(values)
        for f in self.inputs + self.outputs:
            v = values[f.name]
            ptr = v.data_ptr() if hasattr(v, 'data_ptr') else int(v)
            args.append(ctypes.c_void_p(ptr))
        args.append(ctypes.c_void_p(stream))
        check(self.fn_dev(*args), 'pipeline_%s_dev' % self.name)

    def bind_dev(self, values, stream=0):
        """A zero-argument callable that enqueues ``run_dev(values, stream)``
        with the ctypes arguments marshalled once (for launch loops whose
        kernels are shorter than Python's per-call marshalling)."""
        args = self._param_args(values)
        for f in self.inputs + self.outputs:
            v = values[f.name]
            args.append(ctypes.c_void_p(v.data_ptr() if hasattr(v, 'data_ptr') else int(v)))
        args.append(ctypes.c_void_p(stream))
        fn, name = self.fn_dev, 'pipeline_%s_dev' % self.name

        def launch():
            rc = fn(*args)
            if rc:
                check(rc, name)
        return launch


def load(so_path, pipe):
    """Load an already compiled pipeline library for ``pipe``'s signature."""
    return BuiltPipeline(pipe, None, so_path)


def split_source(text):
    """(hip_src, host_src) of a generated pipeline's source text, the
    ``str()`` of ``Pipeline.generate_code()`` (kernels, then host stub)."""
    from .hip_emit import HIP_MARK, HOST_MARK
    if not text.startswith(HIP_MARK) or HOST_MARK not in text:
        raise ValueError('not the source of a generated pipeline (missing the kernel / host stub markers)')
    hip, host = text[len(HIP_MARK):].split(HOST_MARK, 1)
    return hip, host


def build_sources(name, hip_src, host_src):
    """Compile (cached by source hash) the kernels with hiprtc and the host
    stub with g++; returns the path of ``lib<name>.so``."""
    key = hashlib.sha256((hip_src + '\0' + host_src + '\0' +
                          ' '.join(HIPRTC_OPTS) + ARCH).encode()).hexdigest()[:20]
    d = os.path.join(CACHE, '%s_%s' % (name, key))
    so = os.path.join(d, 'lib%s.so' % name)
    with _lock:
        if not os.path.exists(so):
            os.makedirs(d, exist_ok=True)
            hip = os.path.join(d, 'kernels.hip')
            co = os.path.join(d, 'kernels.co')
            with open(hip, 'w') as f:
                f.write(hip_src)
            if not os.path.exists(co):
                compile_hip(hip_src, co)
            cpp = os.path.join(d, 'stub.cpp')
            with open(cpp, 'w') as f:
                f.write(host_src.replace('@CO_PATH@', co))
            tmp = so + '.tmp%d' % os.getpid()
            host_compile(cpp, tmp)
            os.replace(tmp, so)
    return so


def build(pipe, verbose=False):
    """Generate, compile (cached) and load a pipeline."""
    code = pipe.generate_code()
    so = build_sources(pipe.name, code.hip_src, code.host_src)
    return BuiltPipeline(pipe, code, so)


# ---------------------------------------------------------------------------
# the C++/OpenMP backend (cpu_emit): the reference's host code generator
# ---------------------------------------------------------------------------

class CpuPipeline(object):
    """A pipeline compiled by the C++/OpenMP backend: ``run_host`` calls
    ``pipeline_<name>_host`` on host arrays (the reference's entry point,
    returning a status)."""

    def __init__(self, pipe, src, so_path):
        self.pipe = pipe
        self.src = src
        self.so_path = so_path
        self.lib = ctypes.CDLL(so_path)
        self.name = pipe.name
        self.params = pipe.sorted_parameters()
        self.inputs = pipe.sorted_inputs()
        self.outputs = pipe.sorted_outputs()
        self.fn = self.lib['pipeline_%s_host' % self.name]
        self.fn.restype = ctypes.c_int
        self.err = self.lib['pipeline_%s_error' % self.name]
        self.err.restype = ctypes.c_char_p

    def run_host(self, values):
        from .utils import convert_to_ctype
        args = [convert_to_ctype(p.typ.c_type_name(), values[p.name]) for p in self.params]
        keep = []
        for f in self.inputs + self.outputs:
            a = values[f.name]
            assert isinstance(a, np.ndarray) and a.flags['C_CONTIGUOUS'], f.name
            assert a.dtype == f.typ.np_dtype(), (f.name, a.dtype)
            keep.append(a)
            args.append(ctypes.c_void_p(a.ctypes.data))
        rc = self.fn(*args)
        if rc:
            raise RuntimeError('pipeline_%s (cpu) failed (%d): %s' % (self.name, rc,
                                                                     self.err().decode(errors='replace')))


def build_cpu(pipe, flags=None):
    """Emit the pipeline as C++/OpenMP (cpu_emit), compile it with g++
    (cached by source hash and flags; ``-march=native`` code is keyed by the
    host CPU model too) and load it."""
    from .cpu_emit import emit_cpu, CPU_FLAGS
    flags = list(CPU_FLAGS if flags is None else flags)
    src = emit_cpu(pipe)
    salt = ''
    if '-march=native' in flags:
        try:
            with open('/proc/cpuinfo') as f:
                salt = [l for l in f if l.startswith('model name')][0]
        except (OSError, IndexError):
            salt = os.uname().nodename
    key = hashlib.sha256((src + '\0' + ' '.join(flags) + salt).encode()).hexdigest()[:20]
    d = os.path.join(CACHE, 'cpu_%s_%s' % (pipe.name, key))
    so = os.path.join(d, 'libcpu_%s.so' % pipe.name)
    with _lock:
        if not os.path.exists(so):
            os.makedirs(d, exist_ok=True)
            cpp = os.path.join(d, 'pipeline.cpp')
            with open(cpp, 'w') as f:
                f.write(src)
            tmp = so + '.tmp%d' % os.getpid()
            subprocess.check_call(['g++'] + flags + [cpp, '-o', tmp])
            os.replace(tmp, so)
    return CpuPipeline(pipe, src, so)
