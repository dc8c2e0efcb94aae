"""Shared helpers for the test-suite: building pipelines at given sizes
and running them on the GPU and through the oracle."""

import numpy as np

from polymage_amd import buildPipeline
from polymage_amd.constructs import Condition


def build(out, pd, sizes, name, tile_sizes, group_size=11, literal_mode='double',
          gpu_options=None, params=None, grouping=None, inline=None):
    """``grouping`` (lists of function names) and ``inline`` (function
    names) are resolved in ``out``'s DAG."""
    est = [(pd[k], v) for k, v in sizes.items()]
    cons = [Condition(pd[k], '==', v) for k, v in sizes.items()]
    from polymage_amd.pipe import get_funcs
    funcs = dict((f.name, f) for f in get_funcs([out])[0])
    groups = [[funcs[n] for n in g] for g in (grouping or [])]
    return buildPipeline([out], param_estimates=est, param_constraints=cons,
                         grouping=groups, tile_sizes=tile_sizes, group_size=group_size,
                         inline_directives=[funcs[n] for n in (inline or [])],
                         pipe_name=name, literal_mode=literal_mode, gpu_options=gpu_options)


def bits_equal(a, b):
    if a.dtype.kind == 'f':
        return np.array_equal(a.view(np.uint32 if a.dtype == np.float32 else np.uint64),
                              b.view(np.uint32 if b.dtype == np.float32 else np.uint64))
    return np.array_equal(a, b)


def max_rel(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    den = np.maximum(np.abs(b), 1e-30)
    return float(np.max(np.abs(a - b) / den)) if a.size else 0.0


def emulating():
    """PMX_EMULATE=1 (or =asan) runs the GPU tests on the CPU emulator of
    the generated kernels (tests/kernel_emu.py)."""
    import os
    return os.environ.get('PMX_EMULATE', '') not in ('', '0')


def gpu_build(pipe):
    """runtime.build; with PMX_PREBUILD=1 (CPU-only cache warming) build
    and skip the GPU part of the test; with PMX_EMULATE build for the CPU
    emulator instead."""
    import os
    import pytest
    if emulating():
        import kernel_emu
        return kernel_emu.build(pipe, asan=os.environ.get('PMX_EMULATE') == 'asan')
    from polymage_amd import runtime
    bp = runtime.build(pipe)
    if os.environ.get('PMX_PREBUILD'):
        pytest.skip("prebuild only")
    return bp


def skip_if_prebuild():
    """For tests that need a real device (not the build cache or the
    kernel emulator)."""
    import os
    import pytest
    if os.environ.get('PMX_PREBUILD'):
        pytest.skip("prebuild only")
    if emulating():
        pytest.skip("needs a device, not the kernel emulator")
