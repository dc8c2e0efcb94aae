"""The C ABI: libpmx_rt.so exports every function include/pmx_rt.h declares,
and the generated libraries of the benchmark pipelines export every entry
point include/pmx_pipeline.h declares (loading only; no device calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_DECL = re.compile(r'^[A-Za-z_][\w \*]*?\b((?:pmx|pool|pipeline)_\w+)\s*\(', re.M)


def declared(header):
    with open(os.path.join(ROOT, 'include', header)) as f:
        text = f.read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(_DECL.findall(text)))


def test_runtime_exports_every_declared_symbol():
    from polymage_amd import runtime
    lib = runtime.rt()
    names = declared('pmx_rt.h')
    assert 'pmx_launch' in names and 'pmx_mlp_adam' in names and 'pool_init' in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_pipeline_libraries_export_every_declared_entry_point():
    from polymage_amd import runtime
    from polymage_amd.apps import workloads
    names = declared('pmx_pipeline.h')
    libs = {}
    for wl in ('harris', 'unsharp', 'bilateral', 'laplacian', 'campipe'):
        w = getattr(workloads, wl)(with_data=False)
        libs[wl] = runtime.build(w.pipe).lib
    missing = []
    for n in names:
        app = n[len('pipeline_'):].split('_')[0]
        if not hasattr(libs[app], n):
            missing.append(n)
    assert not missing, missing
    # and every generated library has the full generic set
    for wl, lib in libs.items():
        for suffix in ('', '_host', '_dev', '_dev_noalloc', '_status', '_batch_create', '_batch_run',
                       '_batch_run_graph', '_batch_run_frames', '_batch_destroy'):
            assert hasattr(lib, 'pipeline_%s%s' % (wl, suffix)), (wl, suffix)
