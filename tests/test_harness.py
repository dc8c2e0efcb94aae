"""The reference-side entry points: the apps' harness modes
(new | existing | ready | tune), ``cpp_compiler.c_compile`` and
``loader.load_lib`` on the generated source, ``tuner.generate`` /
``tuner.execute``, and the void ``pipeline_<name>`` they call.

The CPU tests check that the modes write, compile and load what the
reference's do (compiling needs no GPU); the GPU tests run the loaded
functions on host arrays and compare with the oracle."""
import ctypes
import os

import numpy as np
import pytest

from oracle import cemit
from polymage_amd.apps import harness

from helpers import bits_equal, skip_if_prebuild


def _run_mode(tmp_path, mode, app='harris', rows=70, cols=94, extra=()):
    argv = [app, '--mode', mode, '--rows', str(rows), '--cols', str(cols), '--dir', str(tmp_path),
            '--tiles', '16x32'] + list(extra)
    app_args = harness.parse_args(argv)
    app_data = harness.init_all(app_args)
    harness.create_lib(harness.build_app, app, app_data)
    return app_data


def test_modes_write_compile_and_load(tmp_path):
    """new writes <app>.cpp and <app>.so and loads pipeline_<app>;
    existing recompiles the file on disk without a pipeline; ready loads
    the library as it is."""
    d = _run_mode(tmp_path, 'new')
    src = tmp_path / 'harris.cpp'
    so = tmp_path / 'harris.so'
    assert src.exists() and so.exists()
    assert 'pipeline_harris' in d
    text = src.read_text()
    assert '__global__' in text and 'extern "C" void pipeline_harris(' in text
    so.unlink()
    d2 = _run_mode(tmp_path, 'existing')
    assert so.exists()
    assert d2['pipeline_harris'] is not None
    d3 = _run_mode(tmp_path, 'ready')
    assert d3['pipeline_harris'] is not None


def test_existing_mode_compiles_the_file_not_a_rebuilt_pipeline(tmp_path):
    """A source edited on disk is what existing compiles: a marker added
    to the host stub shows up as an exported symbol."""
    _run_mode(tmp_path, 'new')
    src = tmp_path / 'harris.cpp'
    src.write_text(src.read_text() + '\nextern "C" int pmx_edited_on_disk(void) { return 7; }\n')
    _run_mode(tmp_path, 'existing')
    # (the process already holds harris.so from the new mode; the reference
    # runs one mode per process, so inspect the rebuilt file instead)
    import subprocess
    syms = subprocess.check_output(['nm', '-D', '--defined-only', str(tmp_path / 'harris.so')]).decode()
    assert 'pmx_edited_on_disk' in syms


def test_c_compile_rejects_other_sources(tmp_path):
    from polymage_amd.compat.cpp_compiler import c_compile
    bad = tmp_path / 'x.cpp'
    bad.write_text('int main() { return 0; }\n')
    with pytest.raises(ValueError):
        c_compile(str(bad), str(tmp_path / 'x.so'))


def test_ordered_args_follow_the_reference_order():
    """codegen.py:957-999: parameters by name (uppercase first), inputs by
    name, then the output."""
    vals = {'R': 5, 'C': 7, 'threshold': 0.5, 'img': np.zeros(3, np.float32),
            'alpha_img': np.zeros(2, np.float32)}
    out = np.zeros(1, np.float32)
    args = harness.ordered_args(vals, 'o', out)
    assert [type(a) for a in args[:3]] == [ctypes.c_int, ctypes.c_int, ctypes.c_float]
    assert [a.value for a in args[:2]] == [7, 5]
    assert args[3].value == vals['alpha_img'].ctypes.data
    assert args[4].value == vals['img'].ctypes.data
    assert args[5].value == out.ctypes.data


def _oracle(app_data):
    w = app_data['workload']
    o = np.zeros(w.out_shape, w.out_func.typ.np_dtype())
    cemit.build([w.out_func], 'harness_%s' % app_data['app'])(
        dict(app_data['img_data']['values'], **{w.out_func.name: o}))
    return o


@pytest.mark.gpu
@pytest.mark.parametrize("app,rows,cols", [('harris', 70, 94), ('unsharp', 52, 77), ('bilateral', 64, 48)])
def test_harness_new_existing_ready_run_bitexact(gpu, tmp_path, app, rows, cols):
    skip_if_prebuild()
    print("case harness", app, rows, cols, flush=True)
    for mode in ('new', 'existing', 'ready'):
        d = _run_mode(tmp_path, mode, app, rows, cols, ['--runs', '2', '--timer'])
        d['img_data']['OUT'][...] = 0
        harness.exec_pipe(d)
        assert bits_equal(d['img_data']['OUT'], _oracle(d)), mode


@pytest.mark.gpu
def test_tuner_generate_and_execute_bitexact(gpu, tmp_path):
    """tuner.generate with 2 tile configs x 1 group size, then
    tuner.execute calling the void pipeline_harris of each library on host
    arrays; every library's output equals the oracle's."""
    skip_if_prebuild()
    from polymage_amd import tuner
    from polymage_amd.apps.harris import harris_pipe
    from polymage_amd.constructs import Condition
    print("case tuner generate/execute", flush=True)
    R, C = 60, 90
    pd = {}
    out = harris_pipe(pd)
    gen = {'_tuner_app_name': 'harris', '_tuner_pipe_name': 'harris', '_tuner_live_outs': [out],
           '_tuner_param_constraints': [Condition(pd['R'], '==', R), Condition(pd['C'], '==', C)],
           '_tuner_param_estimates': [(pd['R'], R), (pd['C'], C)],
           '_tuner_tile_size_configs': [[16, 32], [32, 64]], '_tuner_group_size_configs': [11],
           '_tuner_dst_path': str(tmp_path), '_tuner_debug_flag': False}
    src, n, pipe = tuner.generate(gen)
    assert n == 2
    assert os.path.exists(os.path.join(src, 'configurations.txt'))
    img = np.random.default_rng(4).random((R + 2, C + 2)).astype(np.float32)
    ref = np.zeros((R + 2, C + 2), np.float32)
    cemit.build([out], 'harris_oracle_double')({'R': R, 'C': C, 'img': img, 'harris': ref})
    seen = []

    def executor(lib, fn, params, args, arg, app_data):
        app_data['out'][...] = 0
        fn(*args)
        seen.append(bits_equal(app_data['out'], ref))

    outbuf = np.zeros((R + 2, C + 2), np.float32)
    res = tuner.execute({'_tuner_app_name': 'harris', '_tuner_pipe': pipe, '_tuner_src_path': src,
                         '_tuner_configs_count': n, '_tuner_nruns': 2,
                         '_tuner_pipe_arg_data': {'R': R, 'C': C, 'img': img, 'harris': outbuf},
                         '_tuner_custom_executor': executor, '_tuner_app_data': {'out': outbuf},
                         '_tuner_debug_flag': False})
    assert res['best_config'] in (1, 2) and len(res['times']) == 2
    assert seen == [True] * 4
    # the plain path: execute calls the void entry itself
    outbuf[...] = 0
    res = tuner.execute({'_tuner_app_name': 'harris', '_tuner_pipe': pipe, '_tuner_src_path': src,
                         '_tuner_configs_count': n, '_tuner_nruns': 1,
                         '_tuner_pipe_arg_data': {'R': R, 'C': C, 'img': img, 'harris': outbuf},
                         '_tuner_debug_flag': False})
    assert len(res['times']) == 2
    assert bits_equal(outbuf, ref)
