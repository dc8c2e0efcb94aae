"""``cpp_compiler.c_compile`` (apps/python/cpp_compiler.py:4-43).

The reference compiles the generated C++ file with ``g++ -fopenmp`` into a
shared library.  Here the file holds what ``str(Pipeline.generate_code())``
prints: the HIP kernels, then the C++ host stub.  The kernels are compiled
with hiprtc for gfx950 and the stub with g++ (the code object embedded),
both cached by source hash, and the library is copied to ``out_file``.
Only the file is read, so a source written by an earlier run (the apps'
``existing`` mode) compiles without rebuilding the pipeline.
"""
import os
import re
import shutil


def c_compile(in_file, out_file, app_data=None):
    from polymage_amd import runtime
    with open(in_file) as f:
        text = f.read()
    hip_src, host_src = runtime.split_source(text)
    m = re.search(r'^// host stub for pipeline (\w+) \(generated\)', host_src, re.M)
    if not m:
        raise ValueError('%s: no pipeline name in the host stub' % in_file)
    so = runtime.build_sources(m.group(1), hip_src, host_src)
    if os.path.lexists(out_file):
        os.remove(out_file)
    tmp = out_file + '.tmp%d' % os.getpid()
    shutil.copyfile(so, tmp)
    os.replace(tmp, out_file)
    if app_data is not None:
        app_data['so_path'] = out_file
