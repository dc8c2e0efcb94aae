"""``cpp_compiler.c_compile`` (apps/python/cpp_compiler.py:4-43): the
source is the generated pipeline (kernels + host stub); the library is
built with hiprtc + g++ instead of g++ -fopenmp."""
import os
import shutil


def c_compile(in_file, out_file, app_data=None):
    """``in_file`` is the path written by the app's codegen step; the
    pipeline object itself is taken from app_data['pipe'] when present."""
    from polymage_amd import runtime
    pipe = (app_data or {}).get('pipe')
    if pipe is None:
        raise ValueError("c_compile needs app_data['pipe'] (the Pipeline)")
    bp = runtime.build(pipe)
    if os.path.lexists(out_file):
        os.remove(out_file)
    shutil.copyfile(bp.so_path, out_file)
