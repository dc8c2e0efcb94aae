"""``buildPipeline``: the compiler entry point (sandbox/compiler.py:33-80).

Same signature and defaults as the reference: ``group_size`` 5,
``tile_sizes`` [16, 16, 16], ``size_threshold`` 200*200 and the option
strings ``optimize_storage`` (implies ``flatten_scratchpad``),
``early_free``, ``pool_alloc``, ``multipar``.  Two keyword arguments are
added for the GPU backend and default to the reference's behaviour:

* ``literal_mode``: 'double' prints float literals as C doubles, as the
  current reference code generator does (targetc.py:80-86); 'float12'
  prints them as ``%.12g`` floats with an ``f`` suffix, as the reference's
  checked-in generated files do (e.g. ipp_bench/harris/harris_function.cpp);
* ``gpu_options``: schedule knobs of the HIP emitter (threads per
  workgroup, vector width, XCD-aware tile order, ...).
"""

from .pipe import Pipeline


def buildPipeline(outputs, param_estimates=[], param_constraints=[],
                  grouping=[], group_size=None, inline_directives=[],
                  tile_sizes=[], size_threshold=None, pipe_name=None,
                  options=[], literal_mode='double', gpu_options=None):
    if group_size is None:
        group_size = 5
    if not tile_sizes:
        tile_sizes = [16, 16, 16]
    if size_threshold is None:
        size_threshold = 200 * 200
    options = list(options)
    if 'optimize_storage' in options and 'flatten_scratchpad' not in options:
        options.append('flatten_scratchpad')
    return Pipeline(outputs, param_estimates, param_constraints, grouping,
                    group_size, inline_directives, tile_sizes,
                    size_threshold, options, pipe_name,
                    _literal_mode=literal_mode, _gpu_options=gpu_options)
