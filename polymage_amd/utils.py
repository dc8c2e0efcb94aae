"""ctypes marshalling and argument ordering (sandbox/utils.py:35-118).

``get_ordered_cfunc_params`` orders the arguments of pipeline_<name>:
parameters by name, then input images by name, then outputs by name
(codegen.py:957-999).  ``map_cfunc_args`` turns an argument dict into the
ctypes list, host arrays passed as ``c_void_p(ndarray.ctypes.data)``.
"""

import ctypes

from .liveness import level_order  # noqa: F401  (re-exported, utils.py:120)

_CTYPE = {
    'float': ctypes.c_float, 'double': ctypes.c_double,
    'int32': ctypes.c_int, 'uint32': ctypes.c_uint,
    'int16': ctypes.c_short, 'uint16': ctypes.c_ushort,
    'int8': ctypes.c_byte, 'uint8': ctypes.c_ubyte,
    'int64': ctypes.c_longlong, 'uint64': ctypes.c_ulonglong,
}


def convert_to_ctype(inp_type, inp_value):
    return _CTYPE[inp_type](inp_value)


def get_ordered_cfunc_params(pipe_object):
    params = sorted(pipe_object.get_parameters(), key=lambda x: x.name)
    inputs = sorted(pipe_object.inputs, key=lambda x: x.name)
    outputs = sorted(pipe_object.outputs, key=lambda x: x.name)
    return params, inputs, outputs


def map_cfunc_args(func_params, arg_data):
    params, inputs, outputs = func_params
    args = []
    for p in params:
        args.append(convert_to_ctype(p.typ.c_type_name(), arg_data[p.name]))
    for f in inputs:
        args.append(ctypes.c_void_p(arg_data[f.name].ctypes.data))
    for f in outputs:
        args.append(ctypes.c_void_p(arg_data[f.name].ctypes.data))
    return args


def image_clamp(arr, ghost=(2, 2)):
    """Pad an image by edge replication (utils.py:147-188 helper)."""
    import numpy as np
    pad = [(g, g) for g in ghost] + [(0, 0)] * (arr.ndim - len(ghost))
    return np.pad(arr, pad, mode='edge')
