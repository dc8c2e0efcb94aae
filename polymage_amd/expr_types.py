"""Scalar element types of the PolyMage DSL and their promotion rule.

Mirrors the type classes and ``result_type`` of the reference
(sandbox/expr_types.py:25-104).  Each type is a class used as a tag, exactly
as the reference's application files use them (``Image(Float, ...)``).
On top of the reference's ``c_type_name`` we attach the C spelling, the byte
width and the numpy dtype, which the emitters and the host runtime need.
"""

import numpy as _np


class _ScalarType(object):
    # overridden per subclass
    _cname = None      # C spelling used in emitted code
    _tname = None      # the reference's c_type_name() string
    _nbytes = 0
    _np = None
    _is_float = False
    _is_signed = True

    @classmethod
    def c_type_name(cls):
        return cls._tname

    @classmethod
    def c_name(cls):
        return cls._cname

    @classmethod
    def nbytes(cls):
        return cls._nbytes

    @classmethod
    def np_dtype(cls):
        return cls._np

    @classmethod
    def is_float(cls):
        return cls._is_float

    @classmethod
    def is_signed(cls):
        return cls._is_signed


class Void(_ScalarType):
    _cname, _tname = 'void', 'void'


class Float(_ScalarType):
    _cname, _tname, _nbytes, _np, _is_float = 'float', 'float', 4, _np.float32, True


class Double(_ScalarType):
    _cname, _tname, _nbytes, _np, _is_float = 'double', 'double', 8, _np.float64, True


class Int(_ScalarType):
    _cname, _tname, _nbytes, _np = 'int', 'int32', 4, _np.int32


class UInt(_ScalarType):
    _cname, _tname, _nbytes, _np, _is_signed = 'unsigned int', 'uint32', 4, _np.uint32, False


class Short(_ScalarType):
    _cname, _tname, _nbytes, _np = 'short', 'int16', 2, _np.int16


class UShort(_ScalarType):
    _cname, _tname, _nbytes, _np, _is_signed = 'unsigned short', 'uint16', 2, _np.uint16, False


class Char(_ScalarType):
    _cname, _tname, _nbytes, _np = 'char', 'int8', 1, _np.int8


class UChar(_ScalarType):
    _cname, _tname, _nbytes, _np, _is_signed = 'unsigned char', 'uint8', 1, _np.uint8, False


class Long(_ScalarType):
    # The reference's TypeNameMap has no int64 spelling (targetc.py:175-180,
    # a duplicated dict key); we give it the real one.
    _cname, _tname, _nbytes, _np = 'long long', 'int64', 8, _np.int64


class ULong(_ScalarType):
    _cname, _tname, _nbytes, _np, _is_signed = 'unsigned long long', 'uint64', 8, _np.uint64, False


class Rational(object):
    """Tag for exact fractions produced by affine analysis."""
    pass


# Promotion order of the reference (expr_types.py:83-104).
_PROMOTION = [Double, Float, ULong, Long, UInt, Int, UShort, Short, UChar]


def result_type(a, b):
    for t in _PROMOTION:
        if a is t or b is t:
            return t
    if a is Char and b is Char:
        return Char
    raise TypeError((a, b))


ALL_TYPES = [Float, Double, Int, UInt, Short, UShort, Char, UChar, Long, ULong]
