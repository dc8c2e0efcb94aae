"""Writes the golden fixtures in this folder from the reference's own
generated C++ (oracle/_ref, built from the reference tree by
``make -C oracle REF=/root/reference``).

    python tests/golden/make_golden.py

Each fixture is a compressed .npz holding the seeded input, the parameters
and the reference library's output for one small case:

* ``refharris_RxC.npz``: ``refHarris`` of ``ipp_bench/harris/harris_function.cpp``
  (Harris at 32x256 tiles, float literals), float32 image in;
* ``vd_harris.npz``: ``video_demo/harris_naive.cpp`` (uint8 BGR in; the
  tiled ``harris_polymage.cpp`` is checked to give the same bytes);
* ``vd_unsharp.npz``: ``video_demo/unsharp_naive.cpp`` (uint8 in, planar
  float out; ``unsharp_polymage.cpp`` checked equal);
* ``vd_laplacian.npz``: ``video_demo/laplacian_naive.cpp`` (K=4, J=5,
  uint8 RGB), with ``valid_rows``: the rows its bottom padding (stride C
  for 3C, line 183) cannot reach.

Only data is stored; the reference's sources are not copied."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(os.path.dirname(os.path.dirname(HERE)), 'oracle', '_ref')


def lib(name):
    return ctypes.CDLL(os.path.join(REF, name))


def vp(a):
    return ctypes.c_void_p(a.ctypes.data)


def refharris(R, C):
    img = np.random.default_rng(R * 7 + C).random((R + 2, C + 2)).astype(np.float32)
    out = np.zeros_like(img)
    lib('libref_harris.so').ref_harris(ctypes.c_int(C), ctypes.c_int(R), vp(img), vp(out))
    np.savez_compressed(os.path.join(HERE, 'refharris_%dx%d.npz' % (R, C)),
                        R=R, C=C, img=img, harris=out)


def vd_harris(R=62, C=90):
    img = np.random.default_rng(11).integers(0, 256, (R + 2, C + 2, 3)).astype(np.uint8)
    outs = []
    for so, fn in (('libvd_harris_naive.so', 'pipeline_harris_naive'),
                   ('libvd_harris_polymage.so', 'pipeline_harris')):
        o = np.zeros((R + 2, C + 2), np.float32)
        getattr(lib(so), fn)(ctypes.c_int(C), ctypes.c_int(R), vp(img), vp(o))
        outs.append(o)
    assert np.array_equal(outs[0][2:R, 2:C].view(np.uint32), outs[1][2:R, 2:C].view(np.uint32))
    np.savez_compressed(os.path.join(HERE, 'vd_harris.npz'), R=R, C=C, img=img, harris=outs[0])


def vd_unsharp(R=40, C=72):
    img = np.random.default_rng(12).integers(0, 256, (R + 4, C + 4, 3)).astype(np.uint8)
    outs = []
    for so, fn in (('libvd_unsharp_naive.so', 'pipeline_mask_naive'),
                   ('libvd_unsharp_polymage.so', 'pipeline_mask')):
        o = np.zeros((R, C, 3), np.float32)
        getattr(lib(so), fn)(ctypes.c_int(C), ctypes.c_int(R), ctypes.c_float(0.001),
                             ctypes.c_float(3.0), vp(img), vp(o))
        outs.append(o)
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))
    np.savez_compressed(os.path.join(HERE, 'vd_unsharp.npz'), R=R, C=C, threshold=0.001,
                        weight=3.0, img=img, mask=outs[0])


def vd_laplacian(Rin=128, Cin=96):
    R, C = Rin + 92, Cin + 92
    img = np.random.default_rng(4).integers(0, 256, (Rin, Cin, 3)).astype(np.uint8)
    o = np.zeros((Rin, Cin, 3), np.uint8)
    lib('libvd_laplacian_naive.so').pipeline_laplacian_naive(
        ctypes.c_int(C), ctypes.c_int(R), ctypes.c_float(1.0 / 3), ctypes.c_float(1.0), vp(img), vp(o))
    np.savez_compressed(os.path.join(HERE, 'vd_laplacian.npz'), R=R, C=C, alpha=1.0 / 3, beta=1.0,
                        img=img, laplacian=o, valid_rows=Rin - 40)


if __name__ == '__main__':
    for R, C in ((64, 96), (37, 133), (2, 7)):
        refharris(R, C)
    vd_harris()
    vd_unsharp()
    vd_laplacian()
    for f in sorted(os.listdir(HERE)):
        if f.endswith('.npz'):
            print(f, os.path.getsize(os.path.join(HERE, f)))
