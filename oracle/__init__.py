"""CPU oracle for the PolyMage stencil-group path — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import anything from here, and only as the checker.  The
product (``polymage_amd``) never imports, links or executes this package; it
fails loudly when its HIP build is missing instead of falling back to it.

Contents
  cemit.py   naive C++/OpenMP emitter restating the reference code
             generator's per-point semantics (sandbox/codegen.py:268-372,
             553-715, 920-1067; sandbox/targetc.py:80-86, 174-217, 403-406,
             486-516).  Every stage is evaluated over its whole domain in
             level order into full arrays, exactly like the reference's
             checked-in ``video_demo/*_naive.cpp`` files; values equal those
             of the tiled reference code because tiling never changes a
             point's expression.
  eval64.py  float64 numpy evaluator of DSL programs (independent of both
             emitters) used for the error statement per pipeline and to
             pin integer semantics.
  Makefile   builds ``oracle/_ref/`` from the reference's own checked-in
             generated C++ (ipp_bench/harris/harris_function.cpp and
             video_demo/*.cpp) where they lie under /root/reference.

How the oracle is pinned (each check is a test under tests/):
  * cemit vs oracle/_ref: bit-exact on seeded inputs for refHarris
    (float12 literal mode) and for the video_demo harris / unsharp naive
    files re-expressed in the DSL;
  * cemit vs eval64: exact for integer pipelines, within float tolerance
    for float pipelines, plus closed forms (Sobel weights, binomial blur of
    constant/linear images, Harris response of constant images = 0);
  * brute-force tiny cases in the tests.
"""
