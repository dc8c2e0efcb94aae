"""C++ host stub of a generated pipeline (compiled with g++).

Exports the reference ABI (codegen.py:957-999, utils.py:87-118):

  extern "C" void pipeline_<name>(params sorted by name,
                                  void* inputs sorted by name,
                                  void* outputs sorted by name);

with host pointers and synchronous behaviour (inputs copied to HBM, the
kernels run, outputs copied back; output points no Case covers keep the
caller's bytes, as in the reference), and the device-side entry

  extern "C" int pipeline_<name>_dev(same params, const void* d_inputs...,
                                     void* d_outputs..., void* stream);

which only enqueues kernels on ``stream`` (no allocation after the first
call for a given size, no synchronisation) and returns 0 or a negative
error code (message in pmx_last_error()).  Both validate the parameters
the way the reference's bounds check does (bounds.py:36-118), at run time,
before anything touches the GPU.
"""

from .analysis import classify_access, unconditional_refs
from .constructs import Reference, Image, Reduction, Case, Reduce
from .cprint import Printer, c_type
from .region_ir import bound_c


def _extent(f, d, pr):
    return '((long long)(%s) - (long long)(%s) + 1)' % (pr.expr(f.domain[d].upperBound),
                                                       pr.expr(f.domain[d].lowerBound))


def _size(f, pr):
    if not f.domain:
        return '1LL'
    return ' * '.join(_extent(f, d, pr) for d in range(len(f.domain)))


def _validation_lines(pipe, pr):
    """Run-time restatement of the static bounds check: every affine
    reference of every Case part stays inside its producer's domain."""
    lines = []
    for comp in pipe.comps:
        f = comp.func
        if comp.is_reduction or comp.parts is None:
            continue
        for pi, part in enumerate(comp.parts):
            box = []
            for d in range(len(f.variables)):
                lo = '(long long)(%s)' % pr.expr(f.domain[d].lowerBound)
                hi = '(long long)(%s)' % pr.expr(f.domain[d].upperBound)
                blo, bhi = part.box[d]
                if blo is not None:
                    lo = 'std::max<long long>(%s, %s)' % (lo, bound_c(blo, pr))
                if bhi is not None:
                    hi = 'std::min<long long>(%s, %s)' % (hi, bound_c(bhi, pr))
                box.append((lo, hi))
            nonempty = ' && '.join('(%s) <= (%s)' % b for b in box) or 'true'
            refs = unconditional_refs(part.expr)
            for c in part.pred:
                refs += c.collect(Reference)
            checks = []
            for r in refs:
                p = r.objectRef
                for i, arg in enumerate(r.arguments):
                    acc = classify_access(arg, f.variables)
                    plo = '(long long)(%s)' % pr.expr(p.domain[i].lowerBound)
                    phi = '(long long)(%s)' % pr.expr(p.domain[i].upperBound)
                    if acc.kind == 'const':
                        v = '(long long)(%s)' % pr.expr(arg)
                        checks.append('(%s >= %s && %s <= %s)' % (v, plo, v, phi))
                    elif acc.kind == 'var':
                        lo, hi = box[acc.dim]
                        pl = Printer({acc.var: '(int)(%s)' % lo}, None, pipe.literal_mode)
                        ph = Printer({acc.var: '(int)(%s)' % hi}, None, pipe.literal_mode)
                        a = '(long long)(%s)' % pl.expr(arg)
                        b = '(long long)(%s)' % ph.expr(arg)
                        checks.append('(std::min(%s, %s) >= %s && std::max(%s, %s) <= %s)' % (
                            a, b, plo, a, b, phi))
            if checks:
                lines.append('if ((%s) && !(%s)) { pmx_set_error("%s: reference out of domain in %s"); return -22; }'
                             % (nonempty, ' && '.join(checks), pipe.name, f.name))
    return lines


def emit_host_stub(pipe, kernels, host_consts, co_symbol='pmx_co', co_path='@CO_PATH@'):
    pr = Printer({}, None, pipe.literal_mode)
    name = pipe.name
    params = pipe.sorted_parameters()
    inputs = pipe.sorted_inputs()
    outputs = pipe.sorted_outputs()
    nk = len(kernels)
    L = []
    w = L.append
    w('// host stub for pipeline %s (generated)' % name)
    w('#include <hip/hip_runtime_api.h>')
    w('#include <algorithm>')
    w('#include <cmath>')
    w('#include <math.h>')
    w('#include <cstdio>')
    w('#include <cstdlib>')
    w('#include <cstring>')
    w('#include <climits>')
    w('#include <vector>')
    w('#include "pmx_rt.h"')
    w('extern "C" const unsigned char %s_image[];' % co_symbol)
    w('__asm__(".section .rodata\\n.balign 64\\n.global %s_image\\n%s_image:\\n.incbin \\"%s\\"\\n.byte 0\\n.previous\\n");'
      % (co_symbol, co_symbol, co_path))
    w('static const char* pmx_kernel_names[%d] = {%s};' % (
        max(1, nk), ', '.join('"%s"' % k.name for k in kernels) or '""'))
    w('static void* pmx_mod[64];')
    w('static void* pmx_fn[64][%d];' % max(1, nk))
    w('static int pmx_load(int dev) {')
    w('  if (dev < 0 || dev >= 64) { pmx_set_error("device index"); return -1; }')
    w('  if (pmx_mod[dev]) return 0;')
    w('  void* m = nullptr;')
    w('  int rc = pmx_module_load_data(%s_image, &m); if (rc) return rc;' % co_symbol)
    w('  for (int k = 0; k < %d; ++k) { rc = pmx_module_function(m, pmx_kernel_names[k], &pmx_fn[dev][k]); if (rc) return rc; }' % nk)
    w('  pmx_mod[dev] = m;')
    w('  return 0;')
    w('}')

    # host-side const tables
    for cp in host_consts:
        L.extend(cp.host_code(pr))

    sig = ['%s %s' % (c_type(p.typ), p.name) for p in params]
    dsig = sig + ['const void* %s_dev' % f.name for f in inputs] + \
        ['void* %s_dev' % f.name for f in outputs] + ['void* stream']
    w('extern "C" int pipeline_%s_dev(%s) {' % (name, ', '.join(dsig)))
    w('  int dev = 0; if (hipGetDevice(&dev) != hipSuccess) { pmx_set_error("hipGetDevice"); return -1; }')
    w('  int rc = pmx_load(dev); if (rc) return rc;')
    for l in _validation_lines(pipe, pr):
        w('  ' + l)
    # 32-bit index guard
    all_funcs = []
    for k in kernels:
        for f in k.buffers:
            if f not in all_funcs:
                all_funcs.append(f)
    for f in all_funcs:
        w('  if (%s > (long long)INT_MAX) { pmx_set_error("%s: array too large for 32-bit indexing"); return -75; }'
          % (_size(f, pr), f.name))
    # buffers
    w('  void* pmx_buf_dummy = nullptr; (void)pmx_buf_dummy;')
    bufname = {}
    for f in inputs:
        bufname[f] = '(void*)%s_dev' % f.name
    for f in outputs:
        bufname[f] = '%s_dev' % f.name
    for f in all_funcs:
        if f in bufname:
            continue
        b = 'b_%s' % f.name
        w('  void* %s = nullptr;' % b)
        w('  rc = pmx_arena_get("%s/%s", (size_t)(%s) * sizeof(%s), &%s); if (rc) return rc;'
          % (name, f.name, _size(f, pr), c_type(f.typ), b))
        bufname[f] = b
    for cp in host_consts:
        f = cp.func
        if f in bufname:
            continue
        b = 'b_%s' % f.name
        w('  void* %s = nullptr;' % b)
        w('  rc = %s(%s, stream, &%s); if (rc) return rc;' % (
            cp.upload_fn, ', '.join(p.name for p in cp.params), b))
        bufname[f] = b
    for k in kernels:
        w('  {')
        argv = []
        for p in params:
            w('    %s a_%s = %s;' % (c_type(p.typ), p.name, p.name))
            argv.append('&a_%s' % p.name)
        for f in k.buffers:
            w('    void* a_%s = %s;' % (f.name, bufname[f]))
            argv.append('&a_%s' % f.name)
        w('    void* args[] = {%s};' % (', '.join(argv) or 'nullptr'))
        w('    long long grid = %s;' % k.grid_c)
        w('    if (grid > 0) { rc = pmx_launch(pmx_fn[dev][%d], (unsigned)grid, 1, 1, %d, 1, 1, 0, stream, args); if (rc) return rc; }'
          % (kernels.index(k), k.threads))
        w('  }')
    w('  return 0;')
    w('}')

    # host-pointer entry point (reference ABI)
    hsig = sig + ['void* %s_void_arg' % f.name for f in inputs] + \
        ['void* %s_void_arg' % f.name for f in outputs]
    w('extern "C" int pipeline_%s_host(%s) {' % (name, ', '.join(hsig)))
    w('  int rc = 0;')
    for f in inputs + outputs:
        w('  size_t n_%s = (size_t)(%s) * sizeof(%s);' % (f.name, _size(f, pr), c_type(f.typ)))
        w('  void* d_%s = nullptr; rc = pmx_arena_get("%s/io/%s", n_%s, &d_%s); if (rc) return rc;'
          % (f.name, name, f.name, f.name, f.name))
        w('  rc = pmx_memcpy_h2d(d_%s, %s_void_arg, n_%s, nullptr); if (rc) return rc;' % (
            f.name, f.name, f.name))
    call = [p.name for p in params] + ['d_%s' % f.name for f in inputs + outputs] + ['nullptr']
    w('  rc = pipeline_%s_dev(%s); if (rc) return rc;' % (name, ', '.join(call)))
    for f in outputs:
        w('  rc = pmx_memcpy_d2h(%s_void_arg, d_%s, n_%s, nullptr); if (rc) return rc;' % (
            f.name, f.name, f.name))
    w('  return pmx_stream_sync(nullptr);')
    w('}')
    w('extern "C" void pipeline_%s(%s) {' % (name, ', '.join(hsig)))
    w('  int rc = pipeline_%s_host(%s);' % (name, ', '.join(
        [p.name for p in params] + ['%s_void_arg' % f.name for f in inputs + outputs])))
    w('  if (rc) fprintf(stderr, "pipeline_%s: %%s\\n", pmx_last_error());' % name)
    w('}')
    return '\n'.join(L) + '\n'
