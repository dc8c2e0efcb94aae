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
    w('#include <map>')
    w('#include <set>')
    w('#include "pmx_rt.h"')
    w('extern "C" const unsigned char %s_image[];' % co_symbol)
    w('__asm__(".section .rodata\\n.balign 64\\n.global %s_image\\n%s_image:\\n.incbin \\"%s\\"\\n.byte 0\\n.previous\\n");'
      % (co_symbol, co_symbol, co_path))
    w('static const char* pmx_kernel_names[%d] = {%s};' % (
        max(1, nk), ', '.join('"%s"' % k.name for k in kernels) or '""'))
    w('static void* pmx_mod[64];')
    w('static void* pmx_fn[64][%d];' % max(1, nk))
    w('static void* pmx_fn_batch[64][%d];' % max(1, nk))
    w('static void* pmx_zero_fn[64];')
    w('static int pmx_load(int dev) {')
    w('  if (dev < 0 || dev >= 64) { pmx_set_error("device index"); return -1; }')
    w('  if (pmx_mod[dev]) return 0;')
    w('  void* m = nullptr;')
    w('  int rc = pmx_module_load_data(%s_image, &m); if (rc) return rc;' % co_symbol)
    w('  for (int k = 0; k < %d; ++k) { rc = pmx_module_function(m, pmx_kernel_names[k], &pmx_fn[dev][k]); if (rc) return rc; }' % nk)
    w('  for (int k = 0; k < %d; ++k) { char nm[256]; snprintf(nm, sizeof nm, "%%s_batch", pmx_kernel_names[k]); '
      'rc = pmx_module_function(m, nm, &pmx_fn_batch[dev][k]); if (rc) return rc; }' % nk)
    w('  rc = pmx_module_function(m, "pmx_zero_u64", &pmx_zero_fn[dev]); if (rc) return rc;')
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
    for cp in host_consts:
        f = cp.func
        b = 'b_%s' % f.name
        w('  void* %s = nullptr;' % b)
        w('  rc = %s(%s, stream, &%s); if (rc) return rc;' % (
            cp.upload_fn, ', '.join(p.name for p in cp.params), b))
        bufname[f] = b
    for f in all_funcs:
        if f in bufname:
            continue
        b = 'b_%s' % f.name
        w('  void* %s = nullptr;' % b)
        w('  rc = pmx_arena_get("%s/%s", (size_t)(%s) * sizeof(%s), &%s); if (rc) return rc;'
          % (name, f.name, _size(f, pr), c_type(f.typ), b))
        bufname[f] = b
    w('  void* pmx_err = nullptr; rc = pmx_arena_get_zeroed("%s/err", 64, &pmx_err); if (rc) return rc;' % name)
    _launch_lines(w, kernels, params, bufname)
    w('  return 0;')
    w('}')

    # the same launches with every buffer the library provides (status
    # word, const tables, intermediates) carved from the runtime's static
    # device scratch: no device allocation after the first call on a device
    w('extern "C" int pipeline_%s_dev_noalloc(%s) {' % (name, ', '.join(dsig)))
    w('  int dev = 0; if (hipGetDevice(&dev) != hipSuccess) { pmx_set_error("hipGetDevice"); return -1; }')
    w('  int rc = pmx_load(dev); if (rc) return rc;')
    for l in _validation_lines(pipe, pr):
        w('  ' + l)
    for f in all_funcs:
        w('  if (%s > (long long)INT_MAX) { pmx_set_error("%s: array too large for 32-bit indexing"); return -75; }'
          % (_size(f, pr), f.name))
    w('  void* pmx_scr = nullptr; size_t pmx_cap = 0; rc = pmx_scratch(&pmx_scr, &pmx_cap); if (rc) return rc;')
    w('  size_t pmx_off = 256;  // [0, 64): the status word of the allocation-free calls')
    nb = {}
    for f in inputs:
        nb[f] = '(void*)%s_dev' % f.name
    for f in outputs:
        nb[f] = '%s_dev' % f.name
    offs = []
    for cp in host_consts:
        f = cp.func
        w('  const %s* h_%s = nullptr; size_t n_%s = 0;' % (c_type(f.typ), f.name, f.name))
        w('  rc = %s_host(%s%s&h_%s, &n_%s, nullptr); if (rc) return rc;' % (
            cp.upload_fn, ', '.join(p.name for p in cp.params), ', ' if cp.params else '', f.name, f.name))
        w('  const size_t o_%s = pmx_off; pmx_off += (n_%s + 255) & ~(size_t)255;' % (f.name, f.name))
        offs.append(f)
    for f in all_funcs:
        if f in nb or f in offs:
            continue
        w('  const size_t o_%s = pmx_off; pmx_off += ((size_t)(%s) * sizeof(%s) + 255) & ~(size_t)255;'
          % (f.name, _size(f, pr), c_type(f.typ)))
        offs.append(f)
    w('  if (pmx_off > pmx_cap) { char m[256]; snprintf(m, sizeof m, "%s: these sizes need %%zu bytes of '
      'static scratch, which holds %%zu; use pipeline_%s_dev", pmx_off, pmx_cap); pmx_set_error(m); return -12; }'
      % (name, name))
    for f in offs:
        nb[f] = 'b_%s' % f.name
        w('  void* b_%s = (void*)((unsigned char*)pmx_scr + o_%s);' % (f.name, f.name))
    for cp in host_consts:
        f = cp.func
        w('  rc = pmx_memcpy_h2d(b_%s, h_%s, n_%s, stream); if (rc) return rc;' % (f.name, f.name, f.name))
    w('  void* pmx_err = pmx_scr;')
    _launch_lines(w, kernels, params, nb)
    w('  return 0;')
    w('}')

    # host-pointer entry point (reference ABI)
    hsig = sig + ['void* %s_void_arg' % f.name for f in inputs] + \
        ['void* %s_void_arg' % f.name for f in outputs]
    w('extern "C" int pipeline_%s_status(void) {' % name)
    w('  void* e = nullptr; int rc = pmx_arena_get_zeroed("%s/err", 64, &e); if (rc) return rc;' % name)
    w('  int v = 0; rc = pmx_memcpy_d2h(&v, e, sizeof v, nullptr); if (rc) return rc;')
    w('  rc = pmx_stream_sync(nullptr); if (rc) return rc;')
    w('  if (v) { int z = 0; pmx_memcpy_h2d(e, &z, sizeof z, nullptr); pmx_stream_sync(nullptr); }')
    w('  // the allocation-free calls keep their status word in the static scratch')
    w('  void* s = nullptr; pmx_scratch_peek(&s);')
    w('  if (s) { int u = 0; rc = pmx_memcpy_d2h(&u, s, sizeof u, nullptr); if (rc) return rc;')
    w('    rc = pmx_stream_sync(nullptr); if (rc) return rc;')
    w('    if (u) { int z = 0; pmx_memcpy_h2d(s, &z, sizeof z, nullptr); pmx_stream_sync(nullptr); v |= u; } }')
    w('  if (v) {')
    w('    pmx_set_error("%s: a tile exceeded its planned LDS box; rebuild with estimates closer to these sizes"); return -34; }' % name)
    w('  return 0;')
    w('}')
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
    w('  rc = pmx_stream_sync(nullptr); if (rc) return rc;')
    w('  return pipeline_%s_status();' % name)
    w('}')
    _emit_batch_api(pipe, kernels, host_consts, L, pr, bufname_fn=None)
    w('extern "C" void pipeline_%s(%s) {' % (name, ', '.join(hsig)))
    w('  int rc = pipeline_%s_host(%s);' % (name, ', '.join(
        [p.name for p in params] + ['%s_void_arg' % f.name for f in inputs + outputs])))
    w('  if (rc) fprintf(stderr, "pipeline_%s: %%s\\n", pmx_last_error());' % name)
    w('}')
    return '\n'.join(L) + '\n'


def _launch_lines(w, kernels, params, bufname):
    """One launch per kernel, in order, on ``stream``; ``pmx_err``, ``dev``
    and ``rc`` are in scope."""
    for k in kernels:
        w('  {')
        argv = ['&pmx_err']
        for f in k.buffers:
            w('    void* a_%s = %s;' % (f.name, bufname[f]))
            argv.append('&a_%s' % f.name)
        for p in params:
            w('    %s a_%s = %s;' % (c_type(p.typ), p.name, p.name))
            argv.append('&a_%s' % p.name)
        w('    void* args[] = {%s};' % (', '.join(argv) or 'nullptr'))
        w('    long long grid = %s;' % k.grid_c)
        w('    if (grid > 0) { rc = pmx_launch(pmx_fn[dev][%d], (unsigned)grid, 1, 1, %d, 1, 1, 0, stream, args); if (rc) return rc; }'
          % (kernels.index(k), k.threads))
        w('  }')


def frame_struct_fields(kernel, params):
    """Field declarations of <kernel>_frame (same text as the device side,
    see hip_emit.kernel_args) and the struct size on LP64 / amdgcn."""
    from .hip_emit import kernel_args
    decls, names, fields = kernel_args(params, kernel.buffers, kernel.writes)
    size = 8 * len(kernel.buffers) + 16
    for p in params:
        size += p.typ.nbytes()
    size = (size + 7) // 8 * 8
    return fields, size


def _emit_batch_api(pipe, kernels, host_consts, L, pr, bufname_fn=None):
    """Batched execution of many frames of different sizes.

    pipeline_<name>_batch_create(n, per-frame parameter arrays...,
        per-frame input pointer arrays..., per-frame output pointer arrays...,
        &handle): plans every kernel's work items over all frames, allocates
        per-frame intermediates, uploads the frame records.
    pipeline_<name>_batch_run(handle, stream): one persistent work-queue
        launch per kernel for all frames.
    pipeline_<name>_batch_run_graph(handle, stream): the same sequence
        captured once in a hipGraph and replayed.
    pipeline_<name>_batch_run_frames(handle, stream): the per-frame plain
        launches (pipeline_<name>_dev once per frame) for comparison.
    pipeline_<name>_batch_destroy(handle)."""
    w = L.append
    name = pipe.name
    params = pipe.sorted_parameters()
    inputs = pipe.sorted_inputs()
    outputs = pipe.sorted_outputs()
    nk = len(kernels)
    for k in kernels:
        fields, size = frame_struct_fields(k, params)
        w('struct %s_frame { %s };' % (k.name, ' '.join(fields)))
        w('static_assert(sizeof(%s_frame) == %d, "frame record layout");' % (k.name, size))
    w('struct pmx_batch_%s {' % name)
    w('  int n = 0; int dev = 0;')
    w('  void* frames[%d] = {};' % max(1, nk))
    w('  long long total[%d] = {};' % max(1, nk))
    w('  void* counters = nullptr; void* graph = nullptr;')
    w('  std::vector<void*> owned;  // per-frame intermediates of this handle')
    w('  long long grid[%d] = {};' % max(1, nk))
    for p in params:
        w('  std::vector<%s> p_%s;' % (c_type(p.typ), p.name))
    for f in inputs + outputs:
        w('  std::vector<void*> io_%s;' % f.name)
    w('};')
    # one release path for destroy and for every failure inside create
    w('static int pmx_batch_%s_free(pmx_batch_%s* B) {' % (name, name))
    w('  if (!B) return 0;')
    w('  if (B->graph) pmx_graph_destroy(B->graph);')
    w('  for (int k = 0; k < %d; ++k) if (B->frames[k]) pmx_free(B->frames[k]);' % max(1, nk))
    w('  if (B->counters) pmx_free(B->counters);')
    w('  for (void* p : B->owned) pmx_free(p);')
    w('  delete B;')
    w('  return 0;')
    w('}')
    sig = ['int n'] + ['const %s* %s_arr' % (c_type(p.typ), p.name) for p in params] + \
        ['void* const* %s_arr' % f.name for f in inputs + outputs] + ['void** handle']
    w('extern "C" int pipeline_%s_batch_create(%s) {' % (name, ', '.join(sig)))
    w('  if (n <= 0) { pmx_set_error("batch: no frames"); return -22; }')
    w('  pmx_batch_%s* B = new pmx_batch_%s(); B->n = n;' % (name, name))
    w('  if (hipGetDevice(&B->dev) != hipSuccess) { pmx_batch_%s_free(B); return -1; }' % name)
    w('  int rc = pmx_load(B->dev); if (rc) { pmx_batch_%s_free(B); return rc; }' % name)
    for p in params:
        w('  B->p_%s.assign(%s_arr, %s_arr + n);' % (p.name, p.name, p.name))
    for f in inputs + outputs:
        w('  B->io_%s.assign(%s_arr, %s_arr + n);' % (f.name, f.name, f.name))
    for k in kernels:
        w('  std::vector<%s_frame> rec_%s((size_t)n);' % (k.name, k.name))
    w('  for (int i = 0; i < n; ++i) {')
    for p in params:
        w('    const %s %s = %s_arr[i];' % (c_type(p.typ), p.name, p.name))
    for l in _validation_lines(pipe, pr):
        w('    ' + l.replace('return -22;', 'pmx_batch_%s_free(B); return -22;' % name))
    bufname = {}
    for f in inputs + outputs:
        bufname[f] = '%s_arr[i]' % f.name
    for cp in host_consts:
        f = cp.func
        b = 'b_%s' % f.name
        w('    void* %s = nullptr;' % b)
        w('    rc = %s(%s, nullptr, &%s); if (rc) { pmx_batch_%s_free(B); return rc; }' % (
            cp.upload_fn, ', '.join(p.name for p in cp.params), b, name))
        bufname[f] = b
    for k in kernels:
        for f in k.buffers:
            if f in bufname:
                continue
            b = 'b_%s' % f.name
            w('    void* %s = nullptr;' % b)
            w('    rc = pmx_malloc(&%s, (size_t)(%s) * sizeof(%s)); if (rc) { pmx_batch_%s_free(B); return rc; }'
              % (b, _size(f, pr), c_type(f.typ), name))
            w('    B->owned.push_back(%s);' % b)
            bufname[f] = b
    w('    void* pmx_err = nullptr; rc = pmx_arena_get_zeroed("%s/err", 64, &pmx_err); if (rc) { pmx_batch_%s_free(B); return rc; }' % (name, name))
    for k in kernels:
        r = 'rec_%s[i]' % k.name
        w('    %s.pmx_err = (int*)pmx_err;' % r)
        for f in k.buffers:
            ct = c_type(f.typ)
            q = '' if f in k.writes else 'const '
            w('    %s.G_%s = (%s%s*)%s;' % (r, f.name, q, ct, bufname[f]))
        for p in params:
            w('    %s.%s = %s;' % (r, p.name, p.name))
        w('    %s.pmx_tile0 = (long long)(%s);' % (r, getattr(k, 'tiles_c', k.grid_c)))
    w('  }')
    # prefix sums -> first work item of every frame
    for ki, k in enumerate(kernels):
        w('  { long long acc = 0; for (int i = 0; i < n; ++i) { long long c = rec_%s[i].pmx_tile0; '
          'rec_%s[i].pmx_tile0 = acc; acc += c; } B->total[%d] = acc; }' % (k.name, k.name, ki))
        w('  rc = pmx_malloc(&B->frames[%d], sizeof(%s_frame) * (size_t)n); if (rc) { pmx_batch_%s_free(B); return rc; }'
          % (ki, k.name, name))
        w('  rc = pmx_memcpy_h2d(B->frames[%d], rec_%s.data(), sizeof(%s_frame) * (size_t)n, nullptr); '
          'if (rc) { pmx_batch_%s_free(B); return rc; }' % (ki, k.name, k.name, name))
    # persistent grids: as many workgroups as stay resident on the device
    w('  { int cus = 0; rc = pmx_device_cus(B->dev, &cus); if (rc) { pmx_batch_%s_free(B); return rc; }' % name)
    for ki, k in enumerate(kernels):
        w('    { int occ = 0; rc = pmx_occupancy(pmx_fn_batch[B->dev][%d], %d, 0, &occ); '
          'if (rc) { pmx_batch_%s_free(B); return rc; } if (occ < 1) occ = 1; B->grid[%d] = (long long)occ * cus; }'
          % (ki, k.threads, name, ki))
    w('  }')
    w('  rc = pmx_malloc(&B->counters, 64 * %d); if (rc) { pmx_batch_%s_free(B); return rc; }' % (max(1, nk), name))
    w('  rc = pmx_stream_sync(nullptr); if (rc) { pmx_batch_%s_free(B); return rc; }' % name)
    w('  *handle = (void*)B;')
    w('  return 0;')
    w('}')
    w('extern "C" int pipeline_%s_batch_run(void* handle, void* stream) {' % name)
    w('  pmx_batch_%s* B = (pmx_batch_%s*)handle;' % (name, name))
    w('  int rc = pmx_load(B->dev); if (rc) return rc;')
    w('  { void* c = B->counters; int nc = %d; void* args[] = {&c, &nc};' % (8 * max(1, nk)))
    w('    rc = pmx_launch(pmx_zero_fn[B->dev], 1, 1, 1, 64, 1, 1, 0, stream, args); if (rc) return rc; }')
    for ki, k in enumerate(kernels):
        w('  { void* f = B->frames[%d]; int n = B->n; unsigned long long* c = (unsigned long long*)B->counters + %d; '
          'long long t = B->total[%d];' % (ki, 8 * ki, ki))
        w('    void* args[] = {&f, &n, &c, &t};')
        w('    long long g = t < B->grid[%d] ? t : B->grid[%d];' % (ki, ki))
        w('    if (g > 0) { rc = pmx_launch(pmx_fn_batch[B->dev][%d], (unsigned)g, 1, 1, %d, 1, 1, 0, stream, args); if (rc) return rc; } }'
          % (ki, k.threads))
    w('  return 0;')
    w('}')
    w('extern "C" int pipeline_%s_batch_run_graph(void* handle, void* stream) {' % name)
    w('  pmx_batch_%s* B = (pmx_batch_%s*)handle;' % (name, name))
    w('  int rc = 0;')
    w('  if (!B->graph) {')
    w('    // capture on a private stream (the null stream cannot be captured);')
    w('    // the instantiated graph is then launched on the caller\'s stream')
    w('    void* cs = nullptr; rc = pmx_stream_create(&cs); if (rc) return rc;')
    w('    rc = pmx_graph_begin(cs); if (rc) { pmx_stream_destroy(cs); return rc; }')
    w('    int rc2 = pipeline_%s_batch_run(handle, cs);' % name)
    w('    rc = pmx_graph_end(cs, &B->graph);')
    w('    pmx_stream_destroy(cs);')
    w('    if (rc2) return rc2; if (rc) return rc;')
    w('  }')
    w('  return pmx_graph_launch(B->graph, stream);')
    w('}')
    w('extern "C" int pipeline_%s_batch_run_frames(void* handle, void* stream) {' % name)
    w('  pmx_batch_%s* B = (pmx_batch_%s*)handle;' % (name, name))
    w('  for (int i = 0; i < B->n; ++i) {')
    call = ['B->p_%s[i]' % p.name for p in params] + ['B->io_%s[i]' % f.name for f in inputs + outputs] + ['stream']
    w('    int rc = pipeline_%s_dev(%s); if (rc) return rc;' % (name, ', '.join(call)))
    w('  }')
    w('  return 0;')
    w('}')
    w('extern "C" int pipeline_%s_batch_destroy(void* handle) {' % name)
    w('  return pmx_batch_%s_free((pmx_batch_%s*)handle);' % (name, name))
    w('}')

