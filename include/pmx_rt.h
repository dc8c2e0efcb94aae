/*
 * pmx_rt.h — C ABI of libpmx_rt.so, the native runtime under the
 * MI355X PolyMage backend.  Plain pointers and sizes only.
 *
 * The reference's native layer is (a) the generated pipeline_<name>()
 * function in a g++-built .so, loaded with ctypes
 * (sandbox/apps/python/loader.py:4-21, sandbox/tuner.py:525), and (b) the
 * size-matched pool allocator of sandbox/memory_allocation/
 * simple_pool_allocator.cpp:17-115 (pool_init / pool_allocate /
 * pool_deallocate / pool_destroy).  This library supplies the MI355X
 * equivalents: a hiprtc compile service used by the Python emitter, a
 * device arena, module/launch helpers the generated host stubs call,
 * hipGraph capture, event timing, and the cost model's MLP kernels.
 *
 * Every function returns 0 on success or a negative / hipError_t code;
 * pmx_last_error() gives the message of the last failure on this thread.
 */
#ifndef PMX_RT_H
#define PMX_RT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- errors ------------------------------------------------------------ */
const char* pmx_last_error(void);
void pmx_set_error(const char* msg);

/* ---- compile service (works without a GPU) ------------------------------
 * Compile HIP source with hiprtc for `arch` (e.g. "gfx950") into a code
 * object written to `out_path`.  `opts` are extra compiler options.  The
 * compiler log (NUL-terminated, truncated to log_cap) goes to `log`. */
int pmx_compile_hip(const char* src, const char* name, const char* arch,
                    const char* const* opts, int nopts, const char* out_path,
                    char* log, size_t log_cap);

/* ---- device info ------------------------------------------------------- */
int pmx_device_count(int* n);
int pmx_device_name(int dev, char* buf, size_t cap);

/* ---- device arena -------------------------------------------------------
 * Replaces pool_allocate/pool_deallocate for arrays that live between
 * kernels (live-outs of one group read by the next).  A slot is a named,
 * per-device buffer that only grows; it is kept until pmx_arena_release().
 * Not thread safe (like the reference pool, simple_pool_allocator.cpp:17). */
int pmx_arena_get(const char* slot, size_t bytes, void** ptr);
/* Same, but the slot is zero-filled when it is (re)allocated. */
int pmx_arena_get_zeroed(const char* slot, size_t bytes, void** ptr);
int pmx_arena_release(void);
size_t pmx_arena_bytes(void);

/* ---- modules and launches (used by generated host stubs) --------------- */
int pmx_module_load_data(const void* image, void** module);
int pmx_module_function(void* module, const char* name, void** function);
int pmx_module_unload(void* module);
int pmx_launch(void* function, unsigned gx, unsigned gy, unsigned gz,
               unsigned bx, unsigned by, unsigned bz, unsigned shmem,
               void* stream, void** args);

/* Resident workgroups of `function` per compute unit at `threads` threads
 * and `dyn_lds` bytes of dynamic LDS (sizes persistent launches). */
int pmx_occupancy(void* function, int threads, size_t dyn_lds, int* blocks_per_cu);
/* Compute units of device `dev`. */
int pmx_device_cus(int dev, int* n);

/* ---- streams, events, graphs --------------------------------------------- */
/* Streams are created at the default priority (non-blocking w.r.t. the
 * null stream); graph capture needs a stream other than the null stream. */
int pmx_stream_create(void** stream);
int pmx_stream_destroy(void* stream);
int pmx_stream_sync(void* stream);
int pmx_event_create(void** ev);
int pmx_event_destroy(void* ev);
int pmx_event_record(void* ev, void* stream);
int pmx_event_elapsed_ms(void* start, void* stop, float* ms);
int pmx_graph_begin(void* stream);
int pmx_graph_end(void* stream, void** graph_exec);
int pmx_graph_launch(void* graph_exec, void* stream);
int pmx_graph_destroy(void* graph_exec);

/* ---- memory helpers (host-pointer entry points, tests) ------------------- */
int pmx_malloc(void** ptr, size_t bytes);
int pmx_free(void* ptr);
int pmx_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int pmx_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
int pmx_memset_d8(void* dst, int value, size_t bytes, void* stream);

/* ---- static scratch (pmx_scratch.hip) ------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): a file-scope device array of
 * PMX_SCRATCH_BYTES on every GPU that runs the library, held until the
 * process ends, where the pipeline_<name>_dev_noalloc entry points keep
 * their status word and intermediates.  pmx_scratch() gives the current
 * device's address of the array (looked up once per device with
 * hipGetSymbolAddress, that device current) and its size; pmx_scratch_peek()
 * gives the address only if it was looked up before (else NULL) and never
 * touches the device.  All users on one GPU share its array; calls that
 * overlap on one GPU race. */
#define PMX_SCRATCH_BYTES (64u << 20)
int pmx_scratch(void** ptr, size_t* bytes);
int pmx_scratch_peek(void** ptr);

/* ---- reference pool API kept for callers (video_demo.py:73-83) --------- */
void pool_init(void);
void pool_destroy(void);

/* ---- learned cost model: MLP kernels (pmx_mlp.hip) ----------------------
 * Fully connected ReLU network, row-major float32.  Layer l maps
 * dims[l] -> dims[l+1]; W_l is [dims[l+1], dims[l]], b_l is [dims[l+1]].
 * `params` packs W_0, b_0, W_1, b_1, ... contiguously; `acts` holds the
 * per-layer activations ([batch, dims[l]] for l = 0..nlayers), filled by
 * the forward pass and consumed by the backward pass.  The last layer is
 * linear.  All pointers are device pointers. */
int pmx_mlp_param_count(const int* dims, int nlayers, int* count);
int pmx_mlp_forward(const float* params, const int* dims, int nlayers,
                    const float* x, int batch, float* acts, void* stream);
/* Mean squared error of the last layer against `target` [batch, dims[L]];
 * writes dL/dparams into `grads` (same packing as params) and the loss
 * (sum over the batch / batch) into loss[0]. */
int pmx_mlp_backward(const float* params, const int* dims, int nlayers,
                     const float* acts, const float* target, int batch,
                     float* grads, float* loss, void* stream);
/* Adam update: params -= lr * mhat / (sqrt(vhat) + eps). */
int pmx_mlp_adam(float* params, const float* grads, float* m, float* v,
                 int count, float lr, float beta1, float beta2, float eps,
                 int step, void* stream);
/* `epochs` full-batch epochs (forward, backward, Adam) launched back to
 * back; loss[e] = loss of epoch e (device array of `epochs` floats); Adam
 * steps step0+1 .. step0+epochs.  `acts` and `grads` as above. */
int pmx_mlp_train(float* params, const int* dims, int nlayers, const float* x,
                  const float* target, int batch, int epochs, float lr, float beta1,
                  float beta2, float eps, float* m, float* v, float* acts, float* grads,
                  float* loss, int step0, void* stream);
/* The same loop for `nmodels` networks of one shape trained side by side
 * on the same x / target (an ensemble from different initial weights):
 * params, m, v and grads are [nmodels][count]; `hidden` holds each model's
 * activations of layers 1..nlayers ([nmodels][sum_l batch*dims[l]]; x is
 * read in place as layer 0); loss is [nmodels][epochs].  Each model's
 * arithmetic is exactly that of pmx_mlp_train on its own. */
int pmx_mlp_train_ensemble(float* params, const int* dims, int nlayers, int nmodels,
                           const float* x, const float* target, int batch, int epochs,
                           float lr, float beta1, float beta2, float eps, float* m, float* v,
                           float* hidden, float* grads, float* loss, int step0, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PMX_RT_H */
