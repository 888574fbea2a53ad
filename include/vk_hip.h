/*
 * vk_hip.h -- the C ABI of verkehrt's MI355X-native `hip` kernel variant.
 *
 * verkehrt's plugin API for kernels is its kernel table: every kernel has entries per
 * variant of five shapes (elementwise, reduction, selection, scan, whole;
 * src/kernels.rs:28-154), installed per variant slot (src/kernels.rs:254-336) and run by
 * `Dispatch` (src/kernels/dispatch.rs:56-265). A `hip` variant is one more slot whose
 * entries cross this C ABI. Every entry here mirrors one of those shapes, takes kernels
 * by the reference's names ("dot_f64", "filter_gt_f64", "argsort_u64", ...), checks its
 * arguments with the reference's rules and reports errors with the reference's exact
 * message text (errors are compared as strings, tests/isa_variants.rs:305-330).
 *
 * Arrays live in HBM. A `vk_array` is an immutable, reference-counted view (element
 * offset, length, type) into a device buffer, the device form of `Array`
 * (src/value.rs:215-223). Entries never free or mutate their inputs; results are new
 * arrays owned by the caller (one reference each).
 *
 * All entries of one `vk_ctx` run in order on that context's HIP stream. An entry returns
 * after its result is final: a returned scalar or error is known on the host; a returned
 * array is complete for every later use on the same context (and for vk_array_to_host).
 * A context is not thread safe: use one context per host thread. Arrays may be shared between
 * contexts: reference counts are atomic and memory returns to its creator's pool from any
 * thread. An array made on context A is ready for context B after vk_ctx_sync(A), and B's
 * user keeps a reference until B has finished with it (vk_ctx_sync(B)).
 *
 * Every function returns 0 on success. On failure it returns nonzero and fills `st`
 * (when not NULL): code VK_ERR_RUNTIME carries a runtime error of the kernel, whose text
 * equals the reference entry's `Err(String)`; VK_ERR_USAGE an argument the kernel does not
 * accept (the reference's kind/arity/length errors, also with its text); VK_ERR_DEVICE a
 * HIP failure.
 */
#ifndef VK_HIP_H
#define VK_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Element types, the storage of ScalarType (src/value.rs:40-70). Bool is one byte, 0 or 1. */
typedef enum vk_type {
  VK_F64 = 0,
  VK_F32 = 1,
  VK_I64 = 2,
  VK_U64 = 3,
  VK_I32 = 4,
  VK_U8 = 5,
  VK_BOOL = 6
} vk_type;

enum { VK_OK = 0, VK_ERR_RUNTIME = 1, VK_ERR_USAGE = 2, VK_ERR_DEVICE = 3 };

typedef struct vk_status {
  int32_t code;
  char msg[512];
} vk_status;

typedef struct vk_ctx vk_ctx;
typedef struct vk_array vk_array;
typedef struct vk_reducer vk_reducer;
typedef struct vk_scanner vk_scanner;
typedef struct vk_builder vk_builder;

/* A scalar: its type and its storage bits, little-endian, zero-extended to 64 bits
 * (Scalar, src/value.rs:116-130). */
typedef struct vk_scalar {
  int32_t type;
  uint64_t bits;
} vk_scalar;

/* One operand of an entry: rows of an array view or one scalar repeated for every row
 * (Arg::Rows / Arg::Splat, src/value.rs:461-467). */
enum { VK_ROWS = 0, VK_SPLAT = 1 };
typedef struct vk_arg {
  int32_t kind;
  const vk_array *rows; /* VK_ROWS */
  vk_scalar splat;      /* VK_SPLAT */
} vk_arg;

/* A whole-kernel or Dispatch::call result (Value, src/value.rs:639-644). */
enum { VK_VALUE_ARRAY = 0, VK_VALUE_SCALAR = 1 };
typedef struct vk_value {
  int32_t kind;
  vk_array *array;
  vk_scalar scalar;
} vk_value;

/* Kernel classes, KernelClass (src/kernels.rs:405-412). */
enum { VK_ELEMENTWISE = 0, VK_REDUCTION = 1, VK_SELECTION = 2, VK_SCAN = 3, VK_WHOLE = 4 };

/* ------------------------------------------------------------------ context */

/* A device, its stream and its memory pool. No counterpart in the reference (which runs on
 * the host thread); it is the device handle a `hip` entry needs as its precondition, the
 * analogue of "the caller guarantees that the CPU supports the entry's ISA"
 * (src/kernels/dispatch.rs:1-6). */
int vk_ctx_create(int device, vk_ctx **out, vk_status *st);
void vk_ctx_destroy(vk_ctx *ctx);
/* Waits for all work of the context. */
int vk_ctx_sync(vk_ctx *ctx, vk_status *st);
/* The context's hipStream_t, for callers that order their own work against it. */
void *vk_ctx_stream(vk_ctx *ctx);
/* Returns cached pool memory to the device. */
void vk_ctx_trim(vk_ctx *ctx);
/* Bytes the pool holds (in use and cached). */
uint64_t vk_ctx_pool_bytes(const vk_ctx *ctx, uint64_t *in_use);

/* Trace: with tracing on, every entry call records one line in the reference's trace
 * format, `trace <kernel> isa=hip n=<rows> read=<B> write=<B> time=<us> gbps=<GB/s>`
 * (src/exec.rs:390-420), timed with HIP events on the context stream around the entry's
 * kernels. vk_ctx_trace_take copies the lines recorded since the last take into buf
 * (NUL-terminated, truncated to cap) and returns the full length. */
void vk_ctx_set_trace(vk_ctx *ctx, int on);
size_t vk_ctx_trace_take(vk_ctx *ctx, char *buf, size_t cap);
/* Device time of the kernels of the last entry call, in microseconds (0 without trace). */
double vk_ctx_last_us(const vk_ctx *ctx);

/* ------------------------------------------------------------------- arrays */

/* A new array of `len` rows (uninitialised device memory). Array::try_zeroed
 * (src/value.rs:297-307) zero-fills; device results are always fully written instead.
 * Fails with "array of {len} elements is too large" / "cannot allocate an array of {len}
 * elements" as the reference does. */
int vk_array_new(vk_ctx *ctx, int32_t type, uint64_t len, vk_array **out, vk_status *st);
/* Copies `len` host rows into a new device array. */
int vk_array_from_host(vk_ctx *ctx, int32_t type, const void *host, uint64_t len, vk_array **out,
                       vk_status *st);
/* Copies the rows of `a` to host memory of vk_array_bytes(a) bytes (waits for the stream). */
int vk_array_to_host(vk_ctx *ctx, const vk_array *a, void *host, vk_status *st);
/* Views external device memory (for example a torch tensor) without taking ownership: the
 * caller keeps the memory alive and unchanged while the view or anything derived is used. */
int vk_array_wrap(vk_ctx *ctx, int32_t type, void *device_ptr, uint64_t len, vk_array **out,
                  vk_status *st);
/* Zero-copy sub-view of rows start..start+len (Array::slice, src/value.rs:236-249). */
int vk_array_slice(const vk_array *a, uint64_t start, uint64_t len, vk_array **out, vk_status *st);
void vk_array_retain(vk_array *a);
void vk_array_release(vk_array *a);
uint64_t vk_array_len(const vk_array *a);
int32_t vk_array_type(const vk_array *a);
uint64_t vk_array_bytes(const vk_array *a);
/* The device address of row 0 of the view. */
void *vk_array_device_ptr(const vk_array *a);

/* --------------------------------------------------------------- the table */

/* The kernels of the hip variant, by reference name (Registry::by_name,
 * src/kernels.rs:610-612). vk_kernel_lookup returns the kernel id or -1. */
int32_t vk_kernel_count(void);
int32_t vk_kernel_lookup(const char *name);
const char *vk_kernel_name(int32_t id);
/* KernelDef::class (src/kernels.rs:414-431) as VK_ELEMENTWISE..VK_WHOLE. */
int32_t vk_kernel_class(int32_t id);
/* The variant tag this library installs, "isa=hip" (Variant's Display, src/kernels.rs:284-297). */
const char *vk_variant_name(void);

/* ---------------------------------------------------------------- entries */

/* Dispatch::call (src/kernels/dispatch.rs:56-141): runs kernel `name` over whole arguments.
 * It checks arity and kinds, unifies array lengths ("length mismatch: {n} vs {m}"), then
 * runs an elementwise kernel into a new array, a reduction over all rows (none when empty),
 * a scan into a new array, a selection with base 0, or a whole kernel as is. */
int vk_call(vk_ctx *ctx, const char *name, const vk_arg *args, size_t nargs, vk_value *out,
            vk_status *st);

/* ReducerFn / Reducer (src/kernels.rs:38-63): a reduction state fed tile by tile. Every
 * update but the last has a multiple of 16 rows; finish returns the scalar. The state lives
 * on the device, so updates queue without host round trips. */
int vk_reducer_new(vk_ctx *ctx, const char *name, vk_reducer **out, vk_status *st);
int vk_reducer_update(vk_reducer *r, const vk_arg *args, size_t nargs, vk_status *st);
int vk_reducer_finish(vk_reducer *r, vk_scalar *out, vk_status *st);
void vk_reducer_free(vk_reducer *r);

/* ScanFn / Scanner (src/kernels.rs:65-85): a scan state fed tile by tile; `out` receives
 * exactly the tile's output rows and the result never depends on the split. */
int vk_scanner_new(vk_ctx *ctx, const char *name, vk_scanner **out, vk_status *st);
int vk_scanner_update(vk_scanner *s, const vk_arg *args, size_t nargs, vk_array *out,
                      vk_status *st);
void vk_scanner_free(vk_scanner *s);

/* ArrayBuilder (src/value.rs:469-604) on the device: a growable array selections append to. */
int vk_builder_new(vk_ctx *ctx, int32_t type, vk_builder **out, vk_status *st);
uint64_t vk_builder_len(const vk_builder *b);
/* Turns the builder into an array (one reference, owned by the caller) and frees it. */
int vk_builder_finish(vk_builder *b, vk_array **out, vk_status *st);
void vk_builder_free(vk_builder *b);

/* SelectionFn (src/kernels.rs:87-96) through Dispatch::select_tile (dispatch.rs:231-265):
 * appends the selected rows of one tile whose first row is `base` of the whole array. */
int vk_select_tile(vk_ctx *ctx, const char *name, const vk_arg *args, size_t nargs, uint64_t base,
                   vk_builder *out, vk_status *st);

/* ElementwiseFn (src/kernels.rs:28-36) through Dispatch::elementwise_tile
 * (dispatch.rs:192-229): writes exactly the rows of `out`. */
int vk_elementwise_tile(vk_ctx *ctx, const char *name, const vk_arg *args, size_t nargs,
                        vk_array *out, vk_status *st);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid allocation): the selection
 * `name` (filter_<cmp>_<t>, compress_<t>, where) over whole arguments, as vk_call runs it (base
 * 0), writing the kept rows to the caller's array `out` (of the kernel's output type and at
 * least as many rows as the input) and their count to *kept. Results equal vk_call's bit for
 * bit. Its look-back status words, tile ticket and count live in a fixed device array of the
 * library, one per GPU held until the process ends (about 1 MiB), so the call allocates
 * nothing: inputs of more than VK_SELECT_INTO_MAX_ROWS rows are rejected. Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU race. It records no trace line. */
#define VK_SELECT_INTO_MAX_ROWS (1ull << 31)
int vk_select_into(vk_ctx *ctx, const char *name, const vk_arg *args, size_t nargs, vk_array *out,
                   uint64_t *kept, vk_status *st);

/* A fused region (src/exec/region.rs:1-60, src/exec/region/fused.rs:96-204): elementwise
 * kernels in post-order, each argument a region input or an earlier kernel's result, feeding
 * one reduction `root` (whose arguments are sources too). The result equals evaluating every
 * kernel with vk_call and then the root, bit for bit, and so do the errors. When the kernels
 * form one chain over at most two input arrays (each reads the previous result and a constant,
 * an input array or that result again; y only under a canonical sum) and the root is
 * reduce_add_f64, an integer sum, an extreme, an arg extreme, count_<cmp>_<t>, or count over a
 * final comparison with a constant, the region runs as one device pass with the chain in
 * registers and *fused is set to 1; otherwise it runs kernel by kernel and *fused is 0. */
enum { VK_SRC_INPUT = 0, VK_SRC_KERNEL = 1 };
typedef struct vk_src {
  int32_t kind;
  int32_t index;
} vk_src;
typedef struct vk_rkernel {
  const char *name;
  vk_src args[3];
  int32_t nargs;
} vk_rkernel;
int vk_region_reduce(vk_ctx *ctx, const vk_rkernel *kernels, size_t nkernels, const char *root,
                     const vk_src *root_args, size_t nroot, const vk_arg *inputs, size_t ninputs,
                     vk_scalar *out, int32_t *fused, vk_status *st);

/* Releases the array a vk_value holds, if any. */
void vk_value_release(vk_value *v);

/* ---------------------------------------------------------------- batching */

/* No counterpart in the reference (its interpreter is single threaded); for callers that run
 * one kernel over many independent arrays. Runs kernel `name` on `nitems` argument lists,
 * item i's arguments being args[i * nargs .. (i + 1) * nargs), with one launch per kernel
 * phase for all items instead of one call per item, and writes item i's result to out[i].
 * Every result equals what vk_call returns for that item alone. Natively batched:
 * filter_<cmp>_<t> (one constant for all items), compress_<t>, where, reduce_add_f64,
 * dot_f64, sumsq_f64 and condsum_<cmp>_f64 (one constant pair), the other reductions
 * (count_<cmp>_<t> with one constant), argsort_<t>, group_id_<t> and reduce_into_add_f64;
 * other kernels and mixed constants run item by item. On an error no result is returned and `st` holds the first
 * failing item's error, the kernel's own text. Array results of one call share one device
 * buffer. */
int vk_batch_call(vk_ctx *ctx, const char *name, const vk_arg *args, size_t nargs, size_t nitems,
                  vk_value *out, vk_status *st);

/* ---------------------------------------------------------------- batch plans */

/* No counterpart in the reference; for callers that run the same statements over the same
 * resident arrays again and again (a serving loop, a repeated query). A plan is a sequence of
 * batched statements over `nitems` items: each statement is one kernel by name, run on every
 * item, whose arguments refer to a plan input of the item, to an earlier statement's array
 * result for the item, or to one scalar for all items. Compiling allocates every result,
 * optionally times the tile shapes of each statement on the plan's own inputs and keeps the
 * fastest, and captures into one hipGraph the sequence of kernel launches, whose tiles are
 * planned on the device from the row counts earlier statements leave there (a selection's kept
 * rows); a run replays the graph with no host work and no read-back. Results equal vk_call on
 * each item alone, bit for bit.
 * Plannable kernels: filter_<cmp>_<t> (over a plan input or an earlier selection),
 * compress_<t> (plan inputs), reduce_add_f64, sumsq_f64 and condsum_<cmp>_f64 (over a plan
 * input or a selection), dot_f64 (plan inputs), the other reductions (over a plan input or a
 * selection), argsort_<t> and group_id_<t> (plan inputs), and reduce_into_add_f64(ids, x, g)
 * whose ids and bucket count g both name one group_id statement (its group count per item, on
 * the device) and whose x is a plan input. Argsort, group_id and that bucket sum decide their
 * radix digits, table sizes and bucket counts on the device, so they are captured too.
 * A plan retains its inputs; they must not change while it is in use. */
typedef struct vk_bplan vk_bplan;

enum { VK_BREF_INPUT = 0, VK_BREF_STMT = 1, VK_BREF_SCALAR = 2 };

/* A plan argument: a plan input (index = input number), an earlier statement's array result
 * (index = statement number), or a scalar (`scalar`). */
typedef struct vk_bref {
  int32_t kind;
  int32_t index;
  vk_scalar scalar;
} vk_bref;

/* A plan over `nitems` items with `ninputs` arrays each: item i's input k is
 * inputs[i * ninputs + k]. */
int vk_bplan_new(vk_ctx *ctx, vk_array *const *inputs, size_t ninputs, size_t nitems, vk_bplan **out,
                 vk_status *st);
/* Appends a statement; its number goes to *stmt. Checks kinds, element types and lengths. */
int vk_bplan_add(vk_bplan *plan, const char *name, const vk_bref *args, size_t nargs, size_t *stmt,
                 vk_status *st);
/* Allocates every result, autotunes the tile shapes when `tune` is nonzero (selection rows per
 * thread 8 or 16, reduction tiles of 4096, 8192 or 16384 rows) and captures the graph. */
int vk_bplan_compile(vk_bplan *plan, int tune, vk_status *st);
/* Enqueues one replay of the plan's graph on the context's stream and returns. */
int vk_bplan_run(vk_bplan *plan, vk_status *st);
/* Statement `stmt`'s result for item `item` after the last run (waits for it): an array (a view
 * of the plan's buffer, one reference owned by the caller) or a scalar. */
int vk_bplan_result(vk_bplan *plan, size_t stmt, size_t item, vk_value *out, vk_status *st);
/* Statement `stmt`'s results for every item at once after the last run (waits for it): a
 * selection's kept row counts, an argsort's rows, a group_id's group counts, a bucket sum's
 * bucket counts, or every item's scalar result bits (out has nitems words). */
int vk_bplan_counts(vk_bplan *plan, size_t stmt, uint64_t *out, vk_status *st);
/* The autotuner's measurements and choices, one line per tuned statement; returns the full
 * length (copies at most cap - 1 bytes and a terminating zero). */
size_t vk_bplan_tuning(const vk_bplan *plan, char *buf, size_t cap);
void vk_bplan_free(vk_bplan *plan);

#ifdef __cplusplus
}
#endif

#endif /* VK_HIP_H */
