/* paraiso_hip.h -- C ABI of the MI355X (gfx950) Orthotope Machine runtime.
 *
 * This ABI replaces the class the reference generator emits for each
 * Orthotope Machine (Language/Paraiso/Generator/PlanTrans.hs:50-276, the
 * CS2/CS3 call stacks): construction, one call per OM kernel, accessors to the
 * static variables, and the size queries.  The generated host code for a
 * machine (one shared library per machine) exports a single descriptor,
 * `pm_describe`, whose launch callbacks start that machine's HIP kernels;
 * everything else -- device storage, ping-pong stores, scalar slots,
 * reduction partials, streams, HIP graphs and the adjoint driver -- lives in
 * this runtime.
 *
 * Conventions: every entry point returns 0 on success and a negative status
 * on failure (PR_E_*), and never throws.  `pr_last_error` describes the last
 * failure of a handle.  A handle is not thread-safe; different handles may be
 * driven from different host threads and streams.  All work of a handle is
 * issued on its stream (pr_set_stream; a fresh non-blocking stream of default
 * priority otherwise).  Array statics are dense, x-fastest (index
 * x + nx*(y + ny*z)) over the memory extent: the local size on Cyclic axes,
 * the local size plus the plan's lower and upper margins on Open axes.
 */
#ifndef PARAISO_HIP_H
#define PARAISO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PR_ABI_VERSION 4

enum pr_status {
  PR_OK = 0,
  PR_E_INVALID = -1,     /* bad argument (index, size, pointer)          */
  PR_E_HIP = -2,         /* a HIP runtime call failed                     */
  PR_E_NOMEM = -3,       /* device allocation failed                      */
  PR_E_ABI = -4,         /* descriptor ABI version mismatch               */
  PR_E_UNSUPPORTED = -5  /* the machine does not provide the operation    */
};

enum pr_realm { PR_SCALAR = 0, PR_ARRAY = 1 };
enum pr_ctype { PR_BOOL = 0, PR_INT = 1, PR_FLOAT = 2, PR_DOUBLE = 3 };

typedef struct pr_machine pr_machine;

/* What a generated launch callback receives (one kernel launch of one OM kernel). */
typedef struct pr_launch_ctx {
  const void* const* cur;  /* per static: current device buffer (arrays) or NULL */
  void* const* next;       /* per static: buffer a store writes (stored arrays) or NULL */
  char* scal;              /* device scalar block: statics, slots, counters */
  char* part;              /* this launch's reduction partials */
  unsigned* counter;       /* a zeroed device word per launch, reserved: the generated
                              kernels finish reductions in a separate one-workgroup kernel
                              and do not use it */
  int n[3];                /* grid extent per axis (1 for absent axes) */
  int tiles[3];            /* tiles per axis */
  unsigned ntiles;         /* product of tiles[] */
  void* stream;            /* hipStream_t */
} pr_launch_ctx;

typedef int (*pr_launch_fn)(const pr_launch_ctx* ctx);

/* What a generated batched launcher receives: one launch over the tiles of
 * many grids of the same machine. */
typedef struct pr_batch_ctx {
  const void* args;        /* device array: one packed argument block per grid */
  const int* tile_grid;    /* device: grid of each workgroup */
  const unsigned* tile_base; /* device: first workgroup of each grid */
  unsigned total_tiles;
  int n_grids;
  void* stream;
} pr_batch_ctx;

typedef void (*pr_pack_fn)(const pr_launch_ctx* ctx, void* out);
typedef int (*pr_batch_launch_fn)(const pr_batch_ctx* ctx);

enum pr_static_flags { PR_STATIC_SCRATCH = 1 };

typedef struct pr_static_desc {
  const char* name;
  int realm;               /* pr_realm */
  int ctype;               /* pr_ctype */
  int flags;               /* PR_STATIC_SCRATCH: internal array between two launches of a kernel */
} pr_static_desc;

typedef struct pr_launch_desc {
  pr_launch_fn fn;
  int tile[3];             /* workgroup tile; 0 tiles = one workgroup (scalar-only launch) */
  int n_reduce;            /* reductions finished by this launch's tail kernel */
  pr_pack_fn pack;         /* packs one grid's argument block */
  size_t args_size;
  pr_batch_launch_fn batched;
} pr_launch_desc;

typedef struct pr_kernel_desc {
  const char* name;
  int n_launch;
  const pr_launch_desc* launches;
  int n_stored;            /* array statics this kernel stores (ping-pong flipped after it) */
  const int* stored;
  int n_loaded;            /* statics this kernel reads */
  const int* loaded;
  int n_written;           /* statics this kernel stores, arrays and scalars */
  const int* written;
} pr_kernel_desc;

typedef struct pr_machine_desc {
  int abi;                 /* PR_ABI_VERSION */
  const char* name;
  int ndim;
  int n_static;
  const pr_static_desc* statics;
  int n_kernel;
  const pr_kernel_desc* kernels;
  int n_slot;              /* scalar slots after the statics in the scalar block */
  int halo;                /* minimum grid extent per axis the kernels need */
  int lower_margin[3];     /* ghost cells below / above the local size on Open axes */
  int upper_margin[3];     /* (Plan.lowerMargin / upperMargin, Generator/Plan.hs:33-92) */
} pr_machine_desc;

/* Exported by every generated machine library. */
const pr_machine_desc* pm_describe(void);

/* ---- lifetime (replaces the generated class constructor, PlanTrans.hs:75-86) */
int pr_create(const pr_machine_desc* desc, const int* local_size, pr_machine** out);
int pr_destroy(pr_machine* m);
int pr_set_stream(pr_machine* m, void* hip_stream);
int pr_sync(pr_machine* m);
const char* pr_last_error(const pr_machine* m);

/* ---- OM kernels (replaces `void <kernel>()`, PlanTrans.hs:225-276)
 * pr_run issues `count` consecutive calls of kernel `k`; pr_run_graph does the
 * same through a cached HIP graph of `count` calls. */
int pr_run(pr_machine* m, int k, int count);
int pr_run_graph(pr_machine* m, int k, int count);
/* pr_run with a HIP event pair around every launch; launch_ms[l] receives the
 * summed device time of launch l over the `count` calls (no counterpart in the
 * reference: measurement support). */
int pr_run_timed(pr_machine* m, int k, int count, double* launch_ms);

/* ---- statics (replaces `T& name()` / `T& name(i0,i1,...)`, PlanTrans.hs:117-157) */
int pr_static_devptr(pr_machine* m, int s, void** dev_ptr);
int pr_static_write(pr_machine* m, int s, const void* src, size_t bytes, int src_on_device);
int pr_static_read(pr_machine* m, int s, void* dst, size_t bytes, int dst_on_device);
size_t pr_static_bytes(const pr_machine* m, int s);

/* ---- sizes (replace om_size, om_memory_size and friends, PlanTrans.hs:167-215) */
int pr_size(const pr_machine* m, int axis);
/* Extent of the arrays, margins included (om_memory_size_k, PlanTrans.hs:167-215); equals
 * pr_size on Cyclic axes.  Static arrays hold prod(pr_memory_size) elements. */
int pr_memory_size(const pr_machine* m, int axis);
long long pr_cells(const pr_machine* m);

/* ---- whole-machine state (the paper's "save the entire machine state to a file" and
 * "load from such save-points", paper/om.tex:679-703; the reference generator has no
 * counterpart).  Every user static is written with its name, type and bytes; loading
 * checks the machine, grid size and statics before changing anything. */
int pr_state_save(pr_machine* m, const char* path);
int pr_state_load(pr_machine* m, const char* path);

/* ---- batches: many independent grids of one machine, of uneven sizes,
 * advanced by the same launches (no counterpart in the reference).  Grid g
 * is a full machine (pr_batch_grid) for static access; all grids share the
 * batch's stream.  pr_batch_run_graph captures `count` calls in one HIP graph. */
typedef struct pr_batch pr_batch;
int pr_batch_create(const pr_machine_desc* desc, int n_grids, const int* sizes, pr_batch** out);
int pr_batch_destroy(pr_batch* b);
pr_machine* pr_batch_grid(pr_batch* b, int g);
int pr_batch_run(pr_batch* b, int k, int count);
int pr_batch_run_graph(pr_batch* b, int k, int count);
int pr_batch_run_timed(pr_batch* b, int k, int count, double* launch_ms);
int pr_batch_sync(pr_batch* b);
const char* pr_batch_last_error(const pr_batch* b);
long long pr_batch_cells(const pr_batch* b);

/* ---- checkpointed adjoints (no counterpart in the reference: the reference
 * has no adjoint).  A snapshot holds copies of the listed statics.
 * pr_adjoint_forward runs `nsteps` calls of kernel `fwd`, saving a checkpoint
 * every `every` calls into slots [0, ceil(nsteps/every)).
 * pr_adjoint_backward then walks the segments backwards: restore the
 * segment's checkpoint, replay it storing each state in slots
 * [S, S+every), and run kernel `adj` once per state in reverse order.  With
 * the adjoint statics holding dJ/du_N on entry they hold dJ/du_0 on return,
 * and the snapshotted statics hold u_0.  Needs ceil(nsteps/every) + every
 * slots.  When `adj` reads none of the snapshotted statics that `fwd` writes
 * (a linear scheme: its adjoint does not depend on the state), the replay is
 * skipped and `adj` runs `nsteps` times back to back. */
int pr_snapshot_alloc(pr_machine* m, int n_slots, const int* statics, int n_statics);
int pr_snapshot_save(pr_machine* m, int slot);
int pr_snapshot_load(pr_machine* m, int slot);
int pr_adjoint_forward(pr_machine* m, int fwd, int nsteps, int every);
int pr_adjoint_backward(pr_machine* m, int fwd, int adj, int nsteps, int every);
/* The same for a step made of several kernels: the forward step runs kernels
 * fwd[0..n_fwd) in order, and its adjoint runs adj[0..n_adj) in order.  The
 * adjoint kernels see the state at the tape point, after fwd[0..tape_after)
 * have run (e.g. a two-stage step split into a stage-1 and a stage-2 kernel,
 * tape_after = 1: the adjoint sees the state and the half-step state).
 * The replay keeps each step's state in its own snapshot slot: every kernel
 * reads the slot of its step and writes what it stores straight into the
 * slot that value belongs to, so the backward pass copies no state per step.
 * pr_adjoint_backward(m, f, a, n, e) is the sequence form with one kernel
 * each and tape_after = 0. */
int pr_adjoint_forward_seq(pr_machine* m, const int* fwd, int n_fwd, int nsteps, int every);
int pr_adjoint_backward_seq(pr_machine* m, const int* fwd, int n_fwd, int tape_after,
                            const int* adj, int n_adj, int nsteps, int every);

/* ---- running without allocating (no counterpart in the reference: for callers that cannot
 * allocate device memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Runs `count` calls of kernel `k` of the machine `desc` on a
 * grid of `local_size` cells, on storage the caller owns, and allocates nothing:
 *   buffers  2 * n_static device pointers; buffers[2s] and buffers[2s + 1] are the two
 *            buffers of array static s (each prod(memory extent) elements of its type;
 *            the second may be NULL unless kernel k stores s); scalar entries are ignored;
 *   parity   n_static ints: which of the two holds static s on entry; updated on return
 *            (a store writes the other one, as pr_run's ping-pong does);
 *   scalars  device block of 8 bytes per static, scalar static s at byte 8 s (the layout
 *            pr_static_devptr exposes): read before the first call, written after the last;
 *   stream   hipStream_t the work is issued on (NULL: the default stream);
 *   err      optional buffer for a message on failure.
 * The scalar slots, reduction partials and scratch arrays the kernels need live in one
 * fixed device array of PR_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends;
 * a kernel and grid that need more are rejected with PR_E_INVALID before any work is
 * issued.  Calls on one GPU share that array: calls that overlap on one GPU (other
 * streams, other host threads) race and must be serialised by the caller. */
#define PR_NOALLOC_SCRATCH_BYTES (32u << 20)
int pr_run_noalloc(const pr_machine_desc* desc, const int* local_size, int k, int count,
                   void* const* buffers, int* parity, void* scalars, void* stream,
                   char* err, size_t err_len);

#ifdef __cplusplus
}
#endif
#endif /* PARAISO_HIP_H */
