/*
 * turtle_c.h -- C ABI of the MI355X (gfx950) TurtleSort library
 * (libturtle_hip.so).  Plain pointers, sizes and status codes; no HIP or
 * torch types.  The C++ drop-in header include/turtle-sort.hpp is built on
 * these calls, and so are the Python bindings.
 *
 * Each declaration names the reference interface it replaces
 * (tugrul512bit/TurtleSort, file:line).
 */
#ifndef TURTLE_C_H
#define TURTLE_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference aborts the process on any CUDA error
 * (TurtleGlobals::gpuErrchk, turtle-globals.cuh:18-30) and only prints
 * device-side problems (turtle-sort.cu:146, 608, 860); every entry point here
 * returns one of these instead. */
typedef enum turtle_status {
    TURTLE_OK = 0,
    TURTLE_ERR_INVALID = 1,        /* null pointer, negative size, bad argument */
    TURTLE_ERR_CAPACITY = 2,       /* n > max_n, or num_arrays*arr_size > max_n */
    TURTLE_ERR_OOM = 3,            /* device or pinned allocation failed */
    TURTLE_ERR_DTYPE = 4,          /* unknown dtype, or dtype differs from the handle's */
    TURTLE_ERR_UNSUPPORTED = 5,    /* feature not available on this device */
    TURTLE_ERR_DEVICE = 6,         /* HIP runtime error */
    TURTLE_ERR_QUEUE_OVERFLOW = 7, /* device work queue / segment table full */
    TURTLE_ERR_TIMEOUT = 8,        /* reserved: no kernel of this library waits on another */
    TURTLE_ERR_STATE = 9,          /* call out of order (Sync without StartSorting...) */
    TURTLE_ERR_BUDGET = 10         /* request larger than the serving memory pool */
} turtle_status;

/* Key types.  The reference instantiates its kernels for char, short, int,
 * long and their unsigned forms plus float/double (turtle-sort.cu:956-1174);
 * long is 32-bit on its only platform (MSVC), so 64-bit keys are added here.
 * Floating-point keys sort in IEEE 754 totalOrder: -NaN < -inf < ... < -0 <
 * +0 < ... < +inf < +NaN.  That agrees with < on every pair < orders, so on
 * NaN-free input the keys equal std::sort's; the two zeros compare equal
 * under <, and here -0 precedes +0 (the stable index follows that order). */
typedef enum turtle_dtype {
    TURTLE_INT8 = 1, TURTLE_UINT8 = 2,
    TURTLE_INT16 = 3, TURTLE_UINT16 = 4,
    TURTLE_INT32 = 5, TURTLE_UINT32 = 6,
    TURTLE_INT64 = 7, TURTLE_UINT64 = 8,
    TURTLE_FLOAT32 = 9, TURTLE_FLOAT64 = 10
} turtle_dtype;

typedef struct turtle_sorter *turtle_handle;

/* Per-sort counters read back from the device after a sort. */
typedef struct turtle_stats {
    double device_ms;          /* device time of the last sort (HIP events) */
    uint64_t n;                /* keys sorted */
    uint64_t bytes_count;      /* algorithmic bytes of the count passes (key bytes read) */
    uint64_t bytes_scatter;    /* algorithmic bytes of the partition scatters (read + write) */
    uint64_t bytes_lds;        /* algorithmic bytes of the LDS sorts (read + write) */
    uint64_t tasks_pivot, tasks_count, tasks_scatter, tasks_lds;
    uint64_t segments;         /* partitioned segments created */
    uint32_t max_depth;        /* deepest partition level reached */
    uint32_t error;            /* device error flags of the last run (0 = none) */
} turtle_stats;

/* Tunables of the sort engine (chosen by the autotuner, or set by hand). */
typedef struct turtle_tuning {
    int32_t splitters;         /* pivots per partition pass: 3 (the reference's
                                  3-pivot partition), 7, 15 or 31 */
    int32_t lds_cutoff;        /* segments at or below this length are sorted in LDS;
                                  1024 .. 8192 (4096 for 8-byte keys) */
    int32_t tiles_per_task;    /* tiles a workgroup takes per step of the count and
                                  scatter loops */
    int32_t workgroups_per_cu; /* grid of the tile kernels = CUs x this (0 = occupancy
                                  limit) */
} turtle_tuning;

/* ---- sorter object: Turtle::TurtleSort<Type,TrackIndex> ------------------ */

/* Replaces TurtleSort(int maxElements, bool optInCompression)
 * (turtle-sort.cuh:63-113).  Allocates device buffers for max_n keys (and
 * indices when track_index != 0).  compress is accepted; generic memory
 * compression does not exist on this device, so the handle reports it off
 * (turtle_compression_enabled) and creation still succeeds. */
int turtle_create(int dtype, int64_t max_n, int track_index, int compress, turtle_handle *out);

/* Replaces ~TurtleSort (turtle-sort.cuh:278-282) and the buffer frees of
 * TurtleBuffer::~Buffer (turtle-buffer.cuh:82-99).  TURTLE_ERR_STATE (and
 * nothing freed) while graphs captured on the handle are alive. */
int turtle_destroy(turtle_handle h);

/* Replaces MemoryCompressionSuccessful (turtle-sort.cuh:115-118); always 0. */
int turtle_compression_enabled(turtle_handle h);

/* Replaces StartSorting(std::vector<Type>*, std::vector<int>*)
 * (turtle-sort.cuh:124-184).  keys/idx are host arrays of n elements; idx
 * may be NULL.  Copies them to the device through pinned staging (the call
 * takes about as long as that copy), queues the device sort on the handle's
 * stream and returns without waiting for it.  The caller keeps both arrays
 * alive and unchanged until turtle_sort_sync. */
int turtle_sort_start(turtle_handle h, void *keys, int32_t *idx, int64_t n);

/* Replaces Sync() (turtle-sort.cuh:188-258): waits, copies the sorted keys
 * (and indices) back into the arrays given to turtle_sort_start and stores
 * the wall-clock seconds since turtle_sort_start in *seconds (transfers
 * included, as the reference). */
int turtle_sort_sync(turtle_handle h, double *seconds);

/* Replaces MultiSort<Type,ArrSize,BlockSize,UseSharedMemory>(numArrays,
 * hostPtr) (turtle-sort.cuh:265-276 -> turtle-multi-sort.cuh:189-206):
 * sorts num_arrays consecutive arrays of arr_size keys each, in place in the
 * host buffer, synchronously; *seconds gets the wall time including
 * transfers.  block_size and use_lds are accepted for source compatibility
 * and do not change the result.  Arrays of up to 8192 keys move in chunks
 * through two pinned buffers on two streams, so that the host copies of one
 * chunk overlap the transfers and the sort of the other; host copies use
 * $TURTLE_COPY_THREADS threads (default 8). */
int turtle_multisort(turtle_handle h, int dtype, int64_t arr_size, int block_size, int use_lds,
                     int64_t num_arrays, void *host_data, double *seconds);

/* MultiSort with index tracking (no reference counterpart: the reference's
 * README.md:46 lists it as unsupported).  Sorts num_arrays consecutive
 * arrays of arr_size keys in place in host_keys and moves a 32-bit value
 * with every key, stably: host_idx holds the values on entry, or, when
 * local_positions is non-zero, is output only and receives each key's
 * position within its array (an argsort per array).  Synchronous; *seconds
 * gets the wall time including transfers.  The handle must track indices. */
int turtle_multisort_index(turtle_handle h, int dtype, int64_t arr_size, int64_t num_arrays, void *host_keys,
                           int32_t *host_idx, int local_positions, double *seconds);

/* ---- device-resident forms (inputs already in HBM) ---------------------- */

/* Device half of StartSorting+Sync: replaces the wave chain
 * copyTasksBack -> resetNumTasks -> quickSortWithoutStreamCompaction/bruteSort
 * (turtle-sort.cu:35-188, 573-934) and the chunk merge of Sync
 * (turtle-sort.cu:190-300, turtle-sort.cuh:192-243) with level-synchronous
 * partition passes over a device-resident segment list followed by one LDS
 * sort launch.  Sorts keys_in[0..n) into keys_out (and carries idx_in into
 * idx_out when both are non-NULL; the handle must track indices).  keys_in
 * is not modified; keys_out may equal keys_in.  stream is a hipStream_t;
 * NULL is the null (default) stream, so the work is ordered after whatever
 * produced the inputs there.  Asynchronous: returns once the work is queued
 * (it first resolves the handle's previous run, see turtle_synchronize).
 * The caller keeps keys_in/idx_in unchanged until turtle_synchronize, which
 * makes the output final. */
int turtle_sort_device(turtle_handle h, const void *keys_in, const int32_t *idx_in,
                       void *keys_out, int32_t *idx_out, int64_t n, void *stream);

/* Allocation-free form of turtle_sort_device (no reference counterpart;
 * for callers that cannot allocate device memory: tight serving loops and
 * harnesses that forbid hipMalloc).  No handle: every buffer of the run
 * beyond the caller's inputs and outputs lives in a static device array of
 * the library (43.6 MiB, held on each GPU it runs on until the process
 * ends), so at most 1,048,576 keys (TURTLE_ERR_CAPACITY above that).  Calls
 * no hipMalloc, hipFree or other allocating routine.  dtype is a
 * turtle_dtype; idx_in/idx_out are both NULL or both set; keys_out may
 * equal keys_in (and idx_out idx_in).  Runs on `stream` (NULL = null
 * stream) with splitters 31 and the largest LDS cutoff, and returns when the
 * sort is done, with the same output as turtle_sort_device.  Calls on one
 * GPU share its array and are serialised by a lock; the array keeps the
 * last call's contents.  Not for use under stream capture (it reads back
 * from the device).  A work-table overflow is reported as
 * TURTLE_ERR_QUEUE_OVERFLOW (there is nothing to grow). */
int turtle_sort_device_noalloc(int dtype, const void *keys_in, const int32_t *idx_in,
                               void *keys_out, int32_t *idx_out, int64_t n, void *stream);

/* Device-wide stable merge of two sorted runs: replaces mergeSortedChunks
 * (turtle-sort.cu:190-281), whose per-element rank rule puts a run-1 key
 * before equal run-2 keys (upper bound of a run-2 key in run 1, lower bound
 * of a run-1 key in run 2), and the chunk merges of Sync
 * (turtle-sort.cuh:192-243).  keys_a[0..na) and keys_b[0..nb) are sorted
 * ascending (in the key type's order, floating-point keys in totalOrder);
 * keys_out[0..na+nb) receives their stable merge.  idx_a, idx_b and idx_out
 * are all NULL or all set: a 32-bit value moves with every key (the
 * reference's merge writes keys into its index array, turtle-sort.cu:220,
 * 255; here the index is carried).  The output must not overlap the runs.
 * Asynchronous on `stream` (NULL = null stream). */
int turtle_merge_device(int dtype, const void *keys_a, const int32_t *idx_a, int64_t na,
                        const void *keys_b, const int32_t *idx_b, int64_t nb,
                        void *keys_out, int32_t *idx_out, void *stream);

/* Batched uneven segments in one engine run (no reference
 * counterpart; the reference's only route is a loop of StartSorting+Sync).
 * Segment s is [offsets[s], offsets[s+1]) of the concatenated arrays;
 * offsets is a HOST array of num_segments+1 entries.  Stream rules as
 * turtle_sort_device. */
int turtle_segmented_sort_device(turtle_handle h, const void *keys_in, const int32_t *idx_in,
                                 void *keys_out, int32_t *idx_out,
                                 const int64_t *offsets, int64_t num_segments, void *stream);

/* Device half of MultiSort (turtle-multi-sort.cuh:86-182): num_arrays
 * consecutive arrays of arr_size keys, keys_in -> keys_out (may alias).
 * Asynchronous. */
int turtle_multisort_device(turtle_handle h, int dtype, int64_t arr_size,
                            const void *keys_in, void *keys_out, int64_t num_arrays, void *stream);

/* Device form of turtle_multisort_index: idx_out (required) receives the
 * values of idx_in moved with their keys, or each key's position within its
 * array when idx_in is NULL.  idx buffers hold arr_size*num_arrays entries
 * (at most 2^32).  Arrays longer than 8192 keys go through the sort engine
 * and need a handle that tracks indices.  Asynchronous. */
int turtle_multisort_index_device(turtle_handle h, int dtype, int64_t arr_size, const void *keys_in,
                                  const int32_t *idx_in, void *keys_out, int32_t *idx_out, int64_t num_arrays,
                                  void *stream);

/* hipGraph form of turtle_segmented_sort_device: the reset, root, level
 * and LDS-sort launches are captured once for a fixed number of levels
 * (offsets copied to a device buffer owned by the graph; no memset nodes)
 * and replayed by turtle_graph_launch on any stream (NULL = null stream).  A graph owns its engine
 * tables (later calls on the handle never resize them) but shares the
 * handle's key/index ping-pong buffers, so at most one run per handle may
 * be in flight.  The graph captures the estimated levels plus two; if an
 * input needs more, turtle_synchronize after the replay reports
 * TURTLE_ERR_QUEUE_OVERFLOW (the output is then incomplete).  keys_out must
 * differ from keys_in.  Destroy a handle's graphs before the handle.  No
 * reference counterpart. */
typedef struct turtle_graph *turtle_graph_handle;
int turtle_segmented_graph_create(turtle_handle h, const void *keys_in, const int32_t *idx_in,
                                  void *keys_out, int32_t *idx_out, const int64_t *offsets,
                                  int64_t num_segments, turtle_graph_handle *out);
int turtle_graph_launch(turtle_graph_handle g, void *stream);
int turtle_graph_destroy(turtle_graph_handle g);

/* Waits for the handle's last device work and fills *st (may be NULL).
 * An engine run is queued with the partition levels its input size
 * suggests; this call also finishes runs that need more (it queues further
 * levels), and after a work-table overflow it grows the tables and sorts
 * again from the input, which no level writes (an in-place call's input is
 * kept in the handle's input buffer).  The next call on the handle does the
 * same for the run before it.  The output is final once this returns OK. */
int turtle_synchronize(turtle_handle h, turtle_stats *st);

/* Non-blocking status of the handle's last device work, as hipStreamQuery
 * (no reference counterpart: the reference's only probe is Sync): 1 while
 * it still runs, 0 when it is done (call turtle_synchronize before reading
 * the output, which also resolves its level check), or a negative status
 * code on error. */
int turtle_query(turtle_handle h);

/* Kernels of one engine run (index of turtle_profile's arrays). */
typedef enum turtle_kernel_kind {
    TURTLE_K_RESET = 0, TURTLE_K_ROOTS = 1, TURTLE_K_LEVEL = 2, TURTLE_K_PIVOT = 3,
    TURTLE_K_COUNT = 4, TURTLE_K_SCAN = 5, TURTLE_K_SCATTER = 6, TURTLE_K_CHILDREN = 7,
    TURTLE_K_LDS = 8, TURTLE_K_LDSPREP = 9, TURTLE_K_KINDS = 10
} turtle_kernel_kind;

typedef struct turtle_profile {
    double ms[TURTLE_K_KINDS];          /* summed device time per kernel kind */
    uint64_t launches[TURTLE_K_KINDS];  /* launches per kernel kind */
} turtle_profile;

/* Per-kernel launch timing (no reference counterpart: the reference times
 * only the whole sort, in Sync).  While profiling is on, every engine run
 * records a HIP event pair around each kernel launch, on the stream the
 * kernel runs on (graph captures record none).  turtle_get_profile waits for
 * those events, returns the sums since the previous call (or since profiling
 * was switched on) and starts new sums. */
int turtle_set_profiling(turtle_handle h, int on);
/* As turtle_set_profiling(h, 1), but only launches of the kernel kinds in
 * kind_mask (bit 1 << turtle_kernel_kind) get an event pair, so that timing
 * one kernel disturbs the rest of the run as little as possible; 0 = off. */
int turtle_set_profiling_kinds(turtle_handle h, uint32_t kind_mask);
int turtle_get_profile(turtle_handle h, turtle_profile *out);

/* Tuning.  Unless turtle_set_tuning fixed it, every run picks its tuning
 * from the autotuned table (tuning_table.txt next to the library, or
 * $TURTLE_TUNING_FILE) by key width, index flag and size, else the built-in
 * default.  turtle_get_tuning returns the tuning of the last run (before
 * any run: the default). */
int turtle_get_tuning(turtle_handle h, turtle_tuning *t);
int turtle_set_tuning(turtle_handle h, const turtle_tuning *t);

/* Autotuner (no reference counterpart: the reference fixes its thresholds
 * at compile time, turtle-sort.cu:7,12 and turtle-sort.cuh:130): times the
 * engine on n synthetic keys (uniform over the key type, iota index when
 * with_index) in device memory for every candidate (splitters 7/15/31 x LDS
 * cutoff cap/4, cap/2, cap x tiles per step 1/2), reps timed runs each, and
 * returns the fastest in *best (and its ms per sort).  The handle's own
 * tuning is left as it was. */
int turtle_autotune(turtle_handle h, int64_t n, int with_index, int reps, turtle_tuning *best, double *best_ms);

/* Autotuner for segmented batches (no reference counterpart: the reference
 * sorts one array per call): the same candidates timed on synthetic keys
 * laid out in the given segments (offsets[0] = 0 .. offsets[num_segments],
 * host memory) through turtle_segmented_sort_device.  Its results are the
 * "seg" lines of the tuning table, which segmented runs prefer. */
int turtle_autotune_segmented(turtle_handle h, const int64_t *offsets, int64_t num_segments, int with_index, int reps,
                              turtle_tuning *best, double *best_ms);

/* ---- request loop with continuous batching ------------------------------ */
/* No reference counterpart: a reference TurtleSort object holds one sort in
 * flight (turtle-sort.cuh:53-61).  A server carves a device memory budget
 * into batch slots (per served key type, slots_per_dtype sorters with their
 * own HIP stream and buffers for slot_keys keys); a slot that is free takes
 * every pending request of its type that fits, in arrival order, and sorts
 * them as the segments of one engine run.  Batches of different slots run
 * concurrently on the device.  Requests are host arrays sorted in place,
 * like StartSorting; idx may be NULL.  All calls are thread-safe. */
typedef struct turtle_server *turtle_server_handle;

typedef struct turtle_server_config {
    int64_t pool_bytes;        /* device memory budget of all slots (0: 64 GiB) */
    int64_t slot_keys;         /* keys per batch slot = the largest request accepted
                                  (0: from the budget, at most 32M) */
    int32_t slots_per_dtype;   /* batches of one key type in flight at once (0: 2) */
    int32_t dtype_mask;        /* bit (1 << dtype) for every served key type (0: all) */
    int64_t inflight_bytes;    /* device bytes the requests being sorted may use at
                                  once (0: pool_bytes).  A request of n keys of K
                                  bytes is charged n * (4 K + 17) bytes (input and
                                  output keys and indices, the engine's ping-pong
                                  keys and indices, per-key bins) while its batch
                                  runs; it joins a batch only when its bytes fit. */
} turtle_server_config;

typedef struct turtle_server_stats {
    uint64_t submitted, completed, rejected, failed, batches, keys;
    uint64_t pool_bytes, slot_keys;
    double busy_seconds;           /* summed batch time of all slots */
    uint32_t slots, in_flight_batches, max_in_flight_batches, pad;
    uint64_t inflight_bytes;       /* the budget of requests being sorted */
    uint64_t bytes_in_flight, max_bytes_in_flight;
} turtle_server_stats;

/* What happened to one request (turtle_server_wait_info). */
typedef struct turtle_request_info {
    double latency_seconds;   /* submission to completion */
    double queued_seconds;    /* submission to the start of its batch */
    double batch_seconds;     /* its batch's run: copies in, sort, copies out */
    int64_t batch_keys;       /* keys of its batch */
    int32_t batch_requests;   /* requests of its batch */
    int32_t status;           /* the request's status */
} turtle_request_info;

/* Allocates every slot (device buffers, pinned host buffers) and starts one
 * worker thread per slot.  TURTLE_ERR_BUDGET if the slots do not fit the
 * pool. */
int turtle_server_create(const turtle_server_config *cfg, turtle_server_handle *out);
/* Queues a request (returns at once with its ticket).  TURTLE_ERR_BUDGET if
 * n exceeds slot_keys or its bytes exceed inflight_bytes; TURTLE_ERR_DTYPE
 * if the type is not served.  keys (and idx) must stay alive and untouched
 * until the ticket is waited for. */
int turtle_server_submit(turtle_server_handle s, int dtype, void *keys, int32_t *idx, int64_t n,
                         uint64_t *ticket);
/* Blocks until the request is sorted; *latency_seconds = completion time -
 * submission time.  Returns the request's status.  A ticket is waited for
 * once. */
int turtle_server_wait(turtle_server_handle s, uint64_t ticket, double *latency_seconds);
/* As turtle_server_wait, and fills *info (may be NULL) with the request's
 * queueing and batch figures. */
int turtle_server_wait_info(turtle_server_handle s, uint64_t ticket, turtle_request_info *info);
/* *done = 1 if the request has finished (then wait returns at once). */
int turtle_server_test(turtle_server_handle s, uint64_t ticket, int *done);
int turtle_server_get_stats(turtle_server_handle s, turtle_server_stats *out);
/* Finishes every queued request, stops the workers, frees the slots. */
int turtle_server_destroy(turtle_server_handle s);

/* ---- synthetic inputs on the device (benchmarks and tests) -------------- */

/* keys[i] = (splitmix64 draw i of seed) >> shift, truncated to the dtype
 * width; draw i is mix(seed + (i+1) * 0x9e3779b97f4a7c15). */
int turtle_fill_keys(int dtype, void *keys, int64_t n, uint64_t seed, int shift, void *stream);
/* idx[i] = i */
int turtle_fill_iota(int32_t *idx, int64_t n, void *stream);

/* ---- misc ---------------------------------------------------------------- */
const char *turtle_status_string(int status);
/* Last error message recorded by this thread (empty string if none). */
const char *turtle_last_error(void);
int turtle_dtype_size(int dtype);
/* Runs, once per device, the check of the hardware property the sort
 * kernels' stable ranks rely on: the lanes of one wave's LDS atomic to the
 * same word receive their return values in ascending lane order (measured
 * on gfx950; every handle creation and the allocation-free sort run it
 * too).  TURTLE_OK, or TURTLE_ERR_UNSUPPORTED if the property does not hold.
 * No reference counterpart. */
int turtle_check_device(void);
/* Number of compute units of the current device (0 if no device). */
int turtle_device_cus(void);

#ifdef __cplusplus
}
#endif

#endif /* TURTLE_C_H */
