/*
 * turtle_oracle.c -- CPU oracle for the TurtleSort sort path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the GPU library is
 * compared against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py load it; the product library never links or calls it.
 *
 * What it restates (reference = tugrul512bit/TurtleSort, CUDA C++):
 *   - ts_oracle_turtle_sort: the whole StartSorting()+Sync() algorithm run
 *     sequentially on the CPU: host 8-chunk split (turtle-sort.cuh:130-166),
 *     the 3-pivot / 7-bin partition with pivot-duplicate counting
 *     (turtle-sort.cu:641-851), child routing by size (turtle-sort.cu:857-933),
 *     the <=32 sorting networks (turtle-sort.cu:392-559, 615-631), the 33..128
 *     odd-even transposition "brute" sort (turtle-sort.cu:118-188) and the
 *     3-level rank merge of the 8 chunks (turtle-sort.cuh:188-243,
 *     turtle-sort.cu:190-281).
 *   - ts_oracle_multisort: the per-array heap sort of MultiSort
 *     (turtle-multi-sort.cuh:15-182).
 *
 * Keys are handled through an order-preserving map to uint64_t
 * (ts_oracle_to_ord): unsigned types widen unchanged, signed types widen with
 * sign extension and then have bit 63 flipped.  Every comparison the
 * reference makes with operator< / operator> on the original type gives the
 * same answer on the mapped value, so the restated algorithm takes the same
 * decisions and produces the same key sequence.
 *
 * float and double keys (instantiated by the reference, turtle-sort.cu:
 * 1136-1174, and sorted by its only test, main.cpp:6-30) map through their
 * bit pattern: non-negative values get the sign bit set, negative values
 * every bit inverted, then widen unchanged.  That is IEEE 754 totalOrder
 * (-NaN < -inf < ... < -0 < +0 < ... < +inf < +NaN).  On NaN-free input
 * it answers every comparison < answers, except that -0 < +0 where < sees
 * them equal; the reference's own comparisons (turtle-sort.cu:562-567) are
 * not a strict weak order once NaN is present, so NaN placement is a policy
 * of this build, not a restatement.
 *
 * Two deliberate differences from the reference's observable behaviour, both
 * defects documented against the reference source:
 *   - the merge moves the tracked index (the reference moves the key into the
 *     index array, turtle-sort.cu:220 and 255);
 *   - the heap sort keeps its swap temporary at full key width (the reference
 *     narrows it to int, turtle-multi-sort.cuh:122 and 167).
 * Within a partition bin the reference's element order is decided by
 * shared-memory atomics (turtle-sort.cu:777-822) and is not reproducible; the
 * restatement uses ascending input order, which is one schedule the reference
 * can take.  Index order among equal keys is therefore checked as a multiset.
 *
 * How each function is pinned (tests/test_oracle.py):
 *   ts_oracle_turtle_sort   -> equal to libc qsort on the same keys, brute
 *                              force over all arrangements of small inputs,
 *                              (key,idx) pairs preserved
 *   ts_oracle_network_sort  -> 0-1 principle, exhaustive (2^n inputs, n<=32)
 *   ts_oracle_oddeven_sort  -> every permutation of n<=8, stability
 *   ts_oracle_merge_ranks   -> textbook sequential stable merge
 *   ts_oracle_partition3    -> bin predicates, bin contiguity, count
 *                              conservation
 *   ts_oracle_multisort     -> libc qsort per array
 *   ts_oracle_splitmix64    -> published splitmix64 output for seed 1234567
 *   float/double order map  -> numpy's sort (bit-exact on NaN-free input
 *                              without mixed zeros); with NaN and both zeros:
 *                              value equality with numpy's sort of the
 *                              non-NaN part, -0 before +0, sign-set NaNs
 *                              first and the others last
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

/* dtype codes; the same numbering as the TURTLE_DTYPE_* values of the C ABI
 * (a test checks that they agree) */
enum {
    ORC_INT8 = 1, ORC_UINT8 = 2, ORC_INT16 = 3, ORC_UINT16 = 4,
    ORC_INT32 = 5, ORC_UINT32 = 6, ORC_INT64 = 7, ORC_UINT64 = 8,
    ORC_FLOAT32 = 9, ORC_FLOAT64 = 10
};

int ts_oracle_dtype_size(int dtype)
{
    switch (dtype) {
    case ORC_INT8: case ORC_UINT8: return 1;
    case ORC_INT16: case ORC_UINT16: return 2;
    case ORC_INT32: case ORC_UINT32: case ORC_FLOAT32: return 4;
    case ORC_INT64: case ORC_UINT64: case ORC_FLOAT64: return 8;
    default: return 0;
    }
}

/* ------------------------------------------------------------------------ */
/* order-preserving map                                                      */
/* ------------------------------------------------------------------------ */
#define SIGN64 0x8000000000000000ULL
#define SIGN32 0x80000000u

/* IEEE bit pattern -> totalOrder rank of the same width, and back */
static uint32_t f32_ord(uint32_t b) { return (b & SIGN32) ? ~b : (b | SIGN32); }
static uint32_t f32_unord(uint32_t o) { return (o & SIGN32) ? (o & ~SIGN32) : ~o; }
static uint64_t f64_ord(uint64_t b) { return (b & SIGN64) ? ~b : (b | SIGN64); }
static uint64_t f64_unord(uint64_t o) { return (o & SIGN64) ? (o & ~SIGN64) : ~o; }

int ts_oracle_to_ord(int dtype, const void *in, uint64_t *out, int64_t n)
{
    int64_t i;
    switch (dtype) {
    case ORC_INT8:   for (i = 0; i < n; i++) out[i] = (uint64_t)(int64_t)((const int8_t *)in)[i] ^ SIGN64; break;
    case ORC_UINT8:  for (i = 0; i < n; i++) out[i] = ((const uint8_t *)in)[i]; break;
    case ORC_INT16:  for (i = 0; i < n; i++) out[i] = (uint64_t)(int64_t)((const int16_t *)in)[i] ^ SIGN64; break;
    case ORC_UINT16: for (i = 0; i < n; i++) out[i] = ((const uint16_t *)in)[i]; break;
    case ORC_INT32:  for (i = 0; i < n; i++) out[i] = (uint64_t)(int64_t)((const int32_t *)in)[i] ^ SIGN64; break;
    case ORC_UINT32: for (i = 0; i < n; i++) out[i] = ((const uint32_t *)in)[i]; break;
    case ORC_INT64:  for (i = 0; i < n; i++) out[i] = (uint64_t)((const int64_t *)in)[i] ^ SIGN64; break;
    case ORC_UINT64: for (i = 0; i < n; i++) out[i] = ((const uint64_t *)in)[i]; break;
    case ORC_FLOAT32: for (i = 0; i < n; i++) out[i] = f32_ord(((const uint32_t *)in)[i]); break;
    case ORC_FLOAT64: for (i = 0; i < n; i++) out[i] = f64_ord(((const uint64_t *)in)[i]); break;
    default: return -1;
    }
    return 0;
}

int ts_oracle_from_ord(int dtype, const uint64_t *in, void *out, int64_t n)
{
    int64_t i;
    switch (dtype) {
    case ORC_INT8:   for (i = 0; i < n; i++) ((int8_t *)out)[i] = (int8_t)(int64_t)(in[i] ^ SIGN64); break;
    case ORC_UINT8:  for (i = 0; i < n; i++) ((uint8_t *)out)[i] = (uint8_t)in[i]; break;
    case ORC_INT16:  for (i = 0; i < n; i++) ((int16_t *)out)[i] = (int16_t)(int64_t)(in[i] ^ SIGN64); break;
    case ORC_UINT16: for (i = 0; i < n; i++) ((uint16_t *)out)[i] = (uint16_t)in[i]; break;
    case ORC_INT32:  for (i = 0; i < n; i++) ((int32_t *)out)[i] = (int32_t)(int64_t)(in[i] ^ SIGN64); break;
    case ORC_UINT32: for (i = 0; i < n; i++) ((uint32_t *)out)[i] = (uint32_t)in[i]; break;
    case ORC_INT64:  for (i = 0; i < n; i++) ((int64_t *)out)[i] = (int64_t)(in[i] ^ SIGN64); break;
    case ORC_UINT64: for (i = 0; i < n; i++) ((uint64_t *)out)[i] = in[i]; break;
    case ORC_FLOAT32: for (i = 0; i < n; i++) ((uint32_t *)out)[i] = f32_unord((uint32_t)in[i]); break;
    case ORC_FLOAT64: for (i = 0; i < n; i++) ((uint64_t *)out)[i] = f64_unord(in[i]); break;
    default: return -1;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* sorting networks (published tables of B. Dobbelaere, as embedded in       */
/* turtle-sort.cu:392-452).  Comparators are listed layer by layer; each      */
/* layer is wire-disjoint, so applying them in list order equals the         */
/* reference's layer-parallel execution.                                     */
/* ------------------------------------------------------------------------ */
static const unsigned char NET4[][2] = {
    {0,2},{1,3}, {0,1},{2,3}, {1,2}
};
static const unsigned char NET8[][2] = {
    {0,2},{1,3},{4,6},{5,7}, {0,4},{1,5},{2,6},{3,7}, {0,1},{2,3},{4,5},{6,7},
    {2,4},{3,5}, {1,4},{3,6}, {1,2},{3,4},{5,6}
};
static const unsigned char NET16[][2] = {
    {0,13},{1,12},{2,15},{3,14},{4,8},{5,6},{7,11},{9,10},
    {0,5},{1,7},{2,9},{3,4},{6,13},{8,14},{10,15},{11,12},
    {0,1},{2,3},{4,5},{6,8},{7,9},{10,11},{12,13},{14,15},
    {0,2},{1,3},{4,10},{5,11},{6,7},{8,9},{12,14},{13,15},
    {1,2},{3,12},{4,6},{5,7},{8,10},{9,11},{13,14},
    {1,4},{2,6},{5,8},{7,10},{9,13},{11,14},
    {2,4},{3,6},{9,12},{11,13},
    {3,5},{6,8},{7,9},{10,12},
    {3,4},{5,6},{7,8},{9,10},{11,12},
    {6,7},{8,9}
};
static const unsigned char NET32[][2] = {
    {0,1},{2,3},{4,5},{6,7},{8,9},{10,11},{12,13},{14,15},{16,17},{18,19},{20,21},{22,23},{24,25},{26,27},{28,29},{30,31},
    {0,2},{1,3},{4,6},{5,7},{8,10},{9,11},{12,14},{13,15},{16,18},{17,19},{20,22},{21,23},{24,26},{25,27},{28,30},{29,31},
    {0,4},{1,5},{2,6},{3,7},{8,12},{9,13},{10,14},{11,15},{16,20},{17,21},{18,22},{19,23},{24,28},{25,29},{26,30},{27,31},
    {0,8},{1,9},{2,10},{3,11},{4,12},{5,13},{6,14},{7,15},{16,24},{17,25},{18,26},{19,27},{20,28},{21,29},{22,30},{23,31},
    {0,16},{1,8},{2,4},{3,12},{5,10},{6,9},{7,14},{11,13},{15,31},{17,24},{18,20},{19,28},{21,26},{22,25},{23,30},{27,29},
    {1,2},{3,5},{4,8},{6,22},{7,11},{9,25},{10,12},{13,14},{17,18},{19,21},{20,24},{23,27},{26,28},{29,30},
    {1,17},{2,18},{3,19},{4,20},{5,10},{7,23},{8,24},{11,27},{12,28},{13,29},{14,30},{21,26},
    {3,17},{4,16},{5,21},{6,18},{7,9},{8,20},{10,26},{11,23},{13,25},{14,28},{15,27},{22,24},
    {1,4},{3,8},{5,16},{7,17},{9,21},{10,22},{11,19},{12,20},{14,24},{15,26},{23,28},{27,30},
    {2,5},{7,8},{9,18},{11,17},{12,16},{13,22},{14,20},{15,19},{23,24},{26,29},
    {2,4},{6,12},{9,16},{10,11},{13,17},{14,18},{15,22},{19,25},{20,21},{27,29},
    {5,6},{8,12},{9,10},{11,13},{14,16},{15,17},{18,20},{19,23},{21,22},{25,26},
    {3,5},{6,7},{8,9},{10,12},{11,14},{13,16},{15,18},{17,20},{19,21},{22,23},{24,25},{26,28},
    {3,4},{5,6},{7,8},{9,10},{11,12},{13,14},{15,16},{17,18},{19,20},{21,22},{23,24},{25,26},{27,28}
};

/* width 4/8/16/32 -> table; returns comparator count */
int ts_oracle_network_table(int width, const unsigned char **table)
{
    switch (width) {
    case 4:  *table = &NET4[0][0];  return (int)(sizeof(NET4) / 2);
    case 8:  *table = &NET8[0][0];  return (int)(sizeof(NET8) / 2);
    case 16: *table = &NET16[0][0]; return (int)(sizeof(NET16) / 2);
    case 32: *table = &NET32[0][0]; return (int)(sizeof(NET32) / 2);
    default: *table = 0; return 0;
    }
}

/* turtle-sort.cu:456-559 + 622-629: pick the smallest network >= num and
 * skip comparators that touch wires >= num (equivalent to +inf padding). */
int ts_oracle_network_sort(uint64_t *a, int32_t *id, int num)
{
    const unsigned char *t;
    int width = num <= 4 ? 4 : num <= 8 ? 8 : num <= 16 ? 16 : num <= 32 ? 32 : 0;
    int nc, c;
    if (width == 0) return -1;
    nc = ts_oracle_network_table(width, &t);
    for (c = 0; c < nc; c++) {
        int x = t[2 * c], y = t[2 * c + 1];
        if (x < num && y < num && a[y] < a[x]) {
            uint64_t tk = a[x]; a[x] = a[y]; a[y] = tk;
            if (id) { int32_t ti = id[x]; id[x] = id[y]; id[y] = ti; }
        }
    }
    return 0;
}

/* turtle-sort.cu:161-179: num rounds of an even phase then an odd phase;
 * compare-swap exchanges only when strictly out of order (compSw, 116). */
void ts_oracle_oddeven_sort(uint64_t *a, int32_t *id, int num)
{
    int r, i, ph;
    for (r = 0; r < num; r++) {
        for (ph = 0; ph < 2; ph++) {
            for (i = ph; i + 1 < num; i += 2) {
                if (a[i + 1] < a[i]) {
                    uint64_t tk = a[i]; a[i] = a[i + 1]; a[i + 1] = tk;
                    if (id) { int32_t ti = id[i]; id[i] = id[i + 1]; id[i + 1] = ti; }
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* one 3-pivot partition step (turtle-sort.cu:641-851)                       */
/* bins in memory order: LL, =pL, L, =p, R, =pR, RR (755-761)                */
/* ------------------------------------------------------------------------ */
static int bin_of(uint64_t v, uint64_t pl, uint64_t p, uint64_t pr)
{
    /* equality with the pivots is tested first, left to right (688-699) */
    if (v == pl) return 1;
    if (v == p) return 3;
    if (v == pr) return 5;
    if (v < pl) return 0;
    if (v < p) return 2;
    if (v < pr) return 4;
    return 6;
}

/* Partitions a[0..num) in place using tmp buffers; writes the 7 bin counts
 * to counts[7] and the sorted pivots to piv[3].  Returns 0. */
int ts_oracle_partition3(uint64_t *a, int32_t *id, int64_t num, uint64_t *tmpk, int32_t *tmpi,
                         int64_t counts[7], uint64_t piv[3])
{
    uint64_t pv[3], t;
    int64_t off[7], pos[7], i;
    int b;
    /* pivots: a[mid], a[start], a[stop], then three compare-swaps (642-650) */
    pv[0] = a[num / 2];
    pv[1] = a[0];
    pv[2] = a[num - 1];
    if (pv[1] < pv[0]) { t = pv[0]; pv[0] = pv[1]; pv[1] = t; }
    if (pv[2] < pv[0]) { t = pv[0]; pv[0] = pv[2]; pv[2] = t; }
    if (pv[2] < pv[1]) { t = pv[1]; pv[1] = pv[2]; pv[2] = t; }
    for (b = 0; b < 7; b++) counts[b] = 0;
    for (i = 0; i < num; i++) counts[bin_of(a[i], pv[0], pv[1], pv[2])]++;
    off[0] = 0;
    for (b = 1; b < 7; b++) off[b] = off[b - 1] + counts[b - 1];
    for (b = 0; b < 7; b++) pos[b] = off[b];
    for (i = 0; i < num; i++) {
        b = bin_of(a[i], pv[0], pv[1], pv[2]);
        tmpk[pos[b]] = a[i];
        if (id) tmpi[pos[b]] = id[i];
        pos[b]++;
    }
    /* copy back; pivot bins are rewritten from the pivot values (838-843) */
    for (i = 0; i < num; i++) {
        if (i >= off[1] && i < off[2]) a[i] = pv[0];
        else if (i >= off[3] && i < off[4]) a[i] = pv[1];
        else if (i >= off[5] && i < off[6]) a[i] = pv[2];
        else a[i] = tmpk[i];
        if (id) id[i] = tmpi[i];
    }
    piv[0] = pv[0]; piv[1] = pv[1]; piv[2] = pv[2];
    return 0;
}

/* ------------------------------------------------------------------------ */
/* rank merge of two adjacent sorted runs (turtle-sort.cu:190-281)           */
/* run 2 element: upper bound in run 1 (226-241); run 1 element: lower bound */
/* in run 2 (261-274).  Output rank = own offset + rank in the other run.    */
/* ------------------------------------------------------------------------ */
static int64_t upper_bound_u64(const uint64_t *a, int64_t n, uint64_t v)
{
    int64_t l = 0, r = n;
    while (l < r) { int64_t m = l + (r - l) / 2; if (!(v < a[m])) l = m + 1; else r = m; }
    return l;
}
static int64_t lower_bound_u64(const uint64_t *a, int64_t n, uint64_t v)
{
    int64_t l = 0, r = n;
    while (l < r) { int64_t m = l + (r - l) / 2; if (v > a[m]) l = m + 1; else r = m; }
    return l;
}

void ts_oracle_merge_ranks(const uint64_t *k1, const int32_t *i1, int64_t n1,
                           const uint64_t *k2, const int32_t *i2, int64_t n2,
                           uint64_t *ko, int32_t *io)
{
    int64_t j;
    for (j = 0; j < n2; j++) {
        int64_t d = upper_bound_u64(k1, n1, k2[j]) + j;
        ko[d] = k2[j];
        if (io) io[d] = i2[j];
    }
    for (j = 0; j < n1; j++) {
        int64_t d = lower_bound_u64(k2, n2, k1[j]) + j;
        ko[d] = k1[j];
        if (io) io[d] = i1[j];
    }
}

/* ------------------------------------------------------------------------ */
/* the whole StartSorting()+Sync() path                                     */
/* ------------------------------------------------------------------------ */
#define BRUTE_LIMIT 128   /* turtle-sort.cu:7 */
#define DIRECT_LIMIT 32   /* turtle-sort.cu:12 */

typedef struct { int64_t start, stop; } task_t;   /* inclusive, as the reference */

/* sorts one host chunk [start, stop] with the wave loop of turtle-sort.cu */
static int sort_chunk(uint64_t *a, int32_t *id, int64_t start, int64_t stop,
                      uint64_t *tmpk, int32_t *tmpi)
{
    /* a LIFO stack visits the same task tree as the reference's waves; tasks
     * are disjoint so the visiting order does not change the result */
    int64_t cap = 64, top = 0;
    task_t *quick = (task_t *)malloc((size_t)cap * sizeof(task_t));
    if (!quick) return -1;
    quick[top].start = start; quick[top].stop = stop; top++;
    while (top > 0) {
        task_t t = quick[--top];
        int64_t num = t.stop - t.start + 1, counts[7], base[7];
        uint64_t piv[3];
        int b;
        if (num < 2) continue;                               /* 612-613 */
        if (num <= DIRECT_LIMIT) {                           /* 619-631 */
            ts_oracle_network_sort(a + t.start, id ? id + t.start : 0, (int)num);
            continue;
        }
        ts_oracle_partition3(a + t.start, id ? id + t.start : 0, num,
                             tmpk + t.start, id ? tmpi + t.start : 0, counts, piv);
        base[0] = t.start;
        for (b = 1; b < 7; b++) base[b] = base[b - 1] + counts[b - 1];
        /* children LL, L, R, RR with more than one element (862-932) */
        for (b = 0; b < 7; b += 2) {
            int64_t c = counts[b];
            if (c <= 1) continue;
            if (c <= BRUTE_LIMIT && c > DIRECT_LIMIT) {
                ts_oracle_oddeven_sort(a + base[b], id ? id + base[b] : 0, (int)c);
            } else {
                if (top == cap) {
                    task_t *nq;
                    cap *= 2;
                    nq = (task_t *)realloc(quick, (size_t)cap * sizeof(task_t));
                    if (!nq) { free(quick); return -1; }
                    quick = nq;
                }
                quick[top].start = base[b]; quick[top].stop = base[b] + c - 1; top++;
            }
        }
    }
    free(quick);
    return 0;
}

/* keys: mapped keys (ts_oracle_to_ord); idx may be NULL.  Sorted in place. */
int ts_oracle_turtle_sort(uint64_t *keys, int32_t *idx, int64_t n)
{
    uint64_t *tk, *mk;
    int32_t *ti = 0, *mi = 0;
    int64_t cs, bounds[9];
    int c, rc = 0;
    if (n <= 1) return 0;
    tk = (uint64_t *)malloc((size_t)n * sizeof(uint64_t));
    mk = (uint64_t *)malloc((size_t)n * sizeof(uint64_t));
    if (idx) {
        ti = (int32_t *)malloc((size_t)n * sizeof(int32_t));
        mi = (int32_t *)malloc((size_t)n * sizeof(int32_t));
    }
    if (!tk || !mk || (idx && (!ti || !mi))) { rc = -1; goto done; }

    cs = n / 8;                                          /* turtle-sort.cuh:139 */
    if (cs > 1024) {
        /* 8 chunks; the last takes the remainder (144-150) */
        int level;
        for (c = 0; c < 8; c++) bounds[c] = (int64_t)c * cs;
        bounds[8] = n;
        for (c = 0; c < 8; c++)
            if (sort_chunk(keys, idx, bounds[c], bounds[c + 1] - 1, tk, ti)) { rc = -1; goto done; }
        /* merge rows of Sync(): 8 -> 4 -> 2 -> 1 runs (192-201) */
        for (level = 1; level < 8; level *= 2) {
            for (c = 0; c < 8; c += 2 * level) {
                int64_t s1 = bounds[c], e1 = bounds[c + level], e2 = bounds[c + 2 * level];
                ts_oracle_merge_ranks(keys + s1, idx ? idx + s1 : 0, e1 - s1,
                                      keys + e1, idx ? idx + e1 : 0, e2 - e1,
                                      mk + s1, idx ? mi + s1 : 0);
            }
            memcpy(keys, mk, (size_t)n * sizeof(uint64_t));
            if (idx) memcpy(idx, mi, (size_t)n * sizeof(int32_t));
        }
    } else {
        rc = sort_chunk(keys, idx, 0, n - 1, tk, ti);
    }
done:
    free(tk); free(mk); free(ti); free(mi);
    return rc;
}

/* typed convenience: keys of any dtype, sorted in place */
int ts_oracle_turtle_sort_typed(int dtype, void *keys, int32_t *idx, int64_t n)
{
    uint64_t *m;
    int rc;
    if (ts_oracle_dtype_size(dtype) == 0) return -1;
    if (n <= 1) return 0;
    m = (uint64_t *)malloc((size_t)n * sizeof(uint64_t));
    if (!m) return -1;
    ts_oracle_to_ord(dtype, keys, m, n);
    rc = ts_oracle_turtle_sort(m, idx, n);
    ts_oracle_from_ord(dtype, m, keys, n);
    free(m);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* MultiSort heap sort (turtle-multi-sort.cuh:33-84, 112-126)                */
/* ------------------------------------------------------------------------ */
static void heap_insert(uint64_t *b, int64_t index)
{
    /* sift-up, iterative form of the recursion at 33-47 */
    for (;;) {
        int64_t parent = ((index + 1) / 2) - 1;
        if (parent < 0 || parent == index) return;
        if (b[parent] < b[index]) {
            uint64_t t = b[parent]; b[parent] = b[index]; b[index] = t;
            index = parent;
        } else return;
    }
}

static void heap_remove(uint64_t *b, int64_t index, int64_t size)
{
    /* sift-down, iterative form of 49-84 */
    for (;;) {
        int64_t l = ((index + 1) * 2) - 1, r = l + 1, next = -1;
        if (l < size && r < size) {
            if (b[l] > b[r]) { if (b[index] < b[l]) next = l; }
            else { if (b[index] < b[r]) next = r; }
        } else if (l < size) {
            if (b[index] < b[l]) next = l;
        }
        if (next < 0) return;
        { uint64_t t = b[index]; b[index] = b[next]; b[next] = t; }
        index = next;
    }
}

void ts_oracle_heapsort(uint64_t *b, int64_t n)
{
    int64_t i;
    for (i = 1; i < n; i++) heap_insert(b, i);
    for (i = 0; i < n; i++) {
        int64_t cur = n - i;
        uint64_t top = b[0];            /* full width (the reference used int) */
        b[0] = b[cur - 1];
        heap_remove(b, 0, cur - 1);
        b[cur - 1] = top;
    }
}

int ts_oracle_multisort_typed(int dtype, void *data, int64_t num_arrays, int64_t arr_size)
{
    int64_t a;
    int es = ts_oracle_dtype_size(dtype);
    uint64_t *m;
    if (es == 0 || arr_size < 0 || num_arrays < 0) return -1;
    if (arr_size == 0 || num_arrays == 0) return 0;
    m = (uint64_t *)malloc((size_t)arr_size * sizeof(uint64_t));
    if (!m) return -1;
    for (a = 0; a < num_arrays; a++) {
        char *p = (char *)data + (size_t)a * (size_t)arr_size * (size_t)es;
        ts_oracle_to_ord(dtype, p, m, arr_size);
        ts_oracle_heapsort(m, arr_size);
        ts_oracle_from_ord(dtype, m, p, arr_size);
    }
    free(m);
    return 0;
}

/* segmented form used for batched uneven segments: every segment sorted with
 * the StartSorting algorithm; offsets has num_segments+1 entries */
int ts_oracle_segmented_sort_typed(int dtype, void *keys, int32_t *idx,
                                   const int64_t *offsets, int64_t num_segments)
{
    int64_t s;
    int es = ts_oracle_dtype_size(dtype);
    if (es == 0) return -1;
    for (s = 0; s < num_segments; s++) {
        int64_t b = offsets[s], e = offsets[s + 1];
        if (e < b) return -1;
        if (ts_oracle_turtle_sort_typed(dtype, (char *)keys + (size_t)b * (size_t)es,
                                        idx ? idx + b : 0, e - b)) return -1;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* synthetic inputs: splitmix64 (Steele, Lea, Flood 2014; Vigna's constants) */
/* ------------------------------------------------------------------------ */
uint64_t ts_oracle_splitmix64(uint64_t *state)
{
    uint64_t z = (*state += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

/* key i = (i-th draw from seed) >> shift, truncated to the dtype width */
int ts_oracle_gen_keys(int dtype, uint64_t seed, int shift, void *out, int64_t n)
{
    uint64_t s = seed;
    int64_t i;
    int es = ts_oracle_dtype_size(dtype);
    if (es == 0 || shift < 0 || shift > 63) return -1;
    for (i = 0; i < n; i++) {
        uint64_t v = ts_oracle_splitmix64(&s) >> shift;
        switch (es) {
        case 1: ((uint8_t *)out)[i] = (uint8_t)v; break;
        case 2: ((uint16_t *)out)[i] = (uint16_t)v; break;
        case 4: ((uint32_t *)out)[i] = (uint32_t)v; break;
        default: ((uint64_t *)out)[i] = v; break;
        }
    }
    return 0;
}

/* lengths log-uniform in [lo, hi]: floor(exp(ln lo + u (ln hi - ln lo))),
 * u = 53-bit uniform draw in [0,1) */
int ts_oracle_gen_loguniform(uint64_t seed, int64_t lo, int64_t hi, int64_t *out, int64_t n)
{
    uint64_t s = seed;
    int64_t i;
    double a, b;
    if (lo < 1 || hi < lo) return -1;
    a = log((double)lo); b = log((double)hi);
    for (i = 0; i < n; i++) {
        double u = (double)(ts_oracle_splitmix64(&s) >> 11) * (1.0 / 9007199254740992.0);
        int64_t v = (int64_t)floor(exp(a + u * (b - a)));
        if (v < lo) v = lo;
        if (v > hi) v = hi;
        out[i] = v;
    }
    return 0;
}
