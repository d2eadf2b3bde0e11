// engine.hpp - internal declarations of the MI355X Louvain engine.
//
// Layout in HBM (one arena per context, first-fit with coalescing):
//   input   : ids[n0] (only when ids are not 0..n0-1), el_src/el_dst[e0] positions,
//             el_w[e0] units, 4 or 8 bytes              (persistent until the next load)
//   level l : CSR rows [rowptr[v], rowend[v]) of col int32 / w units (non-self entries;
//             level 0 packed, coarsened levels keep each row where it was aggregated);
//             per vertex k, loop, selfsum (int64), C/next (int32 ping-pong),
//             per community deg/in (int64), size (int32) and their deltas;
//             vertex lists bucketed by row length for the sweep tiers
//   dendro  : per level cfinal[n] and newc[n] (int32), final[n0]   (persistent per run)
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <vector>

#include "../../../include/louvain_amd.h"

namespace lv {

typedef unsigned long long ull;

// Weight sums stay below this many units so that 4*m and sum(deg) fit in int64.
constexpr int64_t kUnitLimit = int64_t(1) << 60;

// Sweep tiers by q = row length + (1 if the vertex has a self-loop).
// Tiers 1..11 keep an LDS table per row (tiers 9 and 10: rows of 513..1024 and
// 1025..2048 candidates, one 256-lane block with a 2048- or 4096-slot table;
// tier 11: rows of 2049..4096, one 512-lane block with an 8192-slot table, when
// hub_q is above 2048); tier 12 is the hub partitions.  Coarsening segments use
// tiers 1..10 and the partitions.
constexpr int kTiers = LV_SWEEP_TIERS;
constexpr int kTierMaxQ[kTiers] = {4,    8,    16,   32,   64,   128,
                                   256,  512,  1024, 2048, 4096, 0x7fffffff};
constexpr int32_t kHubQDefault = 4096;

#define LV_TRY(expr)                                \
    do {                                            \
        int lv_rc_ = (expr);                        \
        if (lv_rc_ != LV_OK) return lv_rc_;         \
    } while (0)

#define LV_HIP(expr)                                \
    do {                                            \
        hipError_t lv_e_ = (expr);                  \
        if (lv_e_ != hipSuccess) return ::lv::hip_fail(lv_e_, #expr, __FILE__, __LINE__); \
    } while (0)

int hip_fail(hipError_t e, const char *what, const char *file, int line);

// ---------------------------------------------------------------------------
// HBM arena: one hipMalloc, host-side first-fit bookkeeping.
class Arena {
public:
    ~Arena();
    int reserve(size_t bytes);  // (re)allocates only when larger than the current block
    // Manages a block the arena does not own (never freed or grown here).
    void adopt(void *base, size_t bytes);
    void *alloc(size_t bytes);
    void release(void *p);
    void release_all();
    // The live allocations now; restore() releases every allocation made since
    // (what a call that failed part-way left behind).
    std::map<size_t, size_t> mark() const { return used_; }
    void restore(const std::map<size_t, size_t> &m);
    size_t capacity() const { return cap_; }
    size_t in_use() const { return used_bytes_; }
    size_t high_water() const { return high_; }

private:
    char *base_ = nullptr;
    size_t cap_ = 0;
    bool owned_ = true;
    size_t used_bytes_ = 0, high_ = 0;
    std::map<size_t, size_t> free_;  // offset -> size
    std::map<size_t, size_t> used_;  // offset -> size
};

// Typed arena allocation helper.
template <typename T>
int arena_alloc(Arena &a, T *&p, int64_t count)
{
    size_t bytes = sizeof(T) * (size_t)(count > 0 ? count : 1);
    p = (T *)a.alloc(bytes);
    return p ? LV_OK : LV_ERR_NOMEM;
}

template <typename T>
void arena_free(Arena &a, T *&p)
{
    if (p) a.release((void *)p);
    p = nullptr;
}

// ---------------------------------------------------------------------------
// Exact modularity aggregates, filled by device reductions.
struct QAcc {
    long long sum_in;  // sum over communities of internals
    ull d2_lo, d2_hi;  // sum over communities of deg^2 (128-bit)
    long long pad;
};

double q_from_aggregates(int64_t sum_in, uint64_t d2_lo, uint64_t d2_hi, int64_t m, double gamma);

// Per-sweep kernel arguments.
// A vertex's view of its community at the start of a pass: the community id
// and that community's size and total degree, in one 16-byte record, so a
// neighbour's community and its total arrive with one gather.
struct alignas(16) Lab {
    int32_t c;
    int32_t sz;
    int64_t deg;
};

// The same record without the size when every community total fits 32 bits
// (the level's total degree does): half the bytes per neighbour gather (the
// size is needed only for the vertex's own community, from the wide record).
struct alignas(8) Lab8 {
    int32_t c;
    uint32_t deg;
};

// A partition entry of a narrow level (every weight below 2^32).
struct alignas(8) PEntry8 {
    int32_t c;
    uint32_t w;
};

// One (community, weight) entry of a hash partition: a single 16-byte store
// per entry when partitions are scattered.
struct alignas(16) PEntry {
    int32_t c;
    int32_t pad;
    int64_t w;
};

struct SweepArgs {
    const Lab *lab;         // label records; narrow levels also lab8 and deg32
    const Lab8 *lab8;
    const uint32_t *deg32;  // deg[] of narrow levels
    const int64_t *rowptr;  // row v is [rowptr[v], rowend[v]) of col/w (rowend = rowptr + 1
    const int64_t *rowend;  //   when the rows are packed; coarsened levels leave gaps)
    const int32_t *col;     // columns, or with cpk > 0 one word col << cpk | weight per
    const int64_t *w;       //   entry (w, w32 null); row entry weights (units), or w32
    const uint32_t *w32;    //   when every weight fits 32 bits
    int cpk;
    const int64_t *kdeg;
    const int64_t *loopw;
    const int64_t *selfsum;
    const int32_t *C;
    int32_t *next;
    const int64_t *deg;
    const int32_t *size;
    // Move deltas per community.  With size_shift > 0 the size delta rides in
    // the low size_shift bits of the degree word (ddeg = delta_deg * 2^size_shift
    // + delta_size, exact while both fit); with 0 it goes to dsize.  The
    // internal weights need no deltas: a mover adds to the joined community
    // what it takes from the one left, so their sum - all the modularity reads
    // - stays the level's self-loop total.
    ull *ddeg;
    int32_t *dsize;
    int size_shift;
    double m2;     // 2.0 * (double)m
    double gamma;
    int last;      // last pass: fuse merge-isolated, drop move deltas
    int32_t *pcnt; // timing only: each vertex's count of distinct candidate communities
    // Graph sets (batched kernels only): component of each vertex, and per
    // component its mode (0 frozen: next = C; 1 sweep; 2 last pass), 2m and
    // resolution.
    const int32_t *vcomp;
    const int32_t *cmode;
    const double *cm2;
    const double *cgam;
};

// Hub tier plan (rows with q > 2048): each hub's entries (plus a virtual
// own-community entry when it has a self-loop) are cut into chunks of
// chunk_span(P) entries (kHubChunk or more); every pass buckets them by a hash of their community into
// P_j = ceil(q_j / kHubPartTarget) partitions, aggregates each partition in an
// LDS table, then reduces the partitions of a hub.
constexpr int kHubChunk = 4096;
constexpr int kChunkMul = 32;  // entries per partition of a chunk (0: kHubChunk fixed)
constexpr int kHubPartTarget = 2048;  // default (lv_config / tuning table override)
constexpr int kHubMaxParts = 2048;
constexpr int kHubAggLog2 = 12;  // 4096 LDS slots per partition block, or
constexpr int kHubAggLog2Small = 11;  // 2048 (16 KiB) when the partition target is 1024 or less
constexpr int kHubAggSmallTarget = 1024;

struct HubArgs {
    const int32_t *hv;        // hub vertex of hub j
    const int64_t *hq;        // q_j
    const int32_t *hP;        // partitions of hub j
    const int64_t *ebase;     // first entry slot of hub j (exclusive scan of q)
    const int64_t *pbase;     // first partition of hub j (exclusive scan of P)
    const int32_t *part_hub;  // hub of each partition
    const int32_t *chunk_hub; // hub of each chunk
    const int32_t *chunk_k;   // chunk number within its hub
    int32_t *tmp_c;           // community of each hub entry (entry order)
    PEntry *buf;              // partitioned (community, weight) entries, or (narrow
    PEntry8 *buf8;            // levels) buf8 instead, or buf4: community << pwb |
    uint32_t *buf4;           // weight when both fit 32 bits (level 0, small weights)
    int pwb;
    int32_t *chist;           // per chunk its P partition counts, at hbase[j] + k * P_j
    const int64_t *hbase;     // exclusive scan over hubs of chunks * P
    int64_t *pcount;          // entries per partition: counted up by the count pass,
                              //   down to zero by the scatter's reservations
    const int64_t *poff;      // partition offsets (exclusive scan of pcount)
    int32_t *fbk;             // fallback global tables, 2 slots per entry
    ull *fbv;
    double *pg;               // per-partition best gain, community, weight, groups
    int32_t *pc;
    int64_t *pw;
    int32_t *pn;
    const int64_t *cbase;     // first chunk of hub j (exclusive scan of chunks)
    int chunk_mul;            // chunk_span(P, chunk_mul) entries per chunk
    int small_tables;         // partitions aggregate in 2048-slot tables
    double *cg;               // per-chunk best of the first-pass sweep
    int32_t *cc;
    int64_t *cw;
    int32_t *cn;
};

// Coarsening: community segments of the member-order entry list.
struct CoarsenArgs {
    const int32_t *ey;      // newc[target] of every entry, member order
    const int64_t *ew;      // w' of every entry, or (narrow levels) ew32 instead
    const uint32_t *ew32;
    const uint32_t *ep;     // or both in one word, newc << ewb | w' (level 0, when the
    int ewb;                //   next level's ids and the largest weight fit 32 bits)
    int n32;                // narrow level: 4-byte LDS sums in the segment tables
    const int64_t *sb;      // first entry of group g
    const int64_t *slen;    // entries of group g
    const int32_t *newid_g; // new id of group g
    int32_t *uy;            // uniques of segment g at sb[g]: row newid_g[g] of the next
    int64_t *uw;            //   level, whose rows keep the segments' places (weights in
    uint32_t *uw32;         //   uw, or uw32 on narrow levels)
    int32_t *nuniq;         // uniques per group
    ull *internal;          // sum of w' over entries inside the group
    ull *ssum;              // sum of w' over all entries of the group
    int64_t big_q;          // segments longer than this take the hash partitions
};

// Segments of more than 2048 entries, partitioned like the hub rows.
struct BigSegArgs {
    const int32_t *bsegs;   // group of big segment j
    const int32_t *bP;      // partitions of big segment j
    const int64_t *pbase;   // first partition of j
    const int32_t *part_big;
    const int32_t *chunk_big;
    const int32_t *chunk_k;
    int64_t *pcount;        // entries per partition (up in the count, down in the scatter)
    int64_t *poff;
    PEntry *pe;             // partitioned entries, or (narrow levels) pe8 instead, or
    PEntry8 *pe8;           // pe4: key << pwb | w' when both fit 32 bits (level 0,
    uint32_t *pe4;          // small weights)
    int pwb;
    int n32;                // narrow level: 4-byte LDS sums in the aggregation
    int32_t *chist;         // per chunk its P partition counts, at hbase[j] + k * P_j
    const int64_t *hbase;   // exclusive scan over big segments of chunks * P
    int32_t *fbk;           // fallback tables
    ull *fbv;
    int chunk_mul;          // chunk_span(P, chunk_mul) entries per chunk
    int small_tables;       // partitions aggregate in 2048-slot tables
};

// ---------------------------------------------------------------------------
// primitives.hip
int scan_exclusive_i64(Arena &a, hipStream_t s, const int64_t *in, int64_t *out, int64_t n,
                       int64_t *total_dev);
int scan_exclusive_i32(Arena &a, hipStream_t s, const int32_t *in, int32_t *out, int64_t n,
                       int32_t *total_dev);
// Stable LSD radix sort over bits [0, bits).  The sorted data ends in
// (*keys_out, *vals_out), which is either the input pair or the tmp pair.
int radix_sort_u64_i64(Arena &a, hipStream_t s, uint64_t *keys, int64_t *vals, uint64_t *ktmp,
                       int64_t *vtmp, int64_t n, int bits, uint64_t **keys_out,
                       int64_t **vals_out);
int radix_sort_u32_i32(Arena &a, hipStream_t s, uint32_t *keys, int32_t *vals, uint32_t *ktmp,
                       int32_t *vtmp, int64_t n, int bits, uint32_t **keys_out,
                       int32_t **vals_out);
// Sorts 64-bit keys by their bits [lo, 64), stably, without values.
int radix_sort_u64_keys(Arena &a, hipStream_t s, uint64_t *keys, uint64_t *ktmp, int64_t n,
                        int lo, uint64_t **keys_out);
int radix_sort_u32_u32(Arena &a, hipStream_t s, uint32_t *keys, uint32_t *vals, uint32_t *ktmp,
                       uint32_t *vtmp, int64_t n, int bits, uint32_t **keys_out,
                       uint32_t **vals_out);
int radix_sort_u32_u64(Arena &a, hipStream_t s, uint32_t *keys, uint64_t *vals, uint32_t *ktmp,
                       uint64_t *vtmp, int64_t n, int bits, uint32_t **keys_out,
                       uint64_t **vals_out);
void fill_i32(hipStream_t s, int32_t *p, int64_t n, int32_t v);
void fill_i64(hipStream_t s, int64_t *p, int64_t n, int64_t v);
// Small results to the host without a runtime copy: k_mailbox copies up to
// kMboxSegs device word ranges into a page-locked buffer and then sets its
// flag word to seq (system-scope release).
constexpr int kMboxSegs = 6;
struct MboxSeg {
    const uint32_t *src;
    uint32_t off;    // first word in box
    uint32_t words;
};
struct MboxArgs {
    MboxSeg seg[kMboxSegs];
    int nseg;
    uint32_t seq;
    uint32_t *box;   // device address of the data words
    uint32_t *flag;  // device address of the flag word
};
void launch_mailbox(hipStream_t s, const MboxArgs &a);
void iota_i32(hipStream_t s, int32_t *p, int64_t n);
void copy_i32(hipStream_t s, int32_t *dst, const int32_t *src, int64_t n);
void sum_i64(hipStream_t s, const int64_t *p, int64_t n, int64_t *out_dev);  // *out += sum
// *out += sum of p[idx[i]], i < n
void sum_gather_i32(hipStream_t s, const int32_t *idx, int64_t n, const int32_t *p, int64_t *out_dev);

int bits_for(uint64_t maxval);  // bits needed to hold values 0..maxval

// ---------------------------------------------------------------------------
// louvain_kernels.hip (host launchers)
// Input edge weights in units 2^-s: 4 bytes each when every weight fits (int32
// weights always do - they are their own units), else 8.
struct WUnits {
    const int64_t *w64 = nullptr;
    const uint32_t *w32 = nullptr;
    WUnits() = default;
    WUnits(const int64_t *a, const uint32_t *b) : w64(a), w32(b) {}
    __device__ __forceinline__ int64_t operator[](int64_t i) const
    {
        return w32 ? (int64_t)w32[i] : w64[i];
    }
};

struct DevFlags {  // error flags and small scalars written by kernels
    int32_t err;              // bit 1: invalid input, bit 2: out of exact range,
                              // bit 4: CSR rows not strictly increasing,
                              // bit 8: src not sorted, bit 16: edge list not
                              // canonical and strictly sorted
    int32_t scale;            // weight unit exponent s
    uint32_t max_bits;        // float32 weights: bit pattern of the largest (non-negative) one
    int32_t pad1;
    long long total_units;    // sum of (w == 0 ? 1 : w) in units
    long long m_units;        // sum of w in units
    unsigned long long max_units;  // largest weight in units
    long long pad2;
};

void launch_weight_exponent(hipStream_t s, const void *w, int wtype, int64_t e, DevFlags *f);
// Units into wu (8 bytes), wu32 (4 bytes) or neither (int32 weights kept as they are).
void launch_weight_units(hipStream_t s, const void *w, int wtype, int64_t e, int scale,
                         int64_t *wu, uint32_t *wu32, DevFlags *f);
void launch_check_endpoints(hipStream_t s, const int32_t *src, const int32_t *dst, int64_t e,
                            int32_t n, DevFlags *f);
// An uploaded edge list as lv_load_edges sent it: sources raw, or as the row
// offsets (n + 1) of a source-sorted list, decoded into src_out; weights as raw
// 32-bit patterns, or int weights packed wbits (2, 4 or 8) each - edge i at bit
// wbits * i - with the all-ones value marking an escape, whose weight is in an
// escape list sorted by edge.
struct LoadedEdges {
    const int32_t *src = nullptr;    // raw sources, or null
    const int32_t *off32 = nullptr;  // row offsets, 32-bit when e <= INT32_MAX
    const int64_t *off64 = nullptr;
    int32_t *tile_v = nullptr;       // first source of every decode tile (ntiles + 1)
    int32_t *src_out = nullptr;      // decoded sources
    const int32_t *dst = nullptr;
    const uint32_t *w = nullptr;     // raw weights, or null
    const uint8_t *wpk = nullptr;    // packed weights
    int wbits = 0;
    const int64_t *esc_i = nullptr;
    const uint32_t *esc_w = nullptr;
    int64_t nesc = 0;
};
constexpr int kDecTile = 4096;             // edges per decode tile
constexpr int64_t kEncodeMinEdges = 1 << 18;  // smaller lists are copied as they are
// Checks of an uploaded edge list (ids = positions): endpoints in [0, n) and the
// weight checks of k_weight_exponent, with int32 weights also the sums of
// k_weight_units (they are their own units); offsets decoded to sources, and
// weight bytes with their escapes to 32-bit units in wu32.
void launch_load_checks(hipStream_t s, const LoadedEdges &L, int wtype, int64_t e, int32_t n,
                        uint32_t *wu32, DevFlags *f);

// Host side of lv_load_edges' upload.  The copy of dst goes out first; host
// threads meanwhile pack int weights into 2, 4 or 8 bits (the width that a
// sample of the weights says sends the fewest bytes; an escape list holds the
// ones that do not fit) and turn a source-sorted src into row offsets, whose
// copies follow: 12 bytes per edge become 4.25 to 5 plus 4-8 per vertex.  What does not encode (unsorted
// or out-of-range sources, more escapes than esc_cap) goes out raw, so the
// device checks see the caller's values either way.
struct UploadBufs {  // device destinations
    int32_t *src = nullptr, *dst = nullptr;
    uint32_t *w = nullptr;
    void *off = nullptr;  // 8 * (n + 1) bytes
    uint8_t *wpk = nullptr;  // e bytes
    int64_t *esc_i = nullptr;
    uint32_t *esc_w = nullptr;
    int64_t esc_cap = 0;
};
struct UploadPlan {
    bool src_off = false, off64 = false;
    int wbits = 0;  // packed weight width, 0 when the weights went raw
    int64_t nesc = 0;
    int64_t bytes = 0;  // copied host to device
};
class HostPool;
class EdgeUploader {
  public:
    EdgeUploader() = default;
    EdgeUploader(const EdgeUploader &) = delete;
    EdgeUploader &operator=(const EdgeUploader &) = delete;
    ~EdgeUploader();
    static int64_t esc_cap(int64_t e) { return e / 32 + 1024; }
    int threads();  // host threads of the pool (created on first use)
    // Enqueues the copies on s; encode = false sends every array raw (from
    // pageable memory through the staging ring when staging is set, else with
    // plain copies, which allocate nothing).  The host buffers must stay
    // unchanged until the stream has finished the copies.
    int upload(hipStream_t s, int32_t n, int64_t e, const int32_t *src, const int32_t *dst,
               const void *w, int wtype, bool encode, bool staging, const UploadBufs &d,
               UploadPlan *plan);
    // Fills the device array dev from host pieces: piece i (host[i]) is bytes
    // [off[i], off[i+1]) of it.  The bytes go through two page-locked buffers
    // of kRingBytes: host threads fill one while the other is on the link (a
    // buffer is refilled once its copy has finished), so pageable sources move
    // at page-locked speed.  Returns with the last copies queued on s; the
    // host pieces may change once the call returns.
    int upload_concat(hipStream_t s, void *dev, int32_t np, const void *const *host,
                      const int64_t *off);
    static constexpr int64_t kRingBytes = 16 << 20;
    // Copies bytes from dev to host after the work queued on s and returns when
    // they are there: into pageable host memory of kRingMin bytes or more
    // through the staging ring (host threads copy each chunk out while the next
    // one is on the link), else directly.
    int download(hipStream_t s, void *host, const void *dev, size_t bytes);
    static constexpr int64_t kRingMin = 256 << 10;

  private:
    int stage(void **p, size_t *cap, size_t bytes);
    int ring();
    int raw(hipStream_t s, void *dev, const void *host, size_t bytes);
    HostPool *pool_ = nullptr;
    int pool_pid_ = 0;  // the process that started the pool
    void *h_wpk_ = nullptr, *h_off_ = nullptr, *h_esc_ = nullptr;
    size_t wpk_cap_ = 0, off_cap_ = 0, esc_bytes_ = 0;
    std::vector<std::vector<int64_t>> esc_parts_;
    void *h_ring_[2] = {nullptr, nullptr};
    hipEvent_t ring_ev_[2] = {nullptr, nullptr};  // the last copy out of each buffer
    bool stage_ok_ = true;
};
void launch_check_sorted_unique(hipStream_t s, const uint32_t *keys, int64_t n, DevFlags *f);
void launch_map_endpoints(hipStream_t s, const uint32_t *sorted_ids, const int32_t *sorted_pos,
                          int32_t n, int32_t *ep, int64_t e, DevFlags *f);
void launch_ids_to_keys(hipStream_t s, const int32_t *ids, uint32_t *keys, int32_t n);

void launch_csr_entries(hipStream_t s, const int32_t *src, const int32_t *dst, int64_t e,
                        int32_t n, uint32_t *keys, uint32_t *codes);
// Entries with the payload (col << wbits | w) when every weight fits in wbits
// bits and every column in 64 - wbits.
void launch_csr_entries_packed(hipStream_t s, const int32_t *src, const int32_t *dst,
                               WUnits w, int64_t e, int32_t n, int wbits, uint32_t *keys,
                               uint64_t *payload);
void launch_csr_unpack(hipStream_t s, const uint32_t *skeys, const uint64_t *spay, int64_t nnz,
                       int wbits, int32_t *col, int64_t *cw, uint32_t *cw32, int64_t *kdeg);
void launch_self_loops(hipStream_t s, const int32_t *src, const int32_t *dst, WUnits w,
                       int64_t e, int64_t *kdeg, int64_t *selfsum, ull *lastloop);
void launch_rowptr_search(hipStream_t s, const uint32_t *skeys, int64_t total, int32_t n,
                          int64_t *rowptr);
void launch_csr_gather(hipStream_t s, const uint32_t *skeys, const uint32_t *scodes, int64_t nnz,
                       const int32_t *src, const int32_t *dst, WUnits w, int32_t *col,
                       int64_t *cw, uint32_t *cw32, int64_t *kdeg);
void launch_loop_gather(hipStream_t s, const ull *lastloop, WUnits w, int32_t n,
                        int64_t *loopw);

void launch_init_vertex(hipStream_t s, int32_t n, const int64_t *rowptr, const int64_t *rowend,
                        const int64_t *kdeg, const int64_t *loopw, const int64_t *selfsum,
                        int64_t *deg, int64_t *in, int32_t *size, ull *ddeg, int32_t *dsize,
                        int32_t *C, int32_t *next, uint32_t *tier_key, int32_t *verts,
                        long long *tier_cnt, long long *tier_nnz, ull *ksum, uint32_t *deg32,
                        int32_t hub_q);
// Q aggregates: per-block partials in part[kQParts], summed into *acc.
constexpr int kQParts = 1024;
void launch_qagg(hipStream_t s, int32_t n, const int64_t *deg, const int64_t *in, QAcc *part,
                 QAcc *acc);
void launch_lab(hipStream_t s, int32_t n, const int32_t *C, const int32_t *size,
                const int64_t *deg, Lab *lab, Lab8 *lab8);
void launch_apply(hipStream_t s, int32_t n, int64_t *deg, const int64_t *in, int32_t *size,
                  ull *ddeg, int32_t *dsize, int size_shift, uint32_t *deg32, QAcc *part,
                  QAcc *acc);
// Move deltas of a pass that is not a last pass, from C and next (own_direct:
// the first pass of a level, where every community left is a singleton).
void launch_move_deltas(hipStream_t s, const SweepArgs &a, int32_t n, bool own_direct, bool batch);
void launch_sweep_tier(hipStream_t s, int tier, const SweepArgs &a, const int32_t *verts,
                       int32_t nv, bool batch);
// First pass of a level on rows that hold each neighbour once: every community
// is a single vertex, so no aggregation table is needed (tiers 1..9; hubs via
// run_hub_sweep1).
void launch_sweep1_tier(hipStream_t s, int tier, const SweepArgs &a, const int32_t *verts,
                        int32_t nv, bool batch);
int run_hub_sweep1(hipStream_t s, const SweepArgs &a, const HubArgs &h, int32_t nh,
                   int64_t nchunks, bool batch);
// CSR build for edges sorted by source (see louvain_kernels.hip).  Edges are
// processed in tiles of kEdgeTile; tself[t] (ntiles + 1 ints) and vself[v]
// (n + 1 ints) receive the self-loop edges of tile t and of source vertex v,
// and after in-place exclusive scans tscan/vscan are what the later kernels read.
constexpr int kEdgeTile = 2048;
void launch_src_sorted(hipStream_t s, const int32_t *src, const int32_t *dst, WUnits w,
                       int64_t e, int32_t n, int32_t *tself, int32_t *vself, int64_t *fs,
                       int64_t *kdeg, int64_t *selfsum, ull *lastloop, DevFlags *f);
void launch_rev_entries(hipStream_t s, const int32_t *src, const int32_t *dst, WUnits w,
                        const int32_t *tscan, int64_t e, int wbits, uint32_t *keys,
                        uint64_t *payload);
void launch_rowptr_sum(hipStream_t s, const int64_t *rs, const int64_t *fs, const int32_t *vscan,
                       int32_t n, int64_t *rowptr, int64_t *ffwd);
void launch_csr_sorted_fill(hipStream_t srev, hipStream_t sfwd, bool check, const int32_t *src,
                            const int32_t *dst, WUnits w, const int32_t *tscan,
                            const int32_t *vscan, int64_t e, int32_t n, int wbits, int vshift,
                            int kb, const int64_t *fs, const uint32_t *skey, const uint64_t *spay,
                            int64_t nr, const int64_t *rs, const int64_t *rowptr,
                            const int64_t *ffwd, int32_t *col, int64_t *cw, uint32_t *cw32, int pk,
                            int64_t *kdeg, DevFlags *f);
// One-word reverse entries dst << (64 - kb) | src << (64 - 2 kb) | w, and row
// starts over such words sorted by their top bits (key = word >> kshift).
// General path as one word per entry (row | col | w, row = n for self-loops);
// the unpack also checks that every row's columns strictly increase (err bit 4).
void launch_csr_entries1(hipStream_t s, const int32_t *src, const int32_t *dst, WUnits w,
                         int64_t e, int32_t n, int rb, int cb, uint64_t *words);
void launch_csr_unpack1(hipStream_t s, const uint64_t *words, int64_t nnz, int rb, int cb,
                        int32_t *col, int64_t *cw, uint32_t *cw32, int64_t *kdeg, DevFlags *f);
void launch_rev_entries1(hipStream_t s, const int32_t *src, const int32_t *dst, WUnits w,
                         const int32_t *tscan, int64_t e, int kb, uint64_t *words);
void launch_row_starts64(hipStream_t s, const uint64_t *keys, int64_t total, int32_t n,
                         int kshift, int64_t *rowptr);
// flags->err |= 4 when some row's columns are not strictly increasing.
void launch_rows_check(hipStream_t s, const uint32_t *skeys, const int32_t *col, int64_t nnz,
                       DevFlags *f);
void launch_hub_setup(hipStream_t s, const int32_t *hv, int32_t nh, const int64_t *rowptr,
                      const int64_t *rowend, const int64_t *selfsum, int32_t target,
                      int chunk_mul, int64_t *hq, int32_t *hP, int64_t *hP64, int64_t *hch,
                      int64_t *hpc);
// owner[i] = segment j of i under exclusive offsets base[0..nseg]; rank[i] =
// i - base[j] (rank may be null).
void launch_expand_owner(hipStream_t s, const int64_t *base, int32_t nseg, int64_t total,
                         int32_t *owner, int32_t *rank);
void launch_widen_i32(hipStream_t s, const int32_t *in, int64_t n, int64_t *out);
int run_hub_sweep(Arena &A, hipStream_t s, const SweepArgs &a, const HubArgs &h, int32_t nh,
                  int64_t nparts, int64_t nchunks, bool batch);

// batch.hip (graph sets)
void launch_b_graph_exponent(hipStream_t s, const uint32_t *bw, int wtype, const int64_t *beoff,
                             int32_t ngraphs, int64_t eb, int *gscale, DevFlags *f);
void launch_b_union_edges(hipStream_t s, const int32_t *bsrc, const int32_t *bdst,
                          const uint32_t *bw, int wtype, const int *gscale, const int64_t *beoff,
                          const int32_t *bn, const int64_t *ueoff, const int64_t *uvoff,
                          int32_t ncomp, int32_t nres, int64_t etot, int32_t *usrc, int32_t *udst,
                          int64_t *uw, DevFlags *f);
void launch_b_edge_sum(hipStream_t s, const int64_t *w, const int64_t *ueoff, int32_t ncomp,
                       int64_t etot, ull *cm);
// The union CSR of nres copies of the graphs' level-0 CSR (geoff: ngraphs + 1
// scratch entries); see batch.hip.
void launch_b_replicate_csr(hipStream_t s, const int64_t *bvoff, int64_t *geoff, int32_t ngraphs,
                            int32_t nres, int32_t nb, int64_t nnz, const int64_t *rowptr,
                            const int32_t *col, const int64_t *w, const uint32_t *w32,
                            const int64_t *kdeg, const int64_t *loopw, const int64_t *selfsum,
                            int64_t *urowptr, int32_t *ucol, int64_t *uw, uint32_t *uw32,
                            int64_t *ukdeg, int64_t *uloopw, int64_t *uselfsum);
void launch_b_qagg(hipStream_t s, int32_t n, int64_t *deg, const int64_t *in, int32_t *size,
                   ull *ddeg, int32_t *dsize, int size_shift, uint32_t *deg32,
                   const int32_t *vcomp, bool apply, QAcc *cq);
void launch_b_groups(hipStream_t s, const int32_t *gstart, const int32_t *perm,
                     const int32_t *vcomp, int32_t K, ull *kcnt);
void launch_b_next_ids(hipStream_t s, const int32_t *gstart, const int32_t *perm,
                       const int32_t *vcomp, const int32_t *newid_g, int32_t K,
                       const int64_t *koff, const int64_t *vnext, int32_t *newid_next);
void launch_b_vertex_sum(hipStream_t s, const int64_t *kdeg, const int64_t *loopw,
                         const int32_t *vcomp, int32_t n, ull *cm);
void launch_b_flatten(hipStream_t s, int32_t *fin, int32_t n0, const int32_t *vcomp0,
                      const int32_t *clev, int32_t l, const int32_t *newc_l, const int64_t *koff_l,
                      const int64_t *vnext_l);

// coarsen.hip
void launch_cr_heads(hipStream_t s, const uint32_t *Cs, int32_t n, int32_t *head);
void launch_cr_groups(hipStream_t s, const int32_t *head, const int32_t *hscan, const int32_t *perm,
                      int32_t n, int32_t *gstart, int32_t *lflag);
void launch_cr_newid(hipStream_t s, const int32_t *gstart, const int32_t *perm, const int32_t *rank,
                     int32_t K, int32_t *newid_g);
void launch_cr_newc(hipStream_t s, const int32_t *head, const int32_t *hscan, const int32_t *perm,
                    const int32_t *newid_g, int32_t n, int32_t *newc);
void launch_cr_memdeg(hipStream_t s, const int32_t *perm, const int64_t *rowptr,
                      const int64_t *rowend, int32_t n, int64_t *mdeg, int64_t *mstart);
void launch_cr_selfl(hipStream_t s, const int32_t *perm, const int64_t *selfsum,
                     const int32_t *head, const int32_t *hscan, int32_t n, ull *selfl);
void launch_cr_gather(hipStream_t s, const int64_t *estart, int32_t *bound, const int64_t *mstart,
                      const int32_t *col, const int64_t *cw, const uint32_t *cw32,
                      const int32_t *newc, int32_t n,
                      int64_t total, int32_t *ey, int64_t *ew, uint32_t *ew32, uint32_t *ep,
                      int ewb, int cpk);
void launch_cr_seg_tier(hipStream_t s, const int32_t *gstart, const int64_t *estart,
                        const int32_t *newid_g, int32_t K, int64_t *sb, int64_t *slen,
                        uint32_t *tkey, int32_t *sid, long long *tier_cnt, int32_t hub_q);
void launch_cr_agg_tier(hipStream_t s, int tier, const CoarsenArgs &a, const int32_t *segs, int32_t ns);
void launch_cr_big_setup(hipStream_t s, const int32_t *bsegs, int32_t nbs, const int64_t *slen,
                         int32_t target, int chunk_mul, int32_t *bP, int64_t *bP64, int64_t *bch,
                         int64_t *blen, int64_t *bpc, int32_t *nuniq);
int run_cr_big(Arena &A, hipStream_t s, const CoarsenArgs &a, const BigSegArgs &B, int32_t nbs,
               int64_t nparts, int64_t nchunks);
// The big segments' uniques from their hash partitions into their rows.
// Per next-level vertex: loop, selfsum, k and row bounds; msum += 2m, nsum += entries.
void launch_cr_vertex(hipStream_t s, const int32_t *newid_g, const ull *ssum, const ull *internal,
                      const ull *selfl, const int64_t *sb, const int32_t *nuniq, int32_t K,
                      int64_t *loop2, int64_t *self2, int64_t *kdeg2, int64_t *rowptr2,
                      int64_t *rowend2, long long *msum, long long *nsum);

void launch_compose(hipStream_t s, int32_t *f, int32_t n, const int32_t *map);
// skey/perm: cluster keys sorted stably and the positions they came from.
void launch_layout(hipStream_t s, const uint32_t *skey, const int32_t *perm, int32_t n,
                   int32_t ncomm, int32_t *first, int32_t *count, int32_t *index_in_cluster,
                   float *x, float *y, DevFlags *f);
void launch_gather_partition(hipStream_t s, const int32_t *perm, const int32_t *ids,
                             const int32_t *comm_by_pos, int32_t n, int32_t *node,
                             int32_t *comm);
void launch_final_q(hipStream_t s, const int32_t *src, const int32_t *dst, WUnits w,
                    int64_t e, const int32_t *comm, ull *lc, ull *dc);
void launch_qagg_ull(hipStream_t s, int32_t n, const ull *dc, const ull *lc, QAcc *part, QAcc *acc);

}  // namespace lv
