// api.hip - the C ABI of liblouvain_amd.so: context, phase entry points and
// the device-resident level loop.  See include/louvain_amd.h.
#include <algorithm>
#include <chrono>
#include <initializer_list>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dlfcn.h>
#include <mutex>
#include <string>
#include <new>
#include <vector>

#include "engine.hpp"

using namespace lv;

namespace {

enum State { ST_EMPTY = 0, ST_LOADED, ST_CSR, ST_INIT, ST_LEVEL_DONE, ST_RENUMBERED, ST_DONE };
enum Cat { CAT_CSR = 0, CAT_INIT, CAT_SWEEP, CAT_APPLY, CAT_RENUMBER, CAT_COARSEN, CAT_FLATTEN };

struct LevelRec {
    int32_t n = 0, k = 0, passes = 0;
    int64_t nnz = 0, m = 0;
    double q = 0, lower = 0;
    int32_t *perm = nullptr;    // positions sorted by (community, position) (device)
    int32_t *newc = nullptr;    // renumbered community of each position (device)
    // graph sets: first community id of each component (ncomp+1) and first
    // vertex of each component at the next level, -1 when it stops (ncomp)
    int64_t *koff = nullptr, *vnext = nullptr;
    std::vector<int64_t> koff_h, voff_h;  // host copies; voff_h = this level's vertex offsets
    std::vector<double> pass_q;
    std::vector<std::vector<int32_t>> trace;
};

// lv_run_partitions: caller host buffers that receive each level's Partition
// while the run goes on (copies on a second stream), and the device buffers the
// copies read from (kept until the run ends).
struct PartSink {
    int32_t max = 0;
    int32_t *const *node = nullptr;
    int32_t *const *comm = nullptr;
    std::vector<void *> temps;
};

struct PendingTime {
    int cat, tier;
    hipEvent_t a, b;
    double bytes;
    int64_t entries;
};

}  // namespace

// Extra streams a pass's independent sweep tiers are spread over.
constexpr int kFanStreams = 3;
// Fan-out pays from about this many entries in a phase: below it the fork and
// join events cost more than the overlap gains (device-resident runs, never /
// always: SBM 3000 2.36 / 3.13 ms, R-MAT 16 2.62 / 2.97, R-MAT 18 2.06 / 2.08,
// R-MAT 20 4.00 / 3.76, R-MAT 22 12.53 / 12.10).
constexpr int64_t kFanMinLoad = (int64_t)1 << 22;
// Mailbox of fetch_small: bytes, first data word, and how long the host polls
// before it waits through the runtime instead.
constexpr size_t kMboxBytes = 64 << 10;
constexpr int kMboxData = 16;
constexpr int kMboxSpinUs = 400;

struct lv_ctx {
    lv_config cfg;
    int device = 0;
    hipStream_t stream = nullptr;
    Arena arena;
    // small device/pinned scalars
    DevFlags *d_flags = nullptr, *h_flags = nullptr;
    QAcc *d_qacc = nullptr, *h_qacc = nullptr;
    QAcc *d_qpart = nullptr;  // per-block partials of the Q reductions
    long long *d_small = nullptr, *h_small = nullptr;  // 64 slots
    // small results fetched through k_mailbox: coherent page-locked words,
    // [0] the flag, data from word kMboxData (h_ and d_ are the host and
    // device addresses of the same memory)
    uint32_t *h_mbox = nullptr, *d_mbox = nullptr;
    uint32_t mbox_seq = 0;
    // input
    bool loaded = false;
    int32_t n0 = 0;
    int64_t e0 = 0;
    int scale = 0;
    int64_t m0 = 0;
    uint64_t max_units0 = 0;  // largest input weight in units
    int32_t *d_ids = nullptr;  // null when ids are positions
    int32_t *d_src = nullptr, *d_dst = nullptr;
    int64_t *d_w = nullptr;
    uint32_t *d_w32 = nullptr;  // instead of d_w when every weight's units fit 32 bits
    // current level
    int state = ST_EMPTY;
    int level = 0;
    int32_t n = 0;
    int64_t nnz = 0, m = 0;
    int64_t *rowptr = nullptr;
    int64_t *rowend_buf = nullptr;  // row ends of a level whose rows leave gaps (else rowptr + 1)
    int64_t span = 0;               // entries of col/w (nnz when the rows are packed)
    int32_t *col = nullptr;
    int64_t *cw = nullptr;
    uint32_t *cw32 = nullptr;  // instead of cw when every weight of the level is below 2^32
    int64_t *kdeg = nullptr, *loopw = nullptr, *selfsum = nullptr;
    int32_t *C = nullptr, *next = nullptr;
    Lab *lab = nullptr;
    Lab8 *lab8 = nullptr;        // narrow levels: 8-byte label records and 32-bit totals
    uint32_t *deg32 = nullptr;
    bool narrow = false;
    int cpk = 0;  // > 0: the level's CSR entries are col << cpk | weight (no cw/cw32)
    int64_t *deg = nullptr, *in = nullptr;
    int32_t *size = nullptr, *dsize = nullptr;
    ull *ddeg = nullptr;
    int size_shift = 0;  // see SweepArgs
    int32_t *verts = nullptr;
    int32_t *pcnt = nullptr;  // timing: distinct candidates of each vertex in the last pass
    int64_t tier_cnt[kTiers + 1] = {0}, tier_nnz[kTiers + 1] = {0}, tier_off[kTiers + 1] = {0};
    // every row of the level lists each neighbour once (coarsened levels always;
    // level 0 when the CSR rows are strictly increasing): the first pass of the
    // level can then skip aggregation
    bool rows_unique = false;
    // hub tier plan (rows with q > hub_q), see HubArgs
    HubArgs hub{};
    int32_t n_hubs = 0;
    int64_t hub_parts = 0, hub_chunks = 0, hub_entries = 0;
    // expected entries per hub partition; the LDS table of a partition holds 4096
    // distinct communities (more fall back to a global table)
    int32_t hub_part_target = kHubPartTarget;
    int chunk_mul = kChunkMul;  // chunk_span(P, chunk_mul)
    // rows and coarsening segments with more candidates take the hub partitions
    int32_t hub_q = kHubQDefault;
    // groups of the last renumbering (sorted order): head flags, their exclusive
    // scan, first sorted index of each group, new id of each group
    int32_t *g_head = nullptr, *g_hscan = nullptr, *g_start = nullptr, *g_newid = nullptr;
    // pass results of the level being processed
    double cur_q = 0, cur_lower = 0;
    int32_t cur_passes = 0;
    std::vector<double> cur_pass_q;
    std::vector<std::vector<int32_t>> cur_trace;
    // dendrogram
    std::vector<LevelRec> levels;
    int32_t *d_final = nullptr;
    bool final_valid = false;
    // graph set (lv_batch_load): ncomp = graphs x resolutions components of one
    // union graph; component c is graph c / nres at resolution c % nres
    bool batch = false;
    int32_t ncomp = 0, nres = 1;
    // the input holds the graphs once (ngraphs graphs, base_n0 vertices and
    // base_e0 edges, graph g's vertices from d_bvoff[g]); level 0 is built over
    // them and then written out for every resolution
    int32_t ngraphs = 0, base_n0 = 0;
    int64_t base_e0 = 0;
    int64_t *d_bvoff = nullptr;
    std::vector<int32_t> comp_n0;           // level-0 vertices per component
    std::vector<int64_t> comp_e0;           // input edges per component
    std::vector<double> comp_gamma;
    std::vector<int64_t> comp_m0;           // m (units) at level 0
    std::vector<int> comp_scale;            // unit exponent of each component's graph
    std::vector<int64_t> comp_m;            // m (units) at the current level
    std::vector<int64_t> comp_voff;         // current level: first vertex of each component
    std::vector<std::vector<lv_level_info>> comp_rec;  // per component, per level
    int32_t *vcomp = nullptr;               // component of each vertex, current level
    int32_t *d_vcomp0 = nullptr;            // component of each level-0 vertex
    int32_t *d_cmode = nullptr;
    double *d_cm2 = nullptr, *d_cgam = nullptr;
    QAcc *d_cq = nullptr;
    ull *d_ccnt = nullptr;
    QAcc *h_cq = nullptr;                   // pinned, ncomp
    ull *h_ccnt = nullptr;                  // pinned, ncomp
    int32_t h_cap = 0;                      // capacity of the pinned arrays
    // timing
    lv_stats stats;
    std::vector<hipEvent_t> ev_pool;
    std::vector<PendingTime> pending;
    double last_run_ms = 0;
    // arena allocations of the loaded input (a run starts from exactly these)
    std::map<size_t, size_t> load_mark;
    // lv_run_partitions
    PartSink *sink = nullptr;
    hipStream_t copy_stream = nullptr;
    std::vector<hipEvent_t> sink_ev;
    // the sweep tiers of a pass fanned out over extra streams (created on first
    // use; not in timing mode or on a borrowed stream)
    hipStream_t fan[kFanStreams] = {};
    hipEvent_t fan_fork = nullptr, fan_join[kFanStreams] = {};
    bool fan_used[kFanStreams] = {};
    // lv_run_static: the device scalars and the arena live in the static
    // scratch array, the stream is the caller's, and the host copies of the
    // scalars are these ordinary (pageable) members
    bool borrowed = false;
    // lv_load_edges: host threads and page-locked staging of the encoded upload
    EdgeUploader uploader;
    int64_t upload_bytes = 0;  // copied to the device by the last load
    DevFlags hs_flags{};
    QAcc hs_qacc{};
    long long hs_small[64] = {0};
};

// ---------------------------------------------------------------------------
// helpers

namespace {

bool timing(lv_ctx *c) { return (c->cfg.flags & LV_FLAG_TIMING) != 0; }

hipEvent_t ev_get(lv_ctx *c)
{
    if (c->ev_pool.empty()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    hipEvent_t e = c->ev_pool.back();
    c->ev_pool.pop_back();
    return e;
}

hipEvent_t t_begin(lv_ctx *c)
{
    if (!timing(c)) return nullptr;
    hipEvent_t e = ev_get(c);
    if (e) (void)hipEventRecord(e, c->stream);
    return e;
}

void t_end(lv_ctx *c, hipEvent_t a, int cat, int tier = -1, double bytes = 0, int64_t entries = 0)
{
    if (!a) return;
    hipEvent_t b = ev_get(c);
    if (!b) return;
    (void)hipEventRecord(b, c->stream);
    c->pending.push_back({cat, tier, a, b, bytes, entries});
}

// Call only after the stream has been synchronised.
void t_resolve(lv_ctx *c)
{
    for (auto &p : c->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) != hipSuccess) ms = 0.f;
        lv_stats &s = c->stats;
        switch (p.cat) {
        case CAT_CSR: s.csr_ms += ms; break;
        case CAT_INIT: s.init_ms += ms; break;
        case CAT_SWEEP:
            s.sweep_ms += ms;
            if (p.tier >= 1 && p.tier <= kTiers) {
                s.tier_ms[p.tier - 1] += ms;
                s.tier_launches[p.tier - 1] += 1;
                s.tier_bytes[p.tier - 1] += p.bytes;
            }
            s.sweep_bytes += p.bytes;
            s.sweep_entries += p.entries;
            break;
        case CAT_APPLY: s.apply_ms += ms; break;
        case CAT_RENUMBER: s.renumber_ms += ms; break;
        case CAT_COARSEN: s.coarsen_ms += ms; break;
        case CAT_FLATTEN: s.flatten_ms += ms; break;
        }
        c->ev_pool.push_back(p.a);
        c->ev_pool.push_back(p.b);
    }
    c->pending.clear();
}

int sync(lv_ctx *c)
{
    LV_HIP(hipStreamSynchronize(c->stream));
    if (timing(c)) t_resolve(c);
    return LV_OK;
}

// One host buffer filled from device memory by fetch_small.
struct HostSeg {
    void *host;
    const void *dev;
    size_t bytes;
};

template <typename T>
HostSeg seg(T *host, const T *dev, int64_t count)
{
    return HostSeg{host, dev, sizeof(T) * (size_t)count};
}

// Waits for the mailbox flag to read seq: polls for up to kMboxSpinUs, then
// leaves the wait (and any device error) to the runtime.
int mbox_wait(lv_ctx *c, uint32_t seq)
{
    const auto t0 = std::chrono::steady_clock::now();
    for (int i = 1;; ++i) {
        if (__atomic_load_n(c->h_mbox, __ATOMIC_ACQUIRE) == seq) return LV_OK;
        if ((i & 63) == 0 && std::chrono::steady_clock::now() - t0 >
                                 std::chrono::microseconds(kMboxSpinUs))
            break;
        __builtin_ia32_pause();
    }
    LV_HIP(hipStreamSynchronize(c->stream));
    if (__atomic_load_n(c->h_mbox, __ATOMIC_ACQUIRE) != seq) return LV_ERR_DEVICE;
    return LV_OK;
}

// Small device values to host buffers after everything queued on the context
// stream, returning when they are there (the stream has then drained up to
// the copy).  A mailbox kernel writes them into coherent page-locked memory
// and raises a flag the host polls: a round trip of about 9 us against 14 us
// for a runtime copy and a stream synchronisation.  Timing mode (its events
// are read at the synchronisation), a borrowed stream and larger requests take
// the runtime copy.
int fetch_small(lv_ctx *c, std::initializer_list<HostSeg> segs)
{
    size_t total = 0;
    bool words = segs.size() <= (size_t)kMboxSegs;
    for (const HostSeg &g : segs) {
        total += g.bytes;
        words = words && g.bytes % 4 == 0 && ((uintptr_t)g.dev & 3) == 0;
    }
    if (timing(c) || c->borrowed || !c->h_mbox || !words ||
        total > kMboxBytes - 4 * (size_t)kMboxData) {
        for (const HostSeg &g : segs)
            if (g.bytes)
                LV_HIP(hipMemcpyAsync(g.host, g.dev, g.bytes, hipMemcpyDeviceToHost, c->stream));
        return sync(c);
    }
    MboxArgs a{};
    uint32_t off = 0;
    for (const HostSeg &g : segs) {
        a.seg[a.nseg++] = MboxSeg{(const uint32_t *)g.dev, off, (uint32_t)(g.bytes / 4)};
        off += (uint32_t)(g.bytes / 4);
    }
    a.seq = ++c->mbox_seq;
    a.box = c->d_mbox + kMboxData;
    a.flag = c->d_mbox;
    launch_mailbox(c->stream, a);
    LV_HIP(hipGetLastError());
    LV_TRY(mbox_wait(c, a.seq));
    const char *box = (const char *)(c->h_mbox + kMboxData);
    for (const HostSeg &g : segs) {
        memcpy(g.host, box, g.bytes);
        box += g.bytes;
    }
    return LV_OK;
}

// Returns when everything queued on the context stream has run (host buffers
// that queued copies read may then go), through the mailbox like fetch_small.
int barrier(lv_ctx *c) { return fetch_small(c, {}); }

// Timing: the candidate counts P of a pass's tiers (the 8*P term of the
// local-moving bytes) are summed per tier into d_small[kPSlot + t] and fetched
// with the pass's other results.
constexpr int kPSlot = 40;

void p_begin(lv_ctx *c)
{
    if (timing(c)) fill_i64(c->stream, (int64_t *)c->d_small + kPSlot, kTiers + 1, 0);
}

void p_tier(lv_ctx *c, int t, int64_t nv)
{
    if (timing(c) && c->pcnt)
        sum_gather_i32(c->stream, c->verts + c->tier_off[t], nv, c->pcnt,
                       (int64_t *)c->d_small + kPSlot + t);
}

int p_fetch(lv_ctx *c);  // below d2h

// After the pass's sync: P into the tier byte counts.
void p_resolve(lv_ctx *c)
{
    if (!timing(c)) return;
    for (int t = 1; t <= kTiers; ++t) {
        const double p = (double)c->h_small[kPSlot + t];
        c->stats.tier_cand[t - 1] += p;
        c->stats.tier_bytes[t - 1] += 8.0 * p;
        c->stats.sweep_bytes += 8.0 * p;
    }
}

template <typename T>
int d2h(lv_ctx *c, T *host, const T *dev, int64_t count)
{
    if (count <= 0) return LV_OK;
    LV_HIP(hipMemcpyAsync(host, dev, sizeof(T) * (size_t)count, hipMemcpyDeviceToHost, c->stream));
    return LV_OK;
}

template <typename T>
int h2d(lv_ctx *c, T *dev, const T *host, int64_t count)
{
    if (count <= 0) return LV_OK;
    LV_HIP(hipMemcpyAsync(dev, host, sizeof(T) * (size_t)count, hipMemcpyHostToDevice, c->stream));
    return LV_OK;
}

// A result array to the caller: after the work queued so far, through the
// staging ring when the caller's buffer is pageable (a context that may not
// allocate copies it directly).  Returns when the values are there.
int fetch(lv_ctx *c, int32_t *host, const int32_t *dev, int64_t count)
{
    if (count <= 0) return sync(c);
    if (c->borrowed) {
        LV_TRY(d2h(c, host, dev, count));
        return sync(c);
    }
    return c->uploader.download(c->stream, host, dev, sizeof(int32_t) * (size_t)count);
}

int p_fetch(lv_ctx *c)
{
    if (!timing(c)) return LV_OK;
    return d2h(c, c->h_small + kPSlot, c->d_small + kPSlot, kTiers + 1);
}

int zero_small(lv_ctx *c, int slots)
{
    fill_i64(c->stream, (int64_t *)c->d_small, slots, 0);
    return LV_OK;
}

// Community state of the sweeps (not needed once the level is renumbered).
void free_sweep_state(lv_ctx *c)
{
    Arena &a = c->arena;
    arena_free(a, c->C);
    arena_free(a, c->next);
    arena_free(a, c->lab);
    arena_free(a, c->lab8);
    arena_free(a, c->deg32);  // (narrow stays: the level's coarsening reads it)
    arena_free(a, c->deg);
    arena_free(a, c->in);
    arena_free(a, c->size);
    arena_free(a, c->dsize);
    arena_free(a, c->ddeg);
    arena_free(a, c->verts);
    arena_free(a, c->pcnt);
    HubArgs &h = c->hub;
    a.release((void *)h.hq);
    a.release((void *)h.hP);
    a.release((void *)h.ebase);
    a.release((void *)h.pbase);
    a.release((void *)h.part_hub);
    a.release((void *)h.chunk_hub);
    a.release((void *)h.chunk_k);
    a.release((void *)h.tmp_c);
    a.release((void *)h.buf);
    a.release((void *)h.buf8);
    a.release((void *)h.buf4);
    a.release((void *)h.chist);
    a.release((void *)h.hbase);
    a.release((void *)h.pcount);
    a.release((void *)h.poff);
    a.release((void *)h.fbk);
    a.release((void *)h.fbv);
    a.release((void *)h.pg);
    a.release((void *)h.pc);
    a.release((void *)h.pw);
    a.release((void *)h.pn);
    a.release((void *)h.cbase);
    a.release((void *)h.cg);
    a.release((void *)h.cc);
    a.release((void *)h.cw);
    a.release((void *)h.cn);
    c->hub = HubArgs{};
    c->n_hubs = 0;
    c->hub_parts = c->hub_chunks = c->hub_entries = 0;
}

void free_level_arrays(lv_ctx *c)
{
    Arena &a = c->arena;
    free_sweep_state(c);
    arena_free(a, c->rowptr);
    arena_free(a, c->rowend_buf);
    arena_free(a, c->col);
    arena_free(a, c->cw);
    arena_free(a, c->cw32);
    c->cpk = 0;
    arena_free(a, c->kdeg);
    arena_free(a, c->loopw);
    arena_free(a, c->selfsum);
    arena_free(a, c->g_head);
    arena_free(a, c->g_hscan);
    arena_free(a, c->g_start);
    arena_free(a, c->g_newid);
    arena_free(a, c->vcomp);
}

void free_dendrogram(lv_ctx *c)
{
    for (auto &L : c->levels) {
        arena_free(c->arena, L.perm);
        arena_free(c->arena, L.newc);
        arena_free(c->arena, L.koff);
        arena_free(c->arena, L.vnext);
    }
    c->levels.clear();
    arena_free(c->arena, c->d_final);
    c->final_valid = false;
}

void free_input(lv_ctx *c)
{
    arena_free(c->arena, c->d_bvoff);
    arena_free(c->arena, c->d_ids);
    arena_free(c->arena, c->d_src);
    arena_free(c->arena, c->d_dst);
    arena_free(c->arena, c->d_w);
    arena_free(c->arena, c->d_w32);
    c->loaded = false;
    c->batch = false;
    c->ncomp = 0;
    c->comp_n0.clear();
    c->comp_e0.clear();
    c->comp_gamma.clear();
    c->comp_m0.clear();
    c->comp_scale.clear();
}

void free_batch_run(lv_ctx *c)
{
    Arena &a = c->arena;
    arena_free(a, c->d_vcomp0);
    arena_free(a, c->d_cmode);
    arena_free(a, c->d_cm2);
    arena_free(a, c->d_cgam);
    arena_free(a, c->d_cq);
    arena_free(a, c->d_ccnt);
    c->comp_rec.clear();
}

// Resets per-run state to "input loaded, level 0".
void reset_run(lv_ctx *c)
{
    free_level_arrays(c);
    free_dendrogram(c);
    free_batch_run(c);
    if (c->loaded) c->arena.restore(c->load_mark);
    c->cur_pass_q.clear();
    c->cur_trace.clear();
    c->level = 0;
    c->state = c->loaded ? ST_LOADED : ST_EMPTY;
}

double current_q(lv_ctx *c, int &rc)
{
    rc = LV_OK;
    launch_qagg(c->stream, c->n, c->deg, c->in, c->d_qpart, c->d_qacc);
    rc = fetch_small(c, {seg(c->h_qacc, c->d_qacc, 1)});
    if (rc != LV_OK) return 0.0;
    return q_from_aggregates(c->h_qacc->sum_in, c->h_qacc->d2_lo, c->h_qacc->d2_hi, c->m,
                             c->cfg.gamma);
}

// ---------------------------------------------------------------------------
// phases

// Units of a non-negative float32 bit pattern at exponent scale (as
// weight_units on the device; saturates past 2^62).
uint64_t f32_units_host(uint32_t x, int scale)
{
    if ((x & 0x7fffffffu) == 0) return 0;
    const uint32_t ex = (x >> 23) & 0xffu, mant = x & 0x7fffffu;
    const uint64_t M = ex ? (mant | 0x800000u) : mant;
    const int sh = (ex ? (int)ex - 150 : -149) + scale;
    if (sh > 38) return 1ull << 62;
    return sh >= 0 ? (M << sh) : (M >> (-sh));
}

// End of every row of the current level.
const int64_t *row_end(const lv_ctx *c) { return c->rowend_buf ? c->rowend_buf : c->rowptr + 1; }

// The input weights in units (4 or 8 bytes each).
WUnits wunits(const lv_ctx *c) { return WUnits(c->d_w, c->d_w32); }

// Row-entry weights of a level: 4 bytes each when every weight fits 32 bits.
int alloc_csr_w(lv_ctx *c, int64_t nnz, bool w32)
{
    if (w32) return arena_alloc(c->arena, c->cw32, nnz);
    return arena_alloc(c->arena, c->cw, nnz);
}

// Level-0 CSR for edges sorted by source (tself/vself = self-loop edges per edge
// tile and per source vertex, fs[v] = first edge with src >= v): forward
// entries stay in edge order, reverse entries are radix-sorted by destination,
// row v = its reverse entries then its forward entries.  A canonical, strictly
// sorted list (canonical) gives rows that list each neighbour once in
// increasing order; otherwise the fills check.
// Bits that hold every row weight and entry weight w' of level 0 (at most the
// largest input weight in units, and at least 1): the width of the weight
// field of a packed CSR or partition entry.
int pack_bits(const lv_ctx *c)
{
    return bits_for(c->max_units0 > 1 ? c->max_units0 : 1);
}

int fan_assign(lv_ctx *c, int n, const int64_t *load, int64_t main_load, hipStream_t *out);
int fan_join(lv_ctx *c);

int build_level0_csr_sorted(lv_ctx *c, int32_t *tself, int32_t *vself, const int64_t *fs,
                            bool canonical, int wbits, const ull *lastloop)
{
    Arena &A = c->arena;
    hipStream_t s = c->stream;
    const int32_t n = c->n0;
    const int64_t e = c->e0;
    const int64_t ntiles = (e + kEdgeTile - 1) / kEdgeTile;
    // in place: tscan[t] = self-loop edges before tile t, vscan[v] = those with src < v
    int32_t *tscan = tself, *vscan = vself;
    LV_TRY(scan_exclusive_i32(A, s, tself, tscan, ntiles, tscan + ntiles));
    LV_TRY(scan_exclusive_i32(A, s, vself, vscan, n, vscan + n));
    int32_t nself = 0;
    LV_TRY(fetch_small(c, {seg(&nself, (const int32_t *)(tscan + ntiles), 1)}));
    const int64_t nr = e - nself;
    const int kb = bits_for(n > 0 ? (uint64_t)(n - 1) : 0);
    // one word per reverse entry (dst | src | w) when the weight fits the bits
    // that two ids leave: the row sort then moves 8 bytes per entry, not 12
    const int wb1 = 64 - 2 * kb;
    const bool one_word = kb >= 1 && wb1 >= 1 && c->max_units0 < (1ull << wb1);
    int64_t *rs = nullptr, *ffwd = nullptr;
    uint32_t *k1 = nullptr, *k2 = nullptr;
    uint64_t *p1 = nullptr, *p2 = nullptr;
    LV_TRY(arena_alloc(A, rs, (int64_t)n + 1));
    LV_TRY(arena_alloc(A, ffwd, (int64_t)n + 1));
    if (!one_word) {
        LV_TRY(arena_alloc(A, k1, nr));
        LV_TRY(arena_alloc(A, k2, nr));
    }
    LV_TRY(arena_alloc(A, p1, nr));
    LV_TRY(arena_alloc(A, p2, nr));
    uint32_t *sk = nullptr;
    uint64_t *sp = nullptr;
    int fbits = wbits, vshift = 0;
    if (one_word) {
        launch_rev_entries1(s, c->d_src, c->d_dst, wunits(c), tscan, e, kb, p1);
        LV_TRY(radix_sort_u64_keys(A, s, p1, p2, nr, 64 - kb, &sp));
        launch_row_starts64(s, sp, nr, n, 64 - kb, rs);
        fbits = wb1;
        vshift = 64 - kb;
    } else {
        launch_rev_entries(s, c->d_src, c->d_dst, wunits(c), tscan, e, wbits, k1, p1);
        LV_TRY(radix_sort_u32_u64(A, s, k1, p1, k2, p2, nr, kb, &sk, &sp));
        launch_rowptr_search(s, sk, nr, n, rs);
    }
    c->nnz = 2 * nr;
    // one word per entry, col << pk | weight, when the ids and the largest
    // weight fit 32 bits: every pass over the rows reads half the bytes
    // (graph sets keep separate columns and weights)
    const int pk = (!c->batch && pack_bits(c) < 32 && kb + pack_bits(c) <= 32) ? pack_bits(c) : 0;
    LV_TRY(arena_alloc(A, c->col, c->nnz));
    if (!pk) LV_TRY(alloc_csr_w(c, c->nnz, c->max_units0 < (1ull << 32)));
    c->cpk = pk;
    LV_HIP(hipMemsetAsync(&c->d_flags->err, 0, sizeof(int), s));
    launch_rowptr_sum(s, rs, fs, vscan, n, c->rowptr, ffwd);
    {
        // the two fills may overlap (disjoint entries, atomic row sums)
        const int64_t load[2] = {nr, e};
        hipStream_t st[2];
        LV_TRY(fan_assign(c, 2, load, 0, st));
        launch_csr_sorted_fill(st[0], st[1], !canonical, c->d_src, c->d_dst, wunits(c), tscan,
                               vscan, e, n, fbits, vshift, kb, fs, sk, sp, nr, rs, c->rowptr, ffwd,
                               c->col, c->cw, c->cw32, pk, c->kdeg, c->d_flags);
    }
    launch_loop_gather(s, lastloop, wunits(c), n, c->loopw);
    LV_TRY(fan_join(c));
    LV_HIP(hipGetLastError());
    if (canonical) {
        c->rows_unique = true;
    } else {
        LV_TRY(fetch_small(c, {seg(c->h_flags, c->d_flags, 1)}));
        c->rows_unique = (c->h_flags->err & 4) == 0;
    }
    arena_free(A, k1);
    arena_free(A, k2);
    arena_free(A, p1);
    arena_free(A, p2);
    arena_free(A, rs);
    arena_free(A, ffwd);
    return LV_OK;
}

// CalculateNeighbour at level 0: symmetric CSR of the loaded edge list.  Each
// edge gives two entries keyed by row (self-loops keyed past the last row); a
// stable radix sort by row keeps input order within a row.  When the columns
// and weights fit in 64 bits together the entry carries (col, weight) through the sort and is
// unpacked in one streaming pass; otherwise it carries its edge code and the
// columns and weights are gathered from the edge list afterwards.
int build_level0_csr(lv_ctx *c)
{
    Arena &A = c->arena;
    hipStream_t s = c->stream;
    c->cpk = 0;  // (the sorted path may pack the entries)
    const int32_t n = c->n0;
    const int64_t e = c->e0;
    const int64_t total = 2 * e;
    const int bits = bits_for((uint64_t)n);
    // (col, weight) share 64 bits when the widest column and weight fit together
    const int wbits = std::min(63, 64 - bits_for(n > 0 ? (uint64_t)(n - 1) : 0));
    const bool packed = c->max_units0 < (1ull << wbits);
    hipEvent_t t0 = t_begin(c);
    ull *lastloop = nullptr;
    uint32_t *k1 = nullptr, *k2 = nullptr, *v1 = nullptr, *v2 = nullptr;
    uint64_t *p1 = nullptr, *p2 = nullptr;
    LV_TRY(arena_alloc(A, c->rowptr, (int64_t)n + 1));
    LV_TRY(arena_alloc(A, c->kdeg, n));
    LV_TRY(arena_alloc(A, c->selfsum, n));
    LV_TRY(arena_alloc(A, c->loopw, n));
    LV_TRY(arena_alloc(A, lastloop, n));
    fill_i64(s, c->kdeg, n, 0);
    fill_i64(s, c->selfsum, n, 0);
    fill_i64(s, (int64_t *)lastloop, n, 0);
    if (packed && e < INT32_MAX) {
        // edges sorted by source: only the reverse entries need the row sort
        // (the same pass over the edges accumulates the self-loops)
        int32_t *tself = nullptr, *vself = nullptr;
        int64_t *fs = nullptr;
        LV_TRY(arena_alloc(A, tself, (e + kEdgeTile - 1) / kEdgeTile + 1));
        LV_TRY(arena_alloc(A, vself, (int64_t)n + 1));
        LV_TRY(arena_alloc(A, fs, (int64_t)n + 1));
        LV_HIP(hipMemsetAsync(&c->d_flags->err, 0, sizeof(int), s));
        launch_src_sorted(s, c->d_src, c->d_dst, wunits(c), e, n, tself, vself, fs, c->kdeg,
                          c->selfsum, lastloop, c->d_flags);
        LV_HIP(hipGetLastError());
        LV_TRY(fetch_small(c, {seg(c->h_flags, c->d_flags, 1)}));
        const int err = c->h_flags->err;
        if ((err & 8) == 0) {
            const int rc =
                build_level0_csr_sorted(c, tself, vself, fs, (err & 16) == 0, wbits, lastloop);
            arena_free(A, tself);
            arena_free(A, vself);
            arena_free(A, fs);
            arena_free(A, lastloop);
            if (rc != LV_OK) return rc;
            t_end(c, t0, CAT_CSR);
            c->n = n;
            c->m = c->m0;
            return LV_OK;
        }
        arena_free(A, tself);
        arena_free(A, vself);
        arena_free(A, fs);
    } else {
        launch_self_loops(s, c->d_src, c->d_dst, wunits(c), e, c->kdeg, c->selfsum, lastloop);
    }
    // one word per entry (row | col | w) when the weight fits the bits the row
    // key (self-loops keyed n) and the column leave: 8 bytes per sorted entry
    const int cb = bits_for(n > 0 ? (uint64_t)(n - 1) : 0);
    const int wb1 = 64 - bits - cb;
    const bool one_word = bits >= 1 && wb1 >= 1 && c->max_units0 < (1ull << wb1);
    if (one_word) {
        uint64_t *x1 = nullptr, *x2 = nullptr, *sx = nullptr;
        LV_TRY(arena_alloc(A, x1, total));
        LV_TRY(arena_alloc(A, x2, total));
        launch_csr_entries1(s, c->d_src, c->d_dst, wunits(c), e, n, bits, cb, x1);
        LV_TRY(radix_sort_u64_keys(A, s, x1, x2, total, 64 - bits, &sx));
        launch_row_starts64(s, sx, total, n, 64 - bits, c->rowptr);
        LV_TRY(fetch_small(c, {seg(&c->h_small[0], (const long long *)(c->rowptr + n), 1)}));
        c->nnz = c->h_small[0];
        LV_TRY(arena_alloc(A, c->col, c->nnz));
        LV_TRY(alloc_csr_w(c, c->nnz, c->max_units0 < (1ull << 32)));
        LV_HIP(hipMemsetAsync(&c->d_flags->err, 0, sizeof(int), s));
        launch_csr_unpack1(s, sx, c->nnz, bits, cb, c->col, c->cw, c->cw32, c->kdeg, c->d_flags);
        launch_loop_gather(s, lastloop, wunits(c), n, c->loopw);
        LV_HIP(hipGetLastError());
        LV_TRY(fetch_small(c, {seg(c->h_flags, c->d_flags, 1)}));
        c->rows_unique = (c->h_flags->err & 4) == 0;
        t_end(c, t0, CAT_CSR);
        arena_free(A, lastloop);
        arena_free(A, x1);
        arena_free(A, x2);
        c->n = n;
        c->m = c->m0;
        return LV_OK;
    }
    LV_TRY(arena_alloc(A, k1, total));
    LV_TRY(arena_alloc(A, k2, total));
    if (packed) {
        LV_TRY(arena_alloc(A, p1, total));
        LV_TRY(arena_alloc(A, p2, total));
    } else {
        LV_TRY(arena_alloc(A, v1, total));
        LV_TRY(arena_alloc(A, v2, total));
    }
    uint32_t *sk = nullptr, *sv = nullptr;
    uint64_t *sp = nullptr;
    if (packed) {
        launch_csr_entries_packed(s, c->d_src, c->d_dst, wunits(c), e, n, wbits, k1, p1);
        LV_TRY(radix_sort_u32_u64(A, s, k1, p1, k2, p2, total, bits, &sk, &sp));
    } else {
        launch_csr_entries(s, c->d_src, c->d_dst, e, n, k1, v1);
        LV_TRY(radix_sort_u32_u32(A, s, k1, v1, k2, v2, total, bits, &sk, &sv));
    }
    launch_rowptr_search(s, sk, total, n, c->rowptr);
    LV_TRY(fetch_small(c, {seg(&c->h_small[0], (const long long *)(c->rowptr + n), 1)}));
    c->nnz = c->h_small[0];
    LV_TRY(arena_alloc(A, c->col, c->nnz));
    LV_TRY(alloc_csr_w(c, c->nnz, c->max_units0 < (1ull << 32)));
    if (packed)
        launch_csr_unpack(s, sk, sp, c->nnz, wbits, c->col, c->cw, c->cw32, c->kdeg);
    else
        launch_csr_gather(s, sk, sv, c->nnz, c->d_src, c->d_dst, wunits(c), c->col, c->cw, c->cw32,
                          c->kdeg);
    launch_loop_gather(s, lastloop, wunits(c), n, c->loopw);
    LV_HIP(hipMemsetAsync(&c->d_flags->err, 0, sizeof(int), s));
    launch_rows_check(s, sk, c->col, c->nnz, c->d_flags);
    LV_HIP(hipGetLastError());
    LV_TRY(fetch_small(c, {seg(c->h_flags, c->d_flags, 1)}));
    c->rows_unique = (c->h_flags->err & 4) == 0;
    t_end(c, t0, CAT_CSR);
    arena_free(A, lastloop);
    arena_free(A, k1);
    arena_free(A, k2);
    if (packed) {
        arena_free(A, p1);
        arena_free(A, p2);
    } else {
        arena_free(A, v1);
        arena_free(A, v2);
    }
    c->n = n;
    c->m = c->m0;
    return LV_OK;
}

template <typename T>
int alloc_into(Arena &A, const T *&p, int64_t count)
{
    T *q = nullptr;
    LV_TRY(arena_alloc(A, q, count));
    p = q;
    return LV_OK;
}

int plan_hubs(lv_ctx *c, int32_t nh)
{
    Arena &A = c->arena;
    hipStream_t s = c->stream;
    HubArgs &h = c->hub;
    const int32_t *hv = c->verts + c->tier_off[kTiers];
    int64_t *hq = nullptr, *hP64 = nullptr, *hch = nullptr, *hpc = nullptr, *ebase = nullptr,
            *pbase = nullptr, *cbase = nullptr, *hbase = nullptr;
    int32_t *hP = nullptr;
    LV_TRY(arena_alloc(A, hq, nh));
    LV_TRY(arena_alloc(A, hP, nh));
    LV_TRY(arena_alloc(A, hP64, nh));
    LV_TRY(arena_alloc(A, hch, nh));
    LV_TRY(arena_alloc(A, hpc, nh));
    LV_TRY(arena_alloc(A, ebase, (int64_t)nh + 1));
    LV_TRY(arena_alloc(A, pbase, (int64_t)nh + 1));
    LV_TRY(arena_alloc(A, cbase, (int64_t)nh + 1));
    LV_TRY(arena_alloc(A, hbase, (int64_t)nh + 1));
    launch_hub_setup(s, hv, nh, c->rowptr, row_end(c), c->selfsum, c->hub_part_target, c->chunk_mul, hq, hP,
                     hP64, hch, hpc);
    h.chunk_mul = c->chunk_mul;
    h.small_tables = c->hub_part_target <= kHubAggSmallTarget ? 1 : 0;
    LV_TRY(scan_exclusive_i64(A, s, hq, ebase, nh, ebase + nh));
    LV_TRY(scan_exclusive_i64(A, s, hP64, pbase, nh, pbase + nh));
    LV_TRY(scan_exclusive_i64(A, s, hch, cbase, nh, cbase + nh));
    LV_TRY(scan_exclusive_i64(A, s, hpc, hbase, nh, hbase + nh));
    LV_TRY(fetch_small(c, {seg(c->h_small, (const long long *)(ebase + nh), 1),
                           seg(c->h_small + 1, (const long long *)(pbase + nh), 1),
                           seg(c->h_small + 2, (const long long *)(cbase + nh), 1),
                           seg(c->h_small + 3, (const long long *)(hbase + nh), 1)}));
    c->hub_entries = c->h_small[0];
    c->hub_parts = c->h_small[1];
    c->hub_chunks = c->h_small[2];
    const int64_t nhist = c->h_small[3];
    c->n_hubs = nh;
    h.hv = hv;
    h.hq = hq;
    h.hP = hP;
    h.ebase = ebase;
    h.pbase = pbase;
    int32_t *part_hub = nullptr, *chunk_hub = nullptr, *chunk_k = nullptr;
    LV_TRY(arena_alloc(A, part_hub, c->hub_parts));
    LV_TRY(arena_alloc(A, chunk_hub, c->hub_chunks));
    LV_TRY(arena_alloc(A, chunk_k, c->hub_chunks));
    launch_expand_owner(s, pbase, nh, c->hub_parts, part_hub, nullptr);
    launch_expand_owner(s, cbase, nh, c->hub_chunks, chunk_hub, chunk_k);
    h.part_hub = part_hub;
    h.chunk_hub = chunk_hub;
    h.chunk_k = chunk_k;
    h.cbase = cbase;
    arena_free(A, hP64);
    arena_free(A, hch);
    arena_free(A, hpc);
    const int64_t E = c->hub_entries, P = c->hub_parts;
    h.hbase = hbase;
    LV_TRY(arena_alloc(A, h.chist, nhist));
    LV_TRY(arena_alloc(A, h.tmp_c, E));
    h.buf4 = nullptr;
    h.buf8 = nullptr;
    h.buf = nullptr;
    h.pwb = 0;
    if (c->level == 0 && pack_bits(c) < 32 && bits_for((uint64_t)c->n - 1) + pack_bits(c) <= 32) {
        // a community id and a row weight of level 0 (at most the largest input
        // weight) in one word
        h.pwb = pack_bits(c);
        LV_TRY(arena_alloc(A, h.buf4, E));
    } else if (c->narrow) {  // a partition's weights sum below the level's total degree, < 2^32
        LV_TRY(arena_alloc(A, h.buf8, E));
    } else {
        LV_TRY(arena_alloc(A, h.buf, E));
    }
    LV_TRY(arena_alloc(A, h.fbk, 2 * E));
    LV_TRY(arena_alloc(A, h.fbv, 2 * E));
    LV_TRY(arena_alloc(A, h.pcount, P + 1));
    fill_i64(c->stream, h.pcount, P + 1, 0);  // every later pass leaves it at zero
    LV_TRY(alloc_into(A, h.poff, P + 1));
    LV_TRY(arena_alloc(A, h.pg, P));
    LV_TRY(arena_alloc(A, h.pc, P));
    LV_TRY(arena_alloc(A, h.pw, P));
    LV_TRY(arena_alloc(A, h.pn, P));
    LV_TRY(arena_alloc(A, h.cg, c->hub_chunks));
    LV_TRY(arena_alloc(A, h.cc, c->hub_chunks));
    LV_TRY(arena_alloc(A, h.cw, c->hub_chunks));
    LV_TRY(arena_alloc(A, h.cn, c->hub_chunks));
    return LV_OK;
}

int init_level(lv_ctx *c)
{
    Arena &A = c->arena;
    hipStream_t s = c->stream;
    const int32_t n = c->n;
    hipEvent_t t0 = t_begin(c);
    LV_TRY(arena_alloc(A, c->deg, n));
    LV_TRY(arena_alloc(A, c->in, n));
    LV_TRY(arena_alloc(A, c->size, n));
    LV_TRY(arena_alloc(A, c->dsize, n));
    LV_TRY(arena_alloc(A, c->ddeg, n));
    LV_TRY(arena_alloc(A, c->C, n));
    LV_TRY(arena_alloc(A, c->next, n));
    LV_TRY(arena_alloc(A, c->deg32, n));
    if (timing(c)) LV_TRY(arena_alloc(A, c->pcnt, n));
    uint32_t *tkey = nullptr, *tkey2 = nullptr;
    int32_t *v1 = nullptr, *v2 = nullptr;
    LV_TRY(arena_alloc(A, tkey, n));
    LV_TRY(arena_alloc(A, tkey2, n));
    LV_TRY(arena_alloc(A, v1, n));
    LV_TRY(arena_alloc(A, v2, n));
    LV_TRY(zero_small(c, 2 * (kTiers + 1) + 1));
    launch_init_vertex(s, n, c->rowptr, row_end(c), c->kdeg, c->loopw, c->selfsum, c->deg, c->in, c->size,
                       c->ddeg, c->dsize, c->C, c->next, tkey, v1, c->d_small,
                       c->d_small + (kTiers + 1), (ull *)(c->d_small + 2 * (kTiers + 1)),
                       c->deg32, c->hub_q);
    uint32_t *ko = nullptr;
    int32_t *vo = nullptr;
    LV_TRY(radix_sort_u32_i32(A, s, tkey, v1, tkey2, v2, n, 4, &ko, &vo));
    LV_TRY(fetch_small(c, {seg(c->h_small, c->d_small, 2 * (kTiers + 1) + 1)}));
    {
        // size deltas share the degree word when |delta_size| <= n < 2^(S-1) and
        // |delta_deg| <= total degree leave the packed sum below 2^63
        const int S = bits_for((uint64_t)n) + 1;
        const int kb = bits_for((uint64_t)c->h_small[2 * (kTiers + 1)]);
        c->size_shift = (kb + S <= 62) ? S : 0;
        // every community total is at most the total degree: 32 bits when it fits
        c->narrow = kb <= 32;
    }
    LV_TRY(arena_alloc(A, c->lab, n));
    if (c->narrow) LV_TRY(arena_alloc(A, c->lab8, n));
    // keep the sorted vertex list, release the other buffers
    if (vo == v1) {
        c->verts = v1;
        arena_free(A, v2);
    } else {
        c->verts = v2;
        arena_free(A, v1);
    }
    arena_free(A, tkey);
    arena_free(A, tkey2);
    int64_t off = 0;
    for (int t = 0; t <= kTiers; ++t) {
        c->tier_cnt[t] = c->h_small[t];
        c->tier_nnz[t] = c->h_small[kTiers + 1 + t];
        c->tier_off[t] = off;
        off += c->tier_cnt[t];
    }
    // hub tier: chunks and partitions per hub (see HubArgs)
    const int64_t nh = c->tier_cnt[kTiers];
    if (nh > 0) LV_TRY(plan_hubs(c, (int32_t)nh));
    LV_HIP(hipGetLastError());
    t_end(c, t0, CAT_INIT);
    return LV_OK;
}

// Streams for n independent launches of estimated sizes load[0..n): each
// (largest first) goes to whichever of the fan-out streams or the context
// stream (already carrying main_load) has the least so far; a fan-out stream
// that gets one waits on a fork event recorded on the context stream, and
// fan_join makes the context stream wait for it again.  Every launch stays on
// the context stream in timing mode (per-phase events), on a borrowed stream,
// with lv_config.fanout = 1, and (fanout = 0) when the phase's loads sum to
// less than kFanMinLoad.
int fan_assign(lv_ctx *c, int n, const int64_t *load, int64_t main_load, hipStream_t *out)
{
    for (int k = 0; k < n; ++k) out[k] = c->stream;
    if (timing(c) || c->borrowed || n == 0 || c->cfg.fanout == 1) return LV_OK;
    if (c->cfg.fanout == 0) {
        int64_t total = main_load;
        for (int k = 0; k < n; ++k) total += load[k];
        if (total < kFanMinLoad) return LV_OK;
    }
    if (!c->fan_fork) {
        for (int i = 0; i < kFanStreams; ++i) {
            LV_HIP(hipStreamCreateWithFlags(&c->fan[i], hipStreamNonBlocking));
            LV_HIP(hipEventCreateWithFlags(&c->fan_join[i], hipEventDisableTiming));
        }
        LV_HIP(hipEventCreateWithFlags(&c->fan_fork, hipEventDisableTiming));
    }
    std::vector<int> order((size_t)n);
    for (int k = 0; k < n; ++k) order[(size_t)k] = k;
    std::sort(order.begin(), order.end(), [&](int x, int y) { return load[x] > load[y]; });
    int64_t acc[kFanStreams + 1] = {};
    acc[kFanStreams] = main_load;
    bool forked = false;
    for (int k : order) {
        int best = kFanStreams;
        for (int i = 0; i < kFanStreams; ++i)
            if (acc[i] < acc[best]) best = i;
        acc[best] += load[k];
        if (best == kFanStreams) continue;
        if (!forked) {
            LV_HIP(hipEventRecord(c->fan_fork, c->stream));
            forked = true;
        }
        if (!c->fan_used[best]) {
            LV_HIP(hipStreamWaitEvent(c->fan[best], c->fan_fork, 0));
            c->fan_used[best] = true;
        }
        out[k] = c->fan[best];
    }
    return LV_OK;
}

// The stream of every sweep tier of a pass.  The tiers write next[] of
// disjoint vertex sets and read only what the pass does not change, so they
// need no order among themselves: the hubs stay on the context stream and the
// other tiers are spread by fan_assign, so small and tail-end launches overlap
// the big ones.
int fan_tiers(lv_ctx *c, hipStream_t *ts)
{
    int64_t load[kTiers - 1];
    for (int t = 1; t < kTiers; ++t) load[t - 1] = c->tier_cnt[t] ? c->tier_nnz[t] + c->tier_cnt[t] : 0;
    ts[0] = ts[kTiers] = c->stream;
    return fan_assign(c, kTiers - 1, load, c->tier_nnz[kTiers], ts + 1);
}

int fan_join(lv_ctx *c)
{
    for (int i = 0; i < kFanStreams; ++i) {
        if (!c->fan_used[i]) continue;
        LV_HIP(hipEventRecord(c->fan_join[i], c->fan[i]));
        LV_HIP(hipStreamWaitEvent(c->stream, c->fan_join[i], 0));
        c->fan_used[i] = false;
    }
    return LV_OK;
}

int one_level(lv_ctx *c, double lower, double *q_out, int32_t *passes_out)
{
    hipStream_t s = c->stream;
    int rc = LV_OK;
    double q_prev = current_q(c, rc);  // Q of the state the first pass reads
    if (rc != LV_OK) return rc;
    c->cur_pass_q.clear();
    c->cur_trace.clear();
    const bool trace = (c->cfg.flags & LV_FLAG_TRACE) != 0;
    SweepArgs a{};
    a.lab = c->lab;
    a.lab8 = c->narrow ? c->lab8 : nullptr;
    a.deg32 = c->narrow ? c->deg32 : nullptr;
    a.rowptr = c->rowptr;
    a.rowend = row_end(c);
    a.col = c->col;
    a.w = c->cw;
    a.w32 = c->cw32;
    a.cpk = c->cpk;
    a.kdeg = c->kdeg;
    a.loopw = c->loopw;
    a.selfsum = c->selfsum;
    a.deg = c->deg;
    a.size = c->size;
    a.ddeg = c->ddeg;
    a.dsize = c->dsize;
    a.size_shift = c->size_shift;
    a.m2 = 2.0 * (double)c->m;
    a.gamma = c->cfg.gamma;
    a.pcnt = c->pcnt;
    double curQ = -1.0;  // LouvainCUDA/OneLevelKernel.cu:528
    int32_t pass = 0;
    double q_level = 0.0;
    // With min_gain > 0 the second pass is always the last: within a level
    // sum(in) stays the sum of the self-loops and sum(deg^2) over any
    // partition is at least its all-singletons value, so Q after the first
    // pass cannot exceed Q before it and its break test cannot pass.  The first
    // pass's Q is then fetched with the second's (slot 1 of d_qacc), one host
    // round trip per level fewer; the break test is still evaluated on it.
    const bool defer = !c->borrowed && !timing(c) && c->cfg.min_gain > 0 &&
                       (c->cfg.max_passes == 0 || c->cfg.max_passes >= 2);
    bool deferred = false;  // pass 1's Q waits in slot 1
    for (;;) {
        ++pass;
        // The break test of pass p compares Q of the state before p's deltas
        // (OneLevelKernel.cu:875-892), known before the sweep runs.
        const bool last = deferred || (q_prev - curQ) < c->cfg.min_gain ||
                          (c->cfg.max_passes > 0 && pass >= c->cfg.max_passes);
        a.C = c->C;
        a.next = c->next;
        a.last = last ? 1 : 0;
        // first pass: every community is one vertex (no aggregation, no labels)
        const bool singles = pass == 1 && c->rows_unique;
        if (!singles) {
            hipEvent_t tl = t_begin(c);
            launch_lab(s, c->n, c->C, c->size, c->deg, c->lab, c->narrow ? c->lab8 : nullptr);
            t_end(c, tl, CAT_APPLY);
        }
        p_begin(c);
        hipStream_t ts[kTiers + 1];
        LV_TRY(fan_tiers(c, ts));
        for (int t = 1; t <= kTiers; ++t) {
            const int64_t nv = c->tier_cnt[t];
            if (nv == 0) continue;
            hipEvent_t t0 = t_begin(c);
            if (t < kTiers && singles)
                launch_sweep1_tier(ts[t], t, a, c->verts + c->tier_off[t], (int32_t)nv, false);
            else if (t < kTiers)
                launch_sweep_tier(ts[t], t, a, c->verts + c->tier_off[t], (int32_t)nv, false);
            else if (singles)
                LV_TRY(run_hub_sweep1(s, a, c->hub, c->n_hubs, c->hub_chunks, false));
            else
                LV_TRY(run_hub_sweep(c->arena, s, a, c->hub, c->n_hubs, c->hub_parts,
                                     c->hub_chunks, false));
            t_end(c, t0, CAT_SWEEP, t, 12.0 * (double)c->tier_nnz[t] + 24.0 * (double)nv,
                  c->tier_nnz[t]);
            p_tier(c, t, nv);
        }
        LV_TRY(fan_join(c));
        LV_TRY(p_fetch(c));
        if (!last) {
            hipEvent_t td = t_begin(c);
            launch_move_deltas(s, a, c->n, pass == 1, false);  // pass 1: C[v] = v
            t_end(c, td, CAT_APPLY);
        }
        LV_HIP(hipGetLastError());
        if (trace) {
            std::vector<int32_t> nx((size_t)c->n);
            LV_TRY(d2h(c, nx.data(), c->next, c->n));
            LV_TRY(sync(c));
            c->cur_trace.push_back(std::move(nx));
        }
        const bool hold = defer && pass == 1 && !last;
        hipEvent_t t1 = t_begin(c);
        launch_apply(s, c->n, c->deg, c->in, c->size, c->ddeg, c->dsize, c->size_shift,
                     c->narrow ? c->deg32 : nullptr, c->d_qpart, c->d_qacc + (hold ? 1 : 0));
        t_end(c, t1, CAT_APPLY);
        LV_HIP(hipGetLastError());
        if (hold) {
            c->cur_pass_q.push_back(q_prev);
            std::swap(c->C, c->next);
            curQ = q_prev;
            if (curQ < lower) curQ = lower;  // :894-896
            deferred = true;
            continue;
        }
        LV_TRY(fetch_small(c, {seg(c->h_qacc, c->d_qacc, deferred ? 2 : 1)}));
        p_resolve(c);
        if (deferred) {
            q_prev = q_from_aggregates(c->h_qacc[1].sum_in, c->h_qacc[1].d2_lo,
                                       c->h_qacc[1].d2_hi, c->m, c->cfg.gamma);
            if (!((q_prev - curQ) < c->cfg.min_gain)) return LV_ERR_DEVICE;  // cannot happen
        }
        const double q_new = q_from_aggregates(c->h_qacc->sum_in, c->h_qacc->d2_lo,
                                               c->h_qacc->d2_hi, c->m, c->cfg.gamma);
        c->cur_pass_q.push_back(q_prev);
        std::swap(c->C, c->next);
        if (last) {
            q_level = q_new;  // Q after the merge-isolated deltas, :941-953
            break;
        }
        curQ = q_prev;
        if (curQ < lower) curQ = lower;  // :894-896
        q_prev = q_new;
    }
    c->cur_q = q_level;
    c->cur_lower = lower;
    c->cur_passes = pass;
    c->stats.passes += pass;
    if (q_out) *q_out = q_level;
    if (passes_out) *passes_out = pass;
    return LV_OK;
}

// renumberCUDA: stable sort of positions by community; group heads are the
// smallest members; new id = rank of the head among heads in position order.
int renumber(lv_ctx *c, int32_t *k_out)
{
    Arena &A = c->arena;
    hipStream_t s = c->stream;
    const int32_t n = c->n;
    hipEvent_t t0 = t_begin(c);
    LevelRec L;
    uint32_t *k1 = nullptr, *k2 = nullptr;
    int32_t *v1 = nullptr, *v2 = nullptr, *lflag = nullptr, *rank = nullptr;
    LV_TRY(arena_alloc(A, k1, n));
    LV_TRY(arena_alloc(A, k2, n));
    LV_TRY(arena_alloc(A, v1, n));
    LV_TRY(arena_alloc(A, v2, n));
    copy_i32(s, (int32_t *)k1, c->C, n);
    iota_i32(s, v1, n);
    uint32_t *Cs = nullptr;
    int32_t *perm = nullptr;
    LV_TRY(radix_sort_u32_i32(A, s, k1, v1, k2, v2, n, bits_for(n > 0 ? (uint64_t)(n - 1) : 0),
                              &Cs, &perm));
    L.perm = perm;
    arena_free(A, perm == v1 ? v2 : v1);
    LV_TRY(arena_alloc(A, c->g_head, n));
    LV_TRY(arena_alloc(A, c->g_hscan, n));
    LV_TRY(arena_alloc(A, c->g_start, (int64_t)n + 1));
    LV_TRY(arena_alloc(A, lflag, n));
    LV_TRY(arena_alloc(A, rank, n));
    launch_cr_heads(s, Cs, n, c->g_head);
    LV_TRY(scan_exclusive_i32(A, s, c->g_head, c->g_hscan, n, (int32_t *)c->d_small));
    fill_i32(s, lflag, n, 0);
    launch_cr_groups(s, c->g_head, c->g_hscan, perm, n, c->g_start, lflag);
    LV_TRY(scan_exclusive_i32(A, s, lflag, rank, n, nullptr));
    int32_t kk = 0;
    LV_TRY(fetch_small(c, {seg(&kk, (const int32_t *)c->d_small, 1)}));
    LV_TRY(arena_alloc(A, c->g_newid, kk));
    LV_TRY(arena_alloc(A, L.newc, n));
    launch_cr_newid(s, c->g_start, perm, rank, kk, c->g_newid);
    launch_cr_newc(s, c->g_head, c->g_hscan, perm, c->g_newid, n, L.newc);
    LV_HIP(hipGetLastError());
    t_end(c, t0, CAT_RENUMBER);
    arena_free(A, k1);
    arena_free(A, k2);
    arena_free(A, lflag);
    arena_free(A, rank);
    free_sweep_state(c);  // the level's communities are final
    L.n = n;
    L.k = kk;
    L.nnz = c->nnz;
    L.m = c->m;
    L.passes = c->cur_passes;
    L.q = c->cur_q;
    L.lower = c->cur_lower;
    L.pass_q = std::move(c->cur_pass_q);
    L.trace = std::move(c->cur_trace);
    c->cur_pass_q.clear();
    c->cur_trace.clear();
    c->levels.push_back(std::move(L));
    if (k_out) *k_out = kk;
    return LV_OK;
}

// inducedGraphCUDA: gather each community's entries in member order, aggregate
// every community segment by neighbour community, copy the unique entries to
// row newid(community) of the next level.
// newc/newid_g: next-level vertex id of every position / group (-1 for the
// groups of a graph set's stopped components), Kn: vertices of the next level.
int induce(lv_ctx *c, const int32_t *newc, const int32_t *newid_g, int32_t Kn)
{
    Arena &A = c->arena;
    hipStream_t s = c->stream;
    const LevelRec &L = c->levels.back();
    const int32_t n = c->n;
    const int32_t K = L.k;
    const int64_t nnz = c->nnz;
    hipEvent_t t0 = t_begin(c);
    // member order entry starts and the member-order gather
    int64_t *mdeg = nullptr, *estart = nullptr;
    ull *selfl = nullptr, *ssum = nullptr, *internal = nullptr;
    int32_t *ey = nullptr, *bound = nullptr, *uy = nullptr, *nuniq = nullptr;
    int64_t *ew = nullptr, *uw = nullptr;
    uint32_t *ew32 = nullptr, *uw32 = nullptr;  // narrow level: every weight is below 2^32
    int64_t *mstart = nullptr;
    const bool nar = c->narrow;
    LV_TRY(arena_alloc(A, mdeg, n));
    LV_TRY(arena_alloc(A, mstart, n));
    LV_TRY(arena_alloc(A, estart, (int64_t)n + 1));
    LV_TRY(arena_alloc(A, selfl, K));
    LV_TRY(arena_alloc(A, ssum, K));
    LV_TRY(arena_alloc(A, internal, K));
    LV_TRY(arena_alloc(A, nuniq, K));
    // level 0 with small weights: each entry as one word newc << ewb | w'
    uint32_t *ep = nullptr;
    const int ewb = pack_bits(c);
    const bool packed = c->level == 0 && ewb < 32 && bits_for((uint64_t)Kn - 1) + ewb <= 32;
    if (packed) {
        LV_TRY(arena_alloc(A, ep, nnz));
    } else {
        LV_TRY(arena_alloc(A, ey, nnz));
        if (nar)
            LV_TRY(arena_alloc(A, ew32, nnz));
        else
            LV_TRY(arena_alloc(A, ew, nnz));
    }
    // uy/uw become the next level's col/w: segment g's uniques at sb[g] are row
    // newid_g[g] (the total degree bounds every weight and is the same at every
    // level, so a narrow level's coarse rows keep 4-byte weights)
    LV_TRY(arena_alloc(A, uy, nnz));
    if (nar)
        LV_TRY(arena_alloc(A, uw32, nnz));
    else
        LV_TRY(arena_alloc(A, uw, nnz));
    LV_TRY(arena_alloc(A, bound, nnz / 1024 + 2));
    launch_cr_memdeg(s, L.perm, c->rowptr, row_end(c), n, mdeg, mstart);
    LV_TRY(scan_exclusive_i64(A, s, mdeg, estart, n, estart + n));
    fill_i64(s, (int64_t *)selfl, K, 0);
    fill_i64(s, (int64_t *)ssum, K, 0);
    fill_i64(s, (int64_t *)internal, K, 0);
    fill_i32(s, nuniq, K, 0);
    {
        // the self-loop sums and the gather run beside the segment sort below
        // (joined before the segment tiers read the gather's entries)
        const int64_t load[2] = {n, nnz};
        hipStream_t st[2];
        LV_TRY(fan_assign(c, 2, load, (int64_t)K * 8, st));
        launch_cr_selfl(st[0], L.perm, c->selfsum, c->g_head, c->g_hscan, n, selfl);
        launch_cr_gather(st[1], estart, bound, mstart, c->col, c->cw, c->cw32, newc, n, nnz, ey, ew,
                         ew32, ep, ewb, c->cpk);
    }
    // segments by length
    int64_t *sb = nullptr, *slen = nullptr;
    uint32_t *tk1 = nullptr, *tk2 = nullptr;
    int32_t *sid1 = nullptr, *sid2 = nullptr;
    LV_TRY(arena_alloc(A, sb, K));
    LV_TRY(arena_alloc(A, slen, K));
    LV_TRY(arena_alloc(A, tk1, K));
    LV_TRY(arena_alloc(A, tk2, K));
    LV_TRY(arena_alloc(A, sid1, K));
    LV_TRY(arena_alloc(A, sid2, K));
    LV_TRY(zero_small(c, kTiers + 1));
    // segments of more than 2048 entries take the hash partitions whatever hub_q
    // is (one block with an 8192-slot table measured slower on them)
    const int32_t seg_q = c->hub_q < kTierMaxQ[kTiers - 3] ? c->hub_q : kTierMaxQ[kTiers - 3];
    launch_cr_seg_tier(s, c->g_start, estart, newid_g, K, sb, slen, tk1, sid1, c->d_small, seg_q);
    uint32_t *tko = nullptr;
    int32_t *segs = nullptr;
    LV_TRY(radix_sort_u32_i32(A, s, tk1, sid1, tk2, sid2, K, 4, &tko, &segs));
    LV_TRY(fetch_small(c, {seg(c->h_small, c->d_small, kTiers + 1)}));
    int64_t tcnt[kTiers + 1], toff[kTiers + 1], off = 0;
    for (int t = 0; t <= kTiers; ++t) {
        tcnt[t] = c->h_small[t];
        toff[t] = off;
        off += tcnt[t];
    }
    CoarsenArgs ca{};
    ca.ey = ey;
    ca.ew = ew;
    ca.ew32 = ew32;
    ca.ep = ep;
    ca.ewb = ewb;
    ca.n32 = nar ? 1 : 0;
    ca.sb = sb;
    ca.slen = slen;
    ca.newid_g = newid_g;
    ca.uy = uy;
    ca.uw = uw;
    ca.uw32 = uw32;
    ca.nuniq = nuniq;
    ca.internal = internal;
    ca.ssum = ssum;
    ca.big_q = seg_q;
    LV_TRY(fan_join(c));  // the gather's entries
    // the segment tiers are independent of each other and of the big segments'
    // partitions below: spread over the fan-out streams (joined before the
    // per-vertex pass reads their sums)
    {
        int64_t load[kTiers - 1];
        hipStream_t st[kTiers - 1];
        for (int t = 1; t < kTiers; ++t) load[t - 1] = tcnt[t] ? tcnt[t] * (int64_t)kTierMaxQ[t - 1] : 0;
        LV_TRY(fan_assign(c, kTiers - 1, load, seg_q * (int64_t)tcnt[kTiers], st));
        for (int t = 1; t < kTiers; ++t)
            launch_cr_agg_tier(st[t - 1], t, ca, segs + toff[t], (int32_t)tcnt[t]);
    }
    // segments above seg_q entries: hash partitions
    BigSegArgs B{};
    const int32_t nbs = (int32_t)tcnt[kTiers];
    int64_t nparts = 0, nchunks = 0, blen = 0;
    int32_t *bP = nullptr, *part_big = nullptr, *chunk_big = nullptr, *chunk_k = nullptr;
    int64_t *bP64 = nullptr, *bch = nullptr, *pbase = nullptr, *cbase = nullptr, *bhbase = nullptr;
    if (nbs > 0) {
        const int32_t *bsegs = segs + toff[kTiers];
        int64_t *blen64 = nullptr, *blscan = nullptr, *bpc = nullptr;
        LV_TRY(arena_alloc(A, bP, nbs));
        LV_TRY(arena_alloc(A, bP64, nbs));
        LV_TRY(arena_alloc(A, bch, nbs));
        LV_TRY(arena_alloc(A, blen64, nbs));
        LV_TRY(arena_alloc(A, bpc, nbs));
        LV_TRY(arena_alloc(A, blscan, (int64_t)nbs + 1));
        LV_TRY(arena_alloc(A, pbase, (int64_t)nbs + 1));
        LV_TRY(arena_alloc(A, cbase, (int64_t)nbs + 1));
        LV_TRY(arena_alloc(A, bhbase, (int64_t)nbs + 1));
        launch_cr_big_setup(s, bsegs, nbs, slen, c->hub_part_target, c->chunk_mul, bP, bP64, bch,
                            blen64, bpc, nuniq);
        B.chunk_mul = c->chunk_mul;
        B.small_tables = c->hub_part_target <= kHubAggSmallTarget ? 1 : 0;
        LV_TRY(scan_exclusive_i64(A, s, bP64, pbase, nbs, pbase + nbs));
        LV_TRY(scan_exclusive_i64(A, s, bch, cbase, nbs, cbase + nbs));
        LV_TRY(scan_exclusive_i64(A, s, blen64, blscan, nbs, blscan + nbs));
        LV_TRY(scan_exclusive_i64(A, s, bpc, bhbase, nbs, bhbase + nbs));
        LV_TRY(fetch_small(c, {seg(c->h_small, (const long long *)(pbase + nbs), 1),
                               seg(c->h_small + 1, (const long long *)(cbase + nbs), 1),
                               seg(c->h_small + 2, (const long long *)(blscan + nbs), 1),
                               seg(c->h_small + 3, (const long long *)(bhbase + nbs), 1)}));
        nparts = c->h_small[0];
        nchunks = c->h_small[1];
        blen = c->h_small[2];  // entries of the big segments
        const int64_t nhist = c->h_small[3];
        arena_free(A, blen64);
        arena_free(A, blscan);
        arena_free(A, bpc);
        B.hbase = bhbase;
        LV_TRY(arena_alloc(A, B.chist, nhist));
        LV_TRY(arena_alloc(A, part_big, nparts));
        LV_TRY(arena_alloc(A, chunk_big, nchunks));
        LV_TRY(arena_alloc(A, chunk_k, nchunks));
        launch_expand_owner(s, pbase, nbs, nparts, part_big, nullptr);
        launch_expand_owner(s, cbase, nbs, nchunks, chunk_big, chunk_k);
        B.bsegs = bsegs;
        B.bP = bP;
        B.pbase = pbase;
        B.part_big = part_big;
        B.chunk_big = chunk_big;
        B.chunk_k = chunk_k;
        LV_TRY(arena_alloc(A, B.pcount, nparts + 1));
        LV_TRY(arena_alloc(A, B.poff, nparts + 1));
        B.n32 = nar ? 1 : 0;
        if (packed) {
            // a next-level id and an entry weight w' of level 0 in one word
            B.pwb = pack_bits(c);
            LV_TRY(arena_alloc(A, B.pe4, blen));
        } else if (nar) {
            LV_TRY(arena_alloc(A, B.pe8, blen));
        } else {
            LV_TRY(arena_alloc(A, B.pe, blen));
        }
        LV_TRY(arena_alloc(A, B.fbk, 2 * blen));
        LV_TRY(arena_alloc(A, B.fbv, 2 * blen));
        LV_TRY(run_cr_big(A, s, ca, B, nbs, nparts, nchunks));
    }
    LV_TRY(fan_join(c));  // the segment tiers' sums and rows
    // next level: row x = newid_g[g] is [sb[g], sb[g] + nuniq[g]) of uy/uw; the
    // rows keep their segments' places (a segment holds at least its uniques),
    // so no copy packs them
    int64_t *rowptr2 = nullptr, *rowend2 = nullptr, *kdeg2 = nullptr, *loop2 = nullptr,
            *self2 = nullptr;
    LV_TRY(arena_alloc(A, rowptr2, (int64_t)Kn + 1));
    LV_TRY(arena_alloc(A, rowend2, Kn));
    LV_TRY(arena_alloc(A, kdeg2, Kn));
    LV_TRY(arena_alloc(A, loop2, Kn));
    LV_TRY(arena_alloc(A, self2, Kn));
    fill_i64(s, rowptr2 + Kn, 1, nnz);
    fill_i64(s, (int64_t *)c->d_small + 1, 2, 0);
    launch_cr_vertex(s, newid_g, ssum, internal, selfl, sb, nuniq, K, loop2, self2, kdeg2,
                     rowptr2, rowend2, c->d_small + 1, c->d_small + 2);
    LV_TRY(fetch_small(c, {seg(c->h_small + 1, c->d_small + 1, 2)}));
    const int64_t m2 = c->h_small[1] / 2;
    const int64_t nnz2 = c->h_small[2];
    LV_HIP(hipGetLastError());
    t_end(c, t0, CAT_COARSEN);
    // release the transient buffers (later launches on this stream may reuse them)
    void *tmp[] = {mdeg, mstart, estart, selfl, ssum, internal, nuniq, ey, ew, ew32, ep, bound, sb, slen,
                   tk1, tk2, sid1, sid2, bP, bP64, bch, pbase, cbase, part_big, chunk_big, chunk_k,
                   B.pcount, B.poff, B.pe, B.pe8, B.pe4, B.fbk, B.fbv,
                   B.chist, bhbase};
    for (void *p : tmp) A.release(p);
    free_level_arrays(c);
    c->rowptr = rowptr2;
    c->rowend_buf = rowend2;
    c->span = nnz;
    c->col = uy;
    c->cw = uw;
    c->cw32 = uw32;
    c->cpk = 0;
    c->kdeg = kdeg2;
    c->loopw = loop2;
    c->selfsum = self2;
    c->n = Kn;
    c->nnz = nnz2;
    c->rows_unique = true;  // one entry per neighbour community
    c->m = m2;
    c->level += 1;
    return LV_OK;
}

// Level 0 of a graph set: the CSR of the graphs (built once) written out for
// every resolution, so each component's rows are its graph's.
int build_level0_csr_sets(lv_ctx *c)
{
    Arena &A = c->arena;
    hipStream_t s = c->stream;
    const int32_t nu = c->n0;
    const int64_t eu = c->e0;
    c->n0 = c->base_n0;
    c->e0 = c->base_e0;
    const int rc = build_level0_csr(c);
    c->n0 = nu;
    c->e0 = eu;
    if (rc != LV_OK) return rc;
    const int32_t nres = c->nres, nb = c->base_n0;
    if (nres == 1) return LV_OK;
    hipEvent_t t0 = t_begin(c);
    const int64_t unnz = (int64_t)nres * c->nnz;
    int64_t *urow = nullptr, *uk = nullptr, *ul = nullptr, *us = nullptr, *uw = nullptr,
            *geoff = nullptr;
    int32_t *ucol = nullptr;
    uint32_t *uw32 = nullptr;
    LV_TRY(arena_alloc(A, urow, (int64_t)nu + 1));
    LV_TRY(arena_alloc(A, uk, nu));
    LV_TRY(arena_alloc(A, ul, nu));
    LV_TRY(arena_alloc(A, us, nu));
    LV_TRY(arena_alloc(A, ucol, unnz));
    if (c->cw32)
        LV_TRY(arena_alloc(A, uw32, unnz));
    else
        LV_TRY(arena_alloc(A, uw, unnz));
    LV_TRY(arena_alloc(A, geoff, (int64_t)c->ngraphs + 1));
    launch_b_replicate_csr(s, c->d_bvoff, geoff, c->ngraphs, nres, nb, c->nnz, c->rowptr, c->col,
                           c->cw, c->cw32, c->kdeg, c->loopw, c->selfsum, urow, ucol, uw, uw32, uk,
                           ul, us);
    LV_HIP(hipGetLastError());
    t_end(c, t0, CAT_CSR);
    arena_free(A, geoff);
    arena_free(A, c->rowptr);
    arena_free(A, c->col);
    arena_free(A, c->cw);
    arena_free(A, c->cw32);
    arena_free(A, c->kdeg);
    arena_free(A, c->loopw);
    arena_free(A, c->selfsum);
    c->rowptr = urow;
    c->col = ucol;
    c->cw = uw;
    c->cw32 = uw32;
    c->kdeg = uk;
    c->loopw = ul;
    c->selfsum = us;
    c->nnz = unnz;
    c->n = nu;
    return LV_OK;
}

int flatten_final(lv_ctx *c)
{
    hipEvent_t t0 = t_begin(c);
    arena_free(c->arena, c->d_final);
    LV_TRY(arena_alloc(c->arena, c->d_final, c->n0));
    copy_i32(c->stream, c->d_final, c->levels[0].newc, c->n0);
    for (size_t l = 1; l < c->levels.size(); ++l)
        launch_compose(c->stream, c->d_final, c->n0, c->levels[l].newc);
    LV_HIP(hipGetLastError());
    t_end(c, t0, CAT_FLATTEN);
    c->final_valid = true;
    return LV_OK;
}

// lv_run_partitions: the newest level's Partition (reference layout) goes to the
// caller's host buffers on the copy stream while the main stream runs on.
int sink_level(lv_ctx *c)
{
    PartSink *k = c->sink;
    const int32_t l = (int32_t)c->levels.size() - 1;
    if (!k || l < 0 || l >= k->max) return LV_OK;
    const LevelRec &L = c->levels[l];
    if (L.n == 0) return LV_OK;
    int32_t *dn = nullptr, *dc = nullptr;
    LV_TRY(arena_alloc(c->arena, dn, L.n));
    k->temps.push_back(dn);
    LV_TRY(arena_alloc(c->arena, dc, L.n));
    k->temps.push_back(dc);
    launch_gather_partition(c->stream, L.perm, l == 0 ? c->d_ids : nullptr, L.newc, L.n, dn, dc);
    LV_HIP(hipGetLastError());
    while ((int32_t)c->sink_ev.size() <= l) {
        hipEvent_t e = nullptr;
        LV_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        c->sink_ev.push_back(e);
    }
    LV_HIP(hipEventRecord(c->sink_ev[l], c->stream));
    LV_HIP(hipStreamWaitEvent(c->copy_stream, c->sink_ev[l], 0));
    const size_t b = sizeof(int32_t) * (size_t)L.n;
    LV_HIP(hipMemcpyAsync(k->node[l], dn, b, hipMemcpyDeviceToHost, c->copy_stream));
    LV_HIP(hipMemcpyAsync(k->comm[l], dc, b, hipMemcpyDeviceToHost, c->copy_stream));
    return LV_OK;
}

// ---------------------------------------------------------------------------
// graph sets: the level loop of every component, run on the union graph

// Frees pinned host memory; a failure is reported and cleared here so that it
// cannot surface later as the error of an unrelated call.
void pinned_free(void *&p)
{
    if (!p) return;
    const hipError_t e = hipHostFree(p);
    if (e != hipSuccess) {
        (void)hip_fail(e, "hipHostFree", __FILE__, __LINE__);
        (void)hipGetLastError();
    }
    p = nullptr;
}

// Pinned per-component buffers, grown geometrically (rarely reallocated).
int pinned_cap(lv_ctx *c, int32_t need)
{
    if (c->h_cap >= need) return LV_OK;
    const int32_t cap = std::max(need, std::max(4096, 2 * c->h_cap));
    void *a = c->h_cq, *b = c->h_ccnt;
    pinned_free(a);
    pinned_free(b);
    c->h_cq = nullptr;
    c->h_ccnt = nullptr;
    c->h_cap = 0;
    if (hipHostMalloc(&c->h_cq, sizeof(QAcc) * (size_t)cap, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&c->h_ccnt, sizeof(ull) * (size_t)cap, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return LV_ERR_NOMEM;
    }
    c->h_cap = cap;
    return LV_OK;
}

// Component of every vertex of the current level from comp_voff, each
// component's m at levels above 0 (from the coarsened k and loops), and the
// per-component 2m / resolution tables.
int batch_level_setup(lv_ctx *c, int32_t level)
{
    Arena &A = c->arena;
    hipStream_t s = c->stream;
    const int32_t nc = c->ncomp;
    int64_t *dvoff = nullptr;
    LV_TRY(arena_alloc(A, dvoff, (int64_t)nc + 1));
    LV_TRY(h2d(c, dvoff, (const int64_t *)c->comp_voff.data(), (int64_t)nc + 1));
    LV_TRY(arena_alloc(A, c->vcomp, c->n));
    launch_expand_owner(s, dvoff, nc, c->n, c->vcomp, nullptr);
    if (level > 0) {
        fill_i64(s, (int64_t *)c->d_ccnt, nc, 0);
        launch_b_vertex_sum(s, c->kdeg, c->loopw, c->vcomp, c->n, c->d_ccnt);
        LV_HIP(hipGetLastError());
        LV_TRY(fetch_small(c, {seg(c->h_ccnt, c->d_ccnt, nc)}));
        for (int32_t k = 0; k < nc; ++k)
            if (c->comp_voff[k + 1] > c->comp_voff[k]) c->comp_m[k] = (int64_t)(c->h_ccnt[k] / 2);
    }
    std::vector<double> m2((size_t)nc), gam((size_t)nc);
    for (int32_t k = 0; k < nc; ++k) {
        m2[k] = 2.0 * (double)c->comp_m[k];
        gam[k] = c->comp_gamma[k];
    }
    LV_TRY(h2d(c, c->d_cm2, (const double *)m2.data(), nc));
    LV_TRY(h2d(c, c->d_cgam, (const double *)gam.data(), nc));
    LV_TRY(barrier(c));  // the host vectors go out of scope
    arena_free(A, dvoff);
    return LV_OK;
}

// Exact Q aggregates per component (apply: add the pass deltas first).
// Per-component Q terms into d_cq + slot * ncomp; fetch = 1 brings slot 0 to
// h_cq, fetch = 2 also slot 1 (a deferred first pass) to h_cq + ncomp.
int batch_qagg(lv_ctx *c, bool apply, int slot = 0, int fetch = 1)
{
    QAcc *out = c->d_cq + (int64_t)slot * c->ncomp;
    fill_i64(c->stream, (int64_t *)out, 4 * (int64_t)c->ncomp, 0);
    launch_b_qagg(c->stream, c->n, c->deg, c->in, c->size, c->ddeg, c->dsize, c->size_shift,
                  c->narrow ? c->deg32 : nullptr, c->vcomp, apply, out);
    LV_HIP(hipGetLastError());
    if (fetch == 2) return fetch_small(c, {seg(c->h_cq, c->d_cq, 2 * (int64_t)c->ncomp)});
    if (fetch == 1) return fetch_small(c, {seg(c->h_cq, c->d_cq, c->ncomp)});
    return LV_OK;
}

double batch_q(const lv_ctx *c, int32_t k)
{
    const QAcc &a = c->h_cq[k];
    return q_from_aggregates(a.sum_in, a.d2_lo, a.d2_hi, c->comp_m[k], c->comp_gamma[k]);
}

// oneLevel for every present component: the same pass rule as one_level, with
// each component's break decided from its own Q; components that have
// finished (or have no edges) are frozen while the others sweep.
int batch_one_level(lv_ctx *c, const std::vector<double> &lower, const std::vector<uint8_t> &present,
                    std::vector<double> &q_level, std::vector<int32_t> &passes)
{
    hipStream_t s = c->stream;
    const int32_t nc = c->ncomp;
    LV_TRY(batch_qagg(c, false));
    std::vector<double> q_prev((size_t)nc, 0.0), curQ((size_t)nc, -1.0);
    std::vector<uint8_t> done((size_t)nc, 1);
    // one mode array per pass parity: a deferred first pass leaves its copy
    // queued while the second pass fills the other
    std::vector<int32_t> modes[2] = {std::vector<int32_t>((size_t)nc, 0),
                                     std::vector<int32_t>((size_t)nc, 0)};
    // as in one_level: with min_gain > 0 the second pass is every component's
    // last, so the first pass's Q terms wait on the device (slot 1 of d_cq)
    const bool defer = !timing(c) && c->cfg.min_gain > 0 &&
                       (c->cfg.max_passes == 0 || c->cfg.max_passes >= 2);
    bool deferred = false;
    q_level.assign((size_t)nc, 0.0);
    passes.assign((size_t)nc, 0);
    for (int32_t k = 0; k < nc; ++k) {
        if (!present[k] || c->comp_e0[k] == 0) continue;
        done[k] = 0;
        q_prev[k] = batch_q(c, k);
    }
    SweepArgs a{};
    a.lab = c->lab;
    a.lab8 = c->narrow ? c->lab8 : nullptr;
    a.deg32 = c->narrow ? c->deg32 : nullptr;
    a.rowptr = c->rowptr;
    a.rowend = row_end(c);
    a.col = c->col;
    a.w = c->cw;
    a.w32 = c->cw32;
    a.cpk = c->cpk;
    a.kdeg = c->kdeg;
    a.loopw = c->loopw;
    a.selfsum = c->selfsum;
    a.deg = c->deg;
    a.size = c->size;
    a.ddeg = c->ddeg;
    a.dsize = c->dsize;
    a.size_shift = c->size_shift;
    a.vcomp = c->vcomp;
    a.cmode = c->d_cmode;
    a.cm2 = c->d_cm2;
    a.cgam = c->d_cgam;
    a.pcnt = c->pcnt;
    int32_t pass = 0;
    for (;;) {
        bool any = false, any_last = false;
        std::vector<int32_t> &mode = modes[pass & 1];
        for (int32_t k = 0; k < nc; ++k) {
            if (done[k]) {
                mode[k] = 0;
                continue;
            }
            any = true;
            passes[k] += 1;
            const bool last = deferred || (q_prev[k] - curQ[k]) < c->cfg.min_gain ||
                              (c->cfg.max_passes > 0 && passes[k] >= c->cfg.max_passes);
            mode[k] = last ? 2 : 1;
            any_last = any_last || last;
        }
        if (!any) break;
        ++pass;
        LV_TRY(h2d(c, c->d_cmode, (const int32_t *)mode.data(), nc));
        a.C = c->C;
        a.next = c->next;
        const bool singles = pass == 1 && c->rows_unique;  // see one_level
        if (!singles) {
            hipEvent_t tl = t_begin(c);
            launch_lab(s, c->n, c->C, c->size, c->deg, c->lab, c->narrow ? c->lab8 : nullptr);
            t_end(c, tl, CAT_APPLY);
        }
        p_begin(c);
        hipStream_t ts[kTiers + 1];
        LV_TRY(fan_tiers(c, ts));
        for (int t = 1; t <= kTiers; ++t) {
            const int64_t nv = c->tier_cnt[t];
            if (nv == 0) continue;
            hipEvent_t t0 = t_begin(c);
            if (t < kTiers && singles)
                launch_sweep1_tier(ts[t], t, a, c->verts + c->tier_off[t], (int32_t)nv, true);
            else if (t < kTiers)
                launch_sweep_tier(ts[t], t, a, c->verts + c->tier_off[t], (int32_t)nv, true);
            else if (singles)
                LV_TRY(run_hub_sweep1(s, a, c->hub, c->n_hubs, c->hub_chunks, true));
            else
                LV_TRY(run_hub_sweep(c->arena, s, a, c->hub, c->n_hubs, c->hub_parts,
                                     c->hub_chunks, true));
            t_end(c, t0, CAT_SWEEP, t, 12.0 * (double)c->tier_nnz[t] + 24.0 * (double)nv,
                  c->tier_nnz[t]);
            p_tier(c, t, nv);
        }
        LV_TRY(fan_join(c));
        LV_TRY(p_fetch(c));
        {
            hipEvent_t td = t_begin(c);
            launch_move_deltas(s, a, c->n, pass == 1, true);  // pass 1: C[v] = v
            t_end(c, td, CAT_APPLY);
        }
        LV_HIP(hipGetLastError());
        const bool hold = defer && pass == 1 && !any_last;
        hipEvent_t t1 = t_begin(c);
        LV_TRY(batch_qagg(c, true, hold ? 1 : 0, hold ? 0 : deferred ? 2 : 1));
        t_end(c, t1, CAT_APPLY);
        if (hold) {
            std::swap(c->C, c->next);
            for (int32_t k = 0; k < nc; ++k) {
                if (mode[k] == 0) continue;
                curQ[k] = q_prev[k];
                if (curQ[k] < lower[k]) curQ[k] = lower[k];
            }
            deferred = true;
            continue;
        }
        p_resolve(c);
        std::swap(c->C, c->next);
        for (int32_t k = 0; k < nc; ++k) {
            if (mode[k] == 0) continue;
            if (deferred) {
                const QAcc &d = c->h_cq[nc + k];
                q_prev[k] = q_from_aggregates(d.sum_in, d.d2_lo, d.d2_hi, c->comp_m[k],
                                              c->comp_gamma[k]);
                if (!((q_prev[k] - curQ[k]) < c->cfg.min_gain)) return LV_ERR_DEVICE;  // cannot happen
            }
            const double q_new = batch_q(c, k);
            if (mode[k] == 2) {
                q_level[k] = q_new;
                done[k] = 1;
                continue;
            }
            curQ[k] = q_prev[k];
            if (curQ[k] < lower[k]) curQ[k] = lower[k];
            q_prev[k] = q_new;
        }
    }
    c->stats.passes += pass;
    return LV_OK;
}

// generate_dendogram for every component: the same level loop as lv_run
// (lower, the level-0 exemption from the break test, max_levels), each
// component deciding for itself; components that stop drop out of the union
// at the next coarsening.
int batch_run(lv_ctx *c)
{
    Arena &A = c->arena;
    hipStream_t s = c->stream;
    const int32_t nc = c->ncomp;
    LV_TRY(pinned_cap(c, 2 * nc));
    LV_TRY(arena_alloc(A, c->d_cmode, nc));
    LV_TRY(arena_alloc(A, c->d_cm2, nc));
    LV_TRY(arena_alloc(A, c->d_cgam, nc));
    LV_TRY(arena_alloc(A, c->d_cq, 2 * (int64_t)nc));  // slot 1: a deferred first pass
    LV_TRY(arena_alloc(A, c->d_ccnt, nc));
    c->comp_rec.assign((size_t)nc, {});
    c->comp_voff.assign((size_t)nc + 1, 0);
    for (int32_t k = 0; k < nc; ++k) c->comp_voff[k + 1] = c->comp_voff[k] + c->comp_n0[k];
    c->comp_m = c->comp_m0;
    std::vector<uint8_t> present((size_t)nc, 1);
    std::vector<double> mod((size_t)nc, 0.0), lower((size_t)nc, 0.0), q;
    std::vector<int32_t> passes;
    std::vector<int64_t> K((size_t)nc, 0);
    LV_TRY(build_level0_csr_sets(c));
    for (int32_t level = 0;; ++level) {
        LV_TRY(batch_level_setup(c, level));
        if (level == 0) {
            LV_TRY(arena_alloc(A, c->d_vcomp0, c->n));
            copy_i32(s, c->d_vcomp0, c->vcomp, c->n);
        }
        LV_TRY(init_level(c));
        if (level == 0) {
            // lower of level 0 = Q of its initial state (LouvainCPP.cpp:356-366)
            LV_TRY(batch_qagg(c, false));
            for (int32_t k = 0; k < nc; ++k)
                if (present[k] && c->comp_e0[k] > 0) lower[k] = batch_q(c, k);
        } else {
            lower = mod;
        }
        LV_TRY(batch_one_level(c, lower, present, q, passes));
        LV_TRY(renumber(c, nullptr));
        LevelRec &L = c->levels.back();
        fill_i64(s, (int64_t *)c->d_ccnt, nc, 0);
        launch_b_groups(s, c->g_start, L.perm, c->vcomp, L.k, c->d_ccnt);
        LV_HIP(hipGetLastError());
        LV_TRY(fetch_small(c, {seg(c->h_ccnt, c->d_ccnt, nc)}));
        L.voff_h = c->comp_voff;
        L.koff_h.assign((size_t)nc + 1, 0);
        for (int32_t k = 0; k < nc; ++k) {
            K[k] = (int64_t)c->h_ccnt[k];
            L.koff_h[k + 1] = L.koff_h[k] + K[k];
        }
        // per-component records and the level-loop decision
        std::vector<int64_t> vnext((size_t)nc, -1), voff_next((size_t)nc + 1, 0);
        int64_t kn = 0;
        for (int32_t k = 0; k < nc; ++k) {
            voff_next[k] = kn;
            if (!present[k]) continue;
            const bool edges = c->comp_e0[k] > 0;
            lv_level_info r{};
            r.n = (int32_t)(c->comp_voff[k + 1] - c->comp_voff[k]);
            r.k = (int32_t)K[k];
            r.nnz = 0;
            r.m_units = edges ? c->comp_m[k] : 0;
            r.passes = edges ? passes[k] : 0;
            r.scale = c->comp_scale[k];
            r.q = edges ? q[k] : 0.0;
            r.lower = edges ? lower[k] : 0.0;
            c->comp_rec[k].push_back(r);
            bool cont = edges;
            if (cont && level > 0 && q[k] - mod[k] < c->cfg.min_gain) cont = false;
            if (cont) mod[k] = q[k];
            if (c->cfg.max_levels > 0 && (int32_t)c->comp_rec[k].size() >= c->cfg.max_levels)
                cont = false;
            present[k] = cont ? 1 : 0;
            if (cont) {
                vnext[k] = kn;
                kn += K[k];
            }
        }
        voff_next[nc] = kn;
        LV_TRY(arena_alloc(A, L.koff, (int64_t)nc + 1));
        LV_TRY(arena_alloc(A, L.vnext, nc));
        LV_TRY(h2d(c, L.koff, (const int64_t *)L.koff_h.data(), (int64_t)nc + 1));
        LV_TRY(h2d(c, L.vnext, (const int64_t *)vnext.data(), nc));
        if (kn == 0) {
            LV_TRY(barrier(c));
            break;
        }
        if (kn > INT32_MAX) return LV_ERR_RANGE;
        int32_t *newid_next = nullptr, *newc_next = nullptr;
        LV_TRY(arena_alloc(A, newid_next, L.k));
        LV_TRY(arena_alloc(A, newc_next, L.n));
        launch_b_next_ids(s, c->g_start, L.perm, c->vcomp, c->g_newid, L.k, L.koff, L.vnext,
                          newid_next);
        launch_cr_newc(s, c->g_head, c->g_hscan, L.perm, newid_next, L.n, newc_next);
        LV_TRY(induce(c, newc_next, newid_next, (int32_t)kn));
        arena_free(A, newid_next);
        arena_free(A, newc_next);
        c->comp_voff = voff_next;
        LV_TRY(barrier(c));  // vnext / voff_next host buffers
    }
    // flatten every component through its own levels
    hipEvent_t t0 = t_begin(c);
    std::vector<int32_t> clev((size_t)nc);
    for (int32_t k = 0; k < nc; ++k) clev[k] = (int32_t)c->comp_rec[k].size();
    int32_t *dclev = nullptr;
    LV_TRY(arena_alloc(A, dclev, nc));
    LV_TRY(h2d(c, dclev, (const int32_t *)clev.data(), nc));
    arena_free(A, c->d_final);
    LV_TRY(arena_alloc(A, c->d_final, c->n0));
    iota_i32(s, c->d_final, c->n0);
    for (size_t l = 0; l < c->levels.size(); ++l)
        launch_b_flatten(s, c->d_final, c->n0, c->d_vcomp0, dclev, (int32_t)l, c->levels[l].newc,
                         c->levels[l].koff, c->levels[l].vnext);
    LV_HIP(hipGetLastError());
    t_end(c, t0, CAT_FLATTEN);
    LV_TRY(barrier(c));
    arena_free(A, dclev);
    c->final_valid = true;
    free_level_arrays(c);
    c->stats.levels += (int32_t)c->levels.size();
    return LV_OK;
}

// ---------------------------------------------------------------------------
// launch-shape tunables: lv_config fields, else the tuning table, else defaults

bool tunables_valid(const lv_config &c)
{
    return (c.hub_q == 0 || (c.hub_q >= 64 && c.hub_q <= kTierMaxQ[kTiers - 2])) &&
           (c.hub_part_target == 0 || (c.hub_part_target >= 64 && c.hub_part_target <= (1 << 30))) &&
           (c.chunk_mul == 0 || (c.chunk_mul >= 1 && c.chunk_mul <= 1024)) && c.upload >= 0 &&
           c.upload <= 2 && c.fanout >= 0 && c.fanout <= 2;
}

struct TuneRow {
    int64_t max_nnz;
    int32_t hub_q, part_target, chunk_mul;
};

// The table: lines "max_nnz hub_q hub_part_target chunk_mul" ('#' comments),
// read once per process from LV_TUNING_FILE or tuning.txt in the directory of
// this library.  Rows with a value out of range are ignored.
const std::vector<TuneRow> &tuning_table()
{
    static std::once_flag once;
    static std::vector<TuneRow> rows;
    std::call_once(once, [] {
        std::string path;
        if (const char *e = getenv("LV_TUNING_FILE")) {
            path = e;
        } else {
            Dl_info info;
            if (dladdr((void *)&tuning_table, &info) && info.dli_fname) {
                path = info.dli_fname;
                const size_t cut = path.find_last_of('/');
                path = (cut == std::string::npos ? std::string(".") : path.substr(0, cut)) +
                       "/tuning.txt";
            }
        }
        FILE *f = path.empty() ? nullptr : fopen(path.c_str(), "r");
        if (!f) return;
        char line[256];
        while (fgets(line, sizeof(line), f)) {
            if (line[0] == '#') continue;
            long long m = 0;
            int q = 0, t = 0, k = 0;
            if (sscanf(line, "%lld %d %d %d", &m, &q, &t, &k) != 4) continue;
            lv_config probe{};
            probe.hub_q = q;
            probe.hub_part_target = t;
            probe.chunk_mul = k;
            if (m <= 0 || q <= 0 || t <= 0 || k <= 0 || !tunables_valid(probe)) continue;
            rows.push_back({m, q, t, k});
        }
        fclose(f);
        std::sort(rows.begin(), rows.end(),
                  [](const TuneRow &a, const TuneRow &b) { return a.max_nnz < b.max_nnz; });
    });
    return rows;
}

// Launch shapes for a graph of nnz directed entries: each configured field, else
// the first table row whose max_nnz covers nnz (the last row beyond it), else
// the built-in default.
void apply_tuning(lv_ctx *c, int64_t nnz)
{
    const std::vector<TuneRow> &t = tuning_table();
    const TuneRow *row = nullptr;
    for (const TuneRow &r : t) {
        row = &r;
        if (nnz <= r.max_nnz) break;
    }
    c->hub_q = c->cfg.hub_q ? c->cfg.hub_q : row ? row->hub_q : kHubQDefault;
    c->hub_part_target = c->cfg.hub_part_target ? c->cfg.hub_part_target
                         : row                  ? row->part_target
                                                : kHubPartTarget;
    c->chunk_mul = c->cfg.chunk_mul ? c->cfg.chunk_mul : row ? row->chunk_mul : kChunkMul;
}

int check_ctx(lv_ctx *c) { return c ? LV_OK : LV_ERR_ARG; }

int set_device(lv_ctx *c)
{
    LV_HIP(hipSetDevice(c->device));
    return LV_OK;
}

// The input is in place: record it as the state every run starts from.
int loaded(lv_ctx *c, int32_t n, int64_t e, int scale)
{
    c->scale = scale;
    c->m0 = c->h_flags->m_units;
    c->max_units0 = c->h_flags->max_units;
    c->n0 = n;
    c->e0 = e;
    c->loaded = true;
    c->load_mark = c->arena.mark();
    c->level = 0;
    c->state = ST_LOADED;
    return LV_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

void lv_config_default(lv_config *cfg)
{
    if (!cfg) return;
    memset(cfg, 0, sizeof(*cfg));
    cfg->gamma = 1.0;
    cfg->min_gain = 1e-7;
    cfg->max_passes = 0;
    cfg->max_levels = 0;
    cfg->device = 0;
    cfg->flags = 0;
    cfg->arena_bytes = 0;
}

const char *lv_strerror(int status)
{
    switch (status) {
    case LV_OK: return "ok";
    case LV_ERR_NOMEM: return "out of device memory (arena exhausted)";
    case LV_ERR_INVALID_INPUT: return "invalid input (unknown or repeated node id, negative or non-finite weight)";
    case LV_ERR_RANGE: return "weight sums exceed the exact int64 range";
    case LV_ERR_ARG: return "invalid argument or call out of order";
    case LV_ERR_DEVICE: return "HIP runtime error";
    case LV_ERR_NO_TRACE: return "trace not recorded (create the context with LV_FLAG_TRACE)";
    case LV_ERR_TOO_LARGE: return "graph too large for the static scratch array of lv_run_static";
    default: return "unknown status";
    }
}

int lv_create(lv_ctx **out, const lv_config *cfg)
{
    if (!out) return LV_ERR_ARG;
    *out = nullptr;
    lv_ctx *c = new (std::nothrow) lv_ctx();
    if (!c) return LV_ERR_NOMEM;
    if (cfg)
        c->cfg = *cfg;
    else
        lv_config_default(&c->cfg);
    memset(&c->stats, 0, sizeof(c->stats));
    if (!(c->cfg.gamma >= 0.0) || !(c->cfg.min_gain == c->cfg.min_gain) || c->cfg.max_passes < 0 ||
        c->cfg.max_levels < 0 || !tunables_valid(c->cfg)) {
        delete c;
        return LV_ERR_ARG;
    }
    c->device = c->cfg.device;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= c->device || c->device < 0) {
        (void)hipGetLastError();
        delete c;
        return LV_ERR_DEVICE;
    }
    int rc = LV_OK;
    do {
        if (hipSetDevice(c->device) != hipSuccess) { rc = LV_ERR_DEVICE; break; }
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { rc = LV_ERR_DEVICE; break; }
        if (hipMalloc(&c->d_flags, sizeof(DevFlags)) != hipSuccess) { rc = LV_ERR_NOMEM; break; }
        if (hipMalloc(&c->d_qacc, 2 * sizeof(QAcc)) != hipSuccess) { rc = LV_ERR_NOMEM; break; }
        if (hipMalloc(&c->d_qpart, sizeof(QAcc) * kQParts) != hipSuccess) { rc = LV_ERR_NOMEM; break; }
        if (hipMalloc(&c->d_small, 64 * sizeof(long long)) != hipSuccess) { rc = LV_ERR_NOMEM; break; }
        if (hipHostMalloc(&c->h_flags, sizeof(DevFlags), hipHostMallocDefault) != hipSuccess) { rc = LV_ERR_NOMEM; break; }
        if (hipHostMalloc(&c->h_qacc, 2 * sizeof(QAcc), hipHostMallocDefault) != hipSuccess) { rc = LV_ERR_NOMEM; break; }
        if (hipHostMalloc(&c->h_small, 64 * sizeof(long long), hipHostMallocDefault) != hipSuccess) { rc = LV_ERR_NOMEM; break; }
        if (hipHostMalloc(&c->h_mbox, kMboxBytes, hipHostMallocCoherent) != hipSuccess) { rc = LV_ERR_NOMEM; break; }
        memset(c->h_mbox, 0, kMboxBytes);
        if (hipHostGetDevicePointer((void **)&c->d_mbox, c->h_mbox, 0) != hipSuccess) { rc = LV_ERR_DEVICE; break; }
        if (c->cfg.arena_bytes) rc = c->arena.reserve(c->cfg.arena_bytes);
    } while (0);
    if (rc != LV_OK) {
        (void)hipGetLastError();
        lv_destroy(c);
        return rc;
    }
    *out = c;
    return LV_OK;
}

int lv_destroy(lv_ctx *c)
{
    if (!c) return LV_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &p : c->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : c->ev_pool) (void)hipEventDestroy(e);
    for (auto e : c->sink_ev) (void)hipEventDestroy(e);
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    for (int i = 0; i < kFanStreams; ++i) {
        if (c->fan[i]) (void)hipStreamSynchronize(c->fan[i]);
        if (c->fan[i]) (void)hipStreamDestroy(c->fan[i]);
        if (c->fan_join[i]) (void)hipEventDestroy(c->fan_join[i]);
    }
    if (c->fan_fork) (void)hipEventDestroy(c->fan_fork);
    if (c->d_flags) (void)hipFree(c->d_flags);
    if (c->d_qacc) (void)hipFree(c->d_qacc);
    if (c->d_qpart) (void)hipFree(c->d_qpart);
    if (c->d_small) (void)hipFree(c->d_small);
    if (c->h_flags) (void)hipHostFree(c->h_flags);
    if (c->h_qacc) (void)hipHostFree(c->h_qacc);
    if (c->h_small) (void)hipHostFree(c->h_small);
    if (c->h_mbox) (void)hipHostFree(c->h_mbox);
    if (c->h_cq) (void)hipHostFree(c->h_cq);
    if (c->h_ccnt) (void)hipHostFree(c->h_ccnt);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;  // the arena destructor frees the HBM block
    return LV_OK;
}

int lv_load_edges(lv_ctx *c, int32_t n, const int32_t *ids, int64_t e, const int32_t *src,
                  const int32_t *dst, const void *w, int wtype)
{
    LV_TRY(check_ctx(c));
    if (n < 0 || e < 0 || (e > 0 && (!src || !dst || !w)) ||
        (wtype != LV_WEIGHT_INT32 && wtype != LV_WEIGHT_FLOAT32))
        return LV_ERR_ARG;
    LV_TRY(set_device(c));
    LV_HIP(hipStreamSynchronize(c->stream));
    reset_run(c);
    free_input(c);
    c->arena.release_all();  // also what an earlier failed call left behind
    c->state = ST_EMPTY;
    apply_tuning(c, 2 * e);
    Arena &A = c->arena;
    hipStream_t s = c->stream;
    // arena: sized for the largest level-0 phase unless the caller fixed it
    if (!c->cfg.arena_bytes) {
        size_t want = (size_t)200 * (size_t)e + (size_t)320 * (size_t)n + ((size_t)256 << 20);
        size_t freeb = 0, totalb = 0;
        if (hipMemGetInfo(&freeb, &totalb) == hipSuccess) {
            size_t avail = (size_t)((double)(freeb + A.capacity()) * 0.92);
            if (want > avail) want = avail;
        }
        if (A.capacity() < want) {
            A.release_all();
            int rc = A.reserve(want);
            if (rc != LV_OK) return rc;
        }
    }
    bool iota_ids = true;
    if (ids)
        for (int32_t i = 0; i < n && iota_ids; ++i) iota_ids = (ids[i] == i);
    LV_TRY(arena_alloc(A, c->d_src, e));
    LV_TRY(arena_alloc(A, c->d_dst, e));
    // int32 weights are their own units: the uploaded array becomes d_w32
    uint32_t *wraw = nullptr;
    LV_TRY(arena_alloc(A, wraw, e));
    LV_HIP(hipMemsetAsync(c->d_flags, 0, sizeof(DevFlags), s));
    // encoded upload (offsets and weight bytes, decoded by the checks) for
    // position ids, unless the context may not allocate (lv_run_static)
    const int mode = c->cfg.upload;
    const bool encode = iota_ids && e > 0 && !c->borrowed && mode != 1 &&
                        (mode == 2 || (e >= kEncodeMinEdges && c->uploader.threads() >= 4));
    UploadBufs ub;
    ub.src = c->d_src;
    ub.dst = c->d_dst;
    ub.w = wraw;
    int64_t *d_off = nullptr, *esc_i = nullptr;
    uint8_t *wpk = nullptr;
    uint32_t *esc_w = nullptr;
    int32_t *tile_v = nullptr;
    if (encode) {
        ub.esc_cap = EdgeUploader::esc_cap(e);
        LV_TRY(arena_alloc(A, d_off, (int64_t)n + 1));
        LV_TRY(arena_alloc(A, wpk, e));
        LV_TRY(arena_alloc(A, esc_i, ub.esc_cap));
        LV_TRY(arena_alloc(A, esc_w, ub.esc_cap));
        LV_TRY(arena_alloc(A, tile_v, (e + kDecTile - 1) / kDecTile + 1));
        ub.off = d_off;
        ub.wpk = wpk;
        ub.esc_i = esc_i;
        ub.esc_w = esc_w;
    }
    UploadPlan plan;
    LV_TRY(c->uploader.upload(s, n, e, src, dst, w, wtype, encode, !c->borrowed, ub, &plan));
    c->upload_bytes = plan.bytes;
    if (!iota_ids) c->upload_bytes += 4 * (int64_t)n;
    if (iota_ids) {
        // one pass: sources decoded, endpoints, weights, and (int32 weights) the units
        LoadedEdges L;
        L.dst = c->d_dst;
        if (plan.src_off) {
            if (plan.off64)
                L.off64 = d_off;
            else
                L.off32 = (const int32_t *)d_off;
            L.tile_v = tile_v;
            L.src_out = c->d_src;
        } else {
            L.src = c->d_src;
        }
        if (plan.wbits) {
            L.wpk = wpk;
            L.wbits = plan.wbits;
            L.esc_i = esc_i;
            L.esc_w = esc_w;
            L.nesc = plan.nesc;
        } else {
            L.w = wraw;
        }
        // weight bytes are widened into the upload's own weight array
        launch_load_checks(s, L, wtype, e, n, plan.wbits ? wraw : nullptr, c->d_flags);
        if (wtype == LV_WEIGHT_INT32) {
            c->d_w32 = wraw;
            LV_TRY(d2h(c, c->h_flags, c->d_flags, 1));
            LV_TRY(sync(c));
            void *tmp[] = {d_off, wpk, esc_i, esc_w, tile_v};
            for (void *p : tmp) A.release(p);
            if (c->h_flags->err) {
                free_input(c);
                return (c->h_flags->err & 1) ? LV_ERR_INVALID_INPUT : LV_ERR_RANGE;
            }
            return loaded(c, n, e, 0);
        }
        void *tmp[] = {d_off, wpk, esc_i, esc_w, tile_v};  // later launches may reuse them
        for (void *p : tmp) A.release(p);
    } else {
        LV_TRY(arena_alloc(A, c->d_ids, n));
        LV_TRY(h2d(c, c->d_ids, ids, n));
        uint32_t *k1 = nullptr, *k2 = nullptr;
        int32_t *v1 = nullptr, *v2 = nullptr;
        LV_TRY(arena_alloc(A, k1, n));
        LV_TRY(arena_alloc(A, k2, n));
        LV_TRY(arena_alloc(A, v1, n));
        LV_TRY(arena_alloc(A, v2, n));
        launch_ids_to_keys(s, c->d_ids, k1, n);
        iota_i32(s, v1, n);
        uint32_t *ko = nullptr;
        int32_t *vo = nullptr;
        LV_TRY(radix_sort_u32_i32(A, s, k1, v1, k2, v2, n, 32, &ko, &vo));
        launch_check_sorted_unique(s, ko, n, c->d_flags);
        launch_map_endpoints(s, ko, vo, n, c->d_src, e, c->d_flags);
        launch_map_endpoints(s, ko, vo, n, c->d_dst, e, c->d_flags);
        LV_TRY(sync(c));
        arena_free(A, k1);
        arena_free(A, k2);
        arena_free(A, v1);
        arena_free(A, v2);
        launch_weight_exponent(s, wraw, wtype, e, c->d_flags);
    }
    LV_HIP(hipGetLastError());
    LV_TRY(d2h(c, c->h_flags, c->d_flags, 1));
    LV_TRY(sync(c));
    if (c->h_flags->err) {
        arena_free(A, wraw);
        free_input(c);
        return (c->h_flags->err & 1) ? LV_ERR_INVALID_INPUT : LV_ERR_RANGE;
    }
    const int scale = c->h_flags->scale;
    if (wtype == LV_WEIGHT_INT32) {
        // their own units: checked and summed in place
        c->d_w32 = wraw;
        launch_weight_units(s, wraw, wtype, e, scale, nullptr, nullptr, c->d_flags);
    } else if (f32_units_host(c->h_flags->max_bits, scale) < (1ull << 32)) {
        LV_TRY(arena_alloc(A, c->d_w32, e));
        launch_weight_units(s, wraw, wtype, e, scale, nullptr, c->d_w32, c->d_flags);
    } else {
        LV_TRY(arena_alloc(A, c->d_w, e));
        launch_weight_units(s, wraw, wtype, e, scale, c->d_w, nullptr, c->d_flags);
    }
    LV_HIP(hipGetLastError());
    LV_TRY(d2h(c, c->h_flags, c->d_flags, 1));
    LV_TRY(sync(c));
    if (wtype != LV_WEIGHT_INT32) arena_free(A, wraw);
    if (c->h_flags->err) {
        free_input(c);
        return (c->h_flags->err & 1) ? LV_ERR_INVALID_INPUT : LV_ERR_RANGE;
    }
    return loaded(c, n, e, scale);
}

int lv_neighbour_compute(lv_ctx *c)
{
    LV_TRY(check_ctx(c));
    LV_TRY(set_device(c));
    if (c->batch) return LV_ERR_ARG;  // graph sets run through lv_batch_run
    if (c->state == ST_LOADED && c->level == 0) {
        if (c->e0 == 0) return LV_ERR_ARG;  // no edges: use lv_run (identity partition)
        LV_TRY(build_level0_csr(c));
        c->state = ST_CSR;
        return LV_OK;
    }
    if (c->state == ST_CSR) return LV_OK;  // lv_induce already produced this level's CSR
    return LV_ERR_ARG;
}

int lv_init(lv_ctx *c)
{
    LV_TRY(check_ctx(c));
    LV_TRY(set_device(c));
    if (c->state != ST_CSR) return LV_ERR_ARG;
    LV_TRY(init_level(c));
    c->state = ST_INIT;
    return LV_OK;
}

int lv_modularity(lv_ctx *c, double *q)
{
    LV_TRY(check_ctx(c));
    LV_TRY(set_device(c));
    if (c->state < ST_INIT || !q) return LV_ERR_ARG;
    int rc = LV_OK;
    *q = current_q(c, rc);
    return rc;
}

int lv_one_level(lv_ctx *c, double lower, double *q, int32_t *passes)
{
    LV_TRY(check_ctx(c));
    LV_TRY(set_device(c));
    if (c->state != ST_INIT) return LV_ERR_ARG;
    LV_TRY(one_level(c, lower, q, passes));
    c->state = ST_LEVEL_DONE;
    return LV_OK;
}

int lv_renumber(lv_ctx *c, int32_t *k)
{
    LV_TRY(check_ctx(c));
    LV_TRY(set_device(c));
    if (c->state != ST_LEVEL_DONE) return LV_ERR_ARG;
    LV_TRY(renumber(c, k));
    c->state = ST_RENUMBERED;
    LV_TRY(flatten_final(c));
    return sync(c);
}

int lv_induce(lv_ctx *c)
{
    LV_TRY(check_ctx(c));
    LV_TRY(set_device(c));
    if (c->state != ST_RENUMBERED || c->batch) return LV_ERR_ARG;
    LV_TRY(induce(c, c->levels.back().newc, c->g_newid, c->levels.back().k));
    c->state = ST_CSR;
    return LV_OK;
}

int lv_run(lv_ctx *c, int32_t *levels_out)
{
    LV_TRY(check_ctx(c));
    if (!c->loaded || c->batch) return LV_ERR_ARG;
    LV_TRY(set_device(c));
    auto t_start = std::chrono::steady_clock::now();
    LV_HIP(hipStreamSynchronize(c->stream));
    reset_run(c);
    if (c->e0 == 0) {
        // identity partition (the reference leaves it undefined,
        // LouvainCUDA/LouvainCPP.cpp:211-230)
        LevelRec L;
        L.n = c->n0;
        L.k = c->n0;
        LV_TRY(arena_alloc(c->arena, L.perm, c->n0));
        LV_TRY(arena_alloc(c->arena, L.newc, c->n0));
        iota_i32(c->stream, L.perm, c->n0);
        iota_i32(c->stream, L.newc, c->n0);
        c->levels.push_back(std::move(L));
        LV_TRY(sink_level(c));
        LV_TRY(flatten_final(c));
        LV_TRY(sync(c));
        c->state = ST_RENUMBERED;
        c->stats.levels += 1;
        if (levels_out) *levels_out = 1;
        c->last_run_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
        return LV_OK;
    }
    // level 0, LouvainCUDA/LouvainCPP.cpp:334-401
    LV_TRY(build_level0_csr(c));
    c->state = ST_CSR;
    LV_TRY(init_level(c));
    c->state = ST_INIT;
    int rc = LV_OK;
    double mod = current_q(c, rc);
    if (rc != LV_OK) return rc;
    double q = 0;
    LV_TRY(one_level(c, mod, &q, nullptr));
    LV_TRY(renumber(c, nullptr));
    LV_TRY(sink_level(c));
    c->state = ST_RENUMBERED;
    mod = q;
    // later levels, :420-507
    for (;;) {
        if (c->cfg.max_levels > 0 && (int32_t)c->levels.size() >= c->cfg.max_levels) break;
        LV_TRY(induce(c, c->levels.back().newc, c->g_newid, c->levels.back().k));
        c->state = ST_CSR;
        LV_TRY(init_level(c));
        c->state = ST_INIT;
        LV_TRY(one_level(c, mod, &q, nullptr));
        LV_TRY(renumber(c, nullptr));
        LV_TRY(sink_level(c));
        c->state = ST_RENUMBERED;
        if (q - mod < c->cfg.min_gain) break;
        mod = q;
    }
    LV_TRY(flatten_final(c));
    LV_TRY(barrier(c));
    // the last level's working arrays are not part of the result
    free_level_arrays(c);
    c->state = ST_DONE;  // finished run: only result queries are valid
    c->stats.levels += (int32_t)c->levels.size();
    if (levels_out) *levels_out = (int32_t)c->levels.size();
    c->last_run_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return LV_OK;
}

int lv_run_partitions(lv_ctx *c, int32_t max_levels, int32_t *const *node, int32_t *const *comm,
                      int32_t *final_comm, int32_t *levels_out)
{
    LV_TRY(check_ctx(c));
    if (max_levels < 0 || (max_levels > 0 && (!node || !comm)) || !final_comm) return LV_ERR_ARG;
    for (int32_t l = 0; l < max_levels; ++l)
        if (!node[l] || !comm[l]) return LV_ERR_ARG;
    if (!c->loaded || c->batch) return LV_ERR_ARG;
    LV_TRY(set_device(c));
    if (!c->copy_stream)
        LV_HIP(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
    PartSink sink;
    sink.max = max_levels;
    sink.node = node;
    sink.comm = comm;
    c->sink = &sink;
    int32_t nl = 0;
    int rc = lv_run(c, &nl);
    c->sink = nullptr;
    if (rc == LV_OK && c->n0 > 0) {
        // the run ended with a sync: the final partition is complete on the device
        const hipError_t e = hipMemcpyAsync(final_comm, c->d_final, sizeof(int32_t) * (size_t)c->n0,
                                            hipMemcpyDeviceToHost, c->copy_stream);
        if (e != hipSuccess) rc = hip_fail(e, "hipMemcpyAsync", __FILE__, __LINE__);
    }
    const hipError_t se = hipStreamSynchronize(c->copy_stream);
    for (void *p : sink.temps) c->arena.release(p);
    if (rc != LV_OK) return rc;
    if (se != hipSuccess) return hip_fail(se, "hipStreamSynchronize", __FILE__, __LINE__);
    if (levels_out) *levels_out = nl;
    return LV_OK;
}

int lv_num_levels(lv_ctx *c, int32_t *levels)
{
    LV_TRY(check_ctx(c));
    if (!levels) return LV_ERR_ARG;
    *levels = (int32_t)c->levels.size();
    return LV_OK;
}

int lv_level_info_get(lv_ctx *c, int32_t level, lv_level_info *out)
{
    LV_TRY(check_ctx(c));
    if (!out || level < 0 || level >= (int32_t)c->levels.size()) return LV_ERR_ARG;
    const LevelRec &L = c->levels[level];
    out->n = L.n;
    out->k = L.k;
    out->nnz = L.nnz;
    out->m_units = L.m;
    out->passes = L.passes;
    out->scale = c->scale;
    out->q = L.q;
    out->lower = L.lower;
    return LV_OK;
}

int lv_level_partition(lv_ctx *c, int32_t level, int32_t *node, int32_t *comm)
{
    LV_TRY(check_ctx(c));
    if (!node || !comm || level < 0 || level >= (int32_t)c->levels.size()) return LV_ERR_ARG;
    LV_TRY(set_device(c));
    const LevelRec &L = c->levels[level];
    if (L.n == 0) return LV_OK;
    int32_t *dn = nullptr, *dc = nullptr;
    LV_TRY(arena_alloc(c->arena, dn, L.n));
    LV_TRY(arena_alloc(c->arena, dc, L.n));
    launch_gather_partition(c->stream, L.perm, level == 0 ? c->d_ids : nullptr, L.newc, L.n, dn, dc);
    LV_HIP(hipGetLastError());
    LV_TRY(fetch(c, node, dn, L.n));
    LV_TRY(fetch(c, comm, dc, L.n));
    arena_free(c->arena, dn);
    arena_free(c->arena, dc);
    return LV_OK;
}

int lv_level_assignment(lv_ctx *c, int32_t level, int32_t *comm)
{
    LV_TRY(check_ctx(c));
    if (!comm || level < 0 || level >= (int32_t)c->levels.size()) return LV_ERR_ARG;
    LV_TRY(set_device(c));
    return fetch(c, comm, c->levels[level].newc, c->levels[level].n);
}

int lv_flatten(lv_ctx *c, int32_t level, int32_t *node, int32_t *comm)
{
    LV_TRY(check_ctx(c));
    if (!node || !comm || level < 0 || level >= (int32_t)c->levels.size()) return LV_ERR_ARG;
    LV_TRY(set_device(c));
    const LevelRec &L0 = c->levels[0];
    if (L0.n == 0) return LV_OK;
    int32_t *comp = nullptr, *dn = nullptr, *dc = nullptr;
    LV_TRY(arena_alloc(c->arena, comp, L0.n));
    copy_i32(c->stream, comp, L0.newc, L0.n);
    for (int32_t l = 1; l <= level; ++l) launch_compose(c->stream, comp, L0.n, c->levels[l].newc);
    LV_TRY(arena_alloc(c->arena, dn, L0.n));
    LV_TRY(arena_alloc(c->arena, dc, L0.n));
    launch_gather_partition(c->stream, L0.perm, c->d_ids, comp, L0.n, dn, dc);
    LV_HIP(hipGetLastError());
    LV_TRY(fetch(c, node, dn, L0.n));
    LV_TRY(fetch(c, comm, dc, L0.n));
    arena_free(c->arena, comp);
    arena_free(c->arena, dn);
    arena_free(c->arena, dc);
    return LV_OK;
}

int lv_final_assignment(lv_ctx *c, int32_t *comm)
{
    LV_TRY(check_ctx(c));
    if (!comm || !c->final_valid) return LV_ERR_ARG;
    LV_TRY(set_device(c));
    return fetch(c, comm, c->d_final, c->n0);
}

int lv_final_modularity(lv_ctx *c, double *q)
{
    LV_TRY(check_ctx(c));
    if (!q || !c->final_valid || c->batch) return LV_ERR_ARG;  // graph sets: per component only
    LV_TRY(set_device(c));
    if (c->e0 == 0 || c->m0 == 0) {
        *q = 0.0;
        return LV_OK;
    }
    ull *lc = nullptr, *dc = nullptr;
    LV_TRY(arena_alloc(c->arena, lc, c->n0));
    LV_TRY(arena_alloc(c->arena, dc, c->n0));
    fill_i64(c->stream, (int64_t *)lc, c->n0, 0);
    fill_i64(c->stream, (int64_t *)dc, c->n0, 0);
    launch_final_q(c->stream, c->d_src, c->d_dst, wunits(c), c->e0, c->d_final, lc, dc);
    launch_qagg_ull(c->stream, c->n0, dc, lc, c->d_qpart, c->d_qacc);
    LV_HIP(hipGetLastError());
    LV_TRY(d2h(c, c->h_qacc, c->d_qacc, 1));
    LV_TRY(sync(c));
    arena_free(c->arena, lc);
    arena_free(c->arena, dc);
    *q = q_from_aggregates(c->h_qacc->sum_in, c->h_qacc->d2_lo, c->h_qacc->d2_hi, c->m0,
                           c->cfg.gamma);
    return LV_OK;
}

int lv_cluster_layout(lv_ctx *c, int32_t n, const int32_t *comm, int32_t ncomm,
                      int32_t *cluster_end, int32_t *index_in_cluster, float *x, float *y)
{
    LV_TRY(check_ctx(c));
    if (n < 0 || ncomm < 0 || (n > 0 && (!comm || !index_in_cluster || !x || !y)) ||
        (ncomm > 0 && !cluster_end))
        return LV_ERR_ARG;
    LV_TRY(set_device(c));
    if (n == 0) {
        for (int32_t k = 0; k < ncomm; ++k) cluster_end[k] = 0;
        return LV_OK;
    }
    if (ncomm == 0) return LV_ERR_INVALID_INPUT;
    Arena &A = c->arena;
    hipStream_t s = c->stream;
    if (A.capacity() == 0) LV_TRY(A.reserve((size_t)64 * (size_t)n + (size_t)16 * (size_t)ncomm + (64u << 20)));
    uint32_t *k1 = nullptr, *k2 = nullptr;
    int32_t *v1 = nullptr, *v2 = nullptr, *first = nullptr, *count = nullptr, *idx = nullptr;
    int64_t *cnt64 = nullptr, *end64 = nullptr;
    float *dx = nullptr, *dy = nullptr;
    LV_TRY(arena_alloc(A, k1, n));
    LV_TRY(arena_alloc(A, k2, n));
    LV_TRY(arena_alloc(A, v1, n));
    LV_TRY(arena_alloc(A, v2, n));
    LV_TRY(arena_alloc(A, first, ncomm));
    LV_TRY(arena_alloc(A, count, ncomm));
    LV_TRY(arena_alloc(A, idx, n));
    LV_TRY(arena_alloc(A, cnt64, ncomm));
    LV_TRY(arena_alloc(A, end64, (int64_t)ncomm + 1));
    LV_TRY(arena_alloc(A, dx, n));
    LV_TRY(arena_alloc(A, dy, n));
    LV_TRY(h2d(c, (int32_t *)k1, comm, n));
    iota_i32(s, v1, n);
    fill_i32(s, count, ncomm, 0);
    LV_HIP(hipMemsetAsync(c->d_flags, 0, sizeof(DevFlags), s));
    uint32_t *sk = nullptr;
    int32_t *perm = nullptr;
    LV_TRY(radix_sort_u32_i32(A, s, k1, v1, k2, v2, n, 32, &sk, &perm));
    launch_layout(s, sk, perm, n, ncomm, first, count, idx, dx, dy, c->d_flags);
    launch_widen_i32(s, count, ncomm, cnt64);
    LV_TRY(scan_exclusive_i64(A, s, cnt64, end64, ncomm, end64 + ncomm));
    LV_HIP(hipGetLastError());
    LV_TRY(d2h(c, c->h_flags, c->d_flags, 1));
    LV_TRY(sync(c));
    int rc = c->h_flags->err ? LV_ERR_INVALID_INPUT : LV_OK;
    if (rc == LV_OK) {
        std::vector<int64_t> e64((size_t)ncomm + 1);
        LV_TRY(d2h(c, e64.data(), (const int64_t *)end64, (int64_t)ncomm + 1));
        LV_TRY(d2h(c, index_in_cluster, idx, n));
        LV_TRY(d2h(c, x, dx, n));
        LV_TRY(d2h(c, y, dy, n));
        LV_TRY(sync(c));
        for (int32_t k = 0; k < ncomm; ++k) cluster_end[k] = (int32_t)e64[k + 1];  // inclusive
    }
    void *tmp[] = {k1, k2, v1, v2, first, count, idx, cnt64, end64, dx, dy};
    for (void *p : tmp) A.release(p);
    return rc;
}

int lv_trace_next(lv_ctx *c, int32_t level, int32_t pass, int32_t *next)
{
    LV_TRY(check_ctx(c));
    if (!next || level < 0 || level >= (int32_t)c->levels.size()) return LV_ERR_ARG;
    if (!(c->cfg.flags & LV_FLAG_TRACE)) return LV_ERR_NO_TRACE;
    const LevelRec &L = c->levels[level];
    if (pass < 1 || pass > (int32_t)L.trace.size()) return LV_ERR_ARG;
    memcpy(next, L.trace[pass - 1].data(), sizeof(int32_t) * (size_t)L.n);
    return LV_OK;
}

int lv_trace_pass_q(lv_ctx *c, int32_t level, int32_t pass, double *q)
{
    LV_TRY(check_ctx(c));
    if (!q || level < 0 || level >= (int32_t)c->levels.size()) return LV_ERR_ARG;
    const LevelRec &L = c->levels[level];
    if (pass < 1 || pass > (int32_t)L.pass_q.size()) return LV_ERR_ARG;
    *q = L.pass_q[pass - 1];
    return LV_OK;
}

int lv_level_csr(lv_ctx *c, int64_t *rowptr, int32_t *col, int64_t *w)
{
    LV_TRY(check_ctx(c));
    if (c->state < ST_CSR || c->state > ST_RENUMBERED || !c->rowptr) return LV_ERR_ARG;
    if (!rowptr || (c->nnz > 0 && (!col || !w))) return LV_ERR_ARG;
    LV_TRY(set_device(c));
    if (c->rowend_buf) {
        // a coarsened level: its rows keep their segments' places; packed here
        const int64_t n = c->n, sp = c->span;
        std::vector<int64_t> rb((size_t)n), re((size_t)n), ww((size_t)sp);
        std::vector<int32_t> cc((size_t)sp);
        std::vector<uint32_t> w32(c->cw32 ? (size_t)sp : 0);
        LV_TRY(d2h(c, rb.data(), c->rowptr, n));
        LV_TRY(d2h(c, re.data(), c->rowend_buf, n));
        LV_TRY(d2h(c, cc.data(), c->col, sp));
        if (c->cw32)
            LV_TRY(d2h(c, w32.data(), c->cw32, sp));
        else
            LV_TRY(d2h(c, ww.data(), c->cw, sp));
        LV_TRY(sync(c));
        int64_t o = 0;
        for (int64_t v = 0; v < n; ++v) {
            rowptr[v] = o;
            for (int64_t p = rb[(size_t)v]; p < re[(size_t)v]; ++p, ++o) {
                col[o] = cc[(size_t)p];
                w[o] = c->cw32 ? (int64_t)w32[(size_t)p] : ww[(size_t)p];
            }
        }
        rowptr[n] = o;
        return o == c->nnz ? LV_OK : LV_ERR_DEVICE;
    }
    LV_TRY(d2h(c, rowptr, c->rowptr, (int64_t)c->n + 1));
    LV_TRY(d2h(c, col, c->col, c->nnz));
    if (c->cpk) {  // one word per entry: split on the host
        LV_TRY(sync(c));
        const uint32_t mask = (1u << c->cpk) - 1u;
        for (int64_t p = 0; p < c->nnz; ++p) {
            const uint32_t x = (uint32_t)col[p];
            col[p] = (int32_t)(x >> c->cpk);
            w[p] = (int64_t)(x & mask);
        }
        return LV_OK;
    }
    if (c->cw32) {  // widened on the host
        std::vector<uint32_t> w32((size_t)c->nnz);
        LV_TRY(d2h(c, w32.data(), c->cw32, c->nnz));
        LV_TRY(sync(c));
        for (int64_t p = 0; p < c->nnz; ++p) w[p] = (int64_t)w32[(size_t)p];
        return LV_OK;
    }
    LV_TRY(d2h(c, w, c->cw, c->nnz));
    return sync(c);
}

int lv_level_vertex(lv_ctx *c, int64_t *k, int64_t *loop, int64_t *selfsum)
{
    LV_TRY(check_ctx(c));
    if (c->state < ST_CSR || c->state > ST_RENUMBERED || !c->kdeg) return LV_ERR_ARG;
    if (!k || !loop || !selfsum) return LV_ERR_ARG;
    LV_TRY(set_device(c));
    LV_TRY(d2h(c, k, c->kdeg, c->n));
    LV_TRY(d2h(c, loop, c->loopw, c->n));
    LV_TRY(d2h(c, selfsum, c->selfsum, c->n));
    return sync(c);
}

int lv_current_level(lv_ctx *c, int32_t *level, int32_t *n, int64_t *nnz)
{
    LV_TRY(check_ctx(c));
    if (level) *level = c->level;
    if (n) *n = (c->state >= ST_CSR && c->state <= ST_RENUMBERED) ? c->n : c->n0;
    if (nnz) *nnz = (c->state >= ST_CSR && c->state <= ST_RENUMBERED) ? c->nnz : 0;
    return LV_OK;
}

int lv_get_stats(lv_ctx *c, lv_stats *out)
{
    LV_TRY(check_ctx(c));
    if (!out) return LV_ERR_ARG;
    LV_TRY(set_device(c));
    LV_TRY(sync(c));
    *out = c->stats;
    out->arena_capacity = (uint64_t)c->arena.capacity();
    out->arena_high_water = (uint64_t)c->arena.high_water();
    out->upload_bytes = (uint64_t)c->upload_bytes;
    return LV_OK;
}

int lv_reset_stats(lv_ctx *c)
{
    LV_TRY(check_ctx(c));
    LV_TRY(sync(c));
    memset(&c->stats, 0, sizeof(c->stats));
    return LV_OK;
}

int lv_last_run_ms(lv_ctx *c, double *ms)
{
    LV_TRY(check_ctx(c));
    if (!ms) return LV_ERR_ARG;
    *ms = c->last_run_ms;
    return LV_OK;
}

// ---------------------------------------------------------------------------
// graph sets

}  // extern "C"

namespace {

// Uploads the edge lists of graphs given by per-graph pointers into the base
// arrays at beoff: graphs of at least kDirect edges are copied straight from
// the caller's memory; runs of smaller graphs are packed into a pinned staging
// buffer first, so a set of many small graphs costs a few DMA transfers.
int upload_graphs(lv_ctx *c, int32_t ng, const int64_t *e, const std::vector<int64_t> &beoff,
                  const int32_t *const *srcs, const int32_t *const *dsts, const void *const *ws,
                  int32_t *bsrc, int32_t *bdst, uint32_t *bw)
{
    // each union array is its graphs' arrays end to end: one staged stream per
    // array, whatever the mix of small and large graphs
    for (int32_t g = 0; g < ng; ++g)
        if (e[g] > 0 && (!srcs[g] || !dsts[g] || !ws[g])) return LV_ERR_ARG;
    std::vector<int64_t> boff((size_t)ng + 1);
    for (int32_t g = 0; g <= ng; ++g) boff[g] = 4 * beoff[g];
    LV_TRY(c->uploader.upload_concat(c->stream, bsrc, ng, (const void *const *)srcs, boff.data()));
    LV_TRY(c->uploader.upload_concat(c->stream, bdst, ng, (const void *const *)dsts, boff.data()));
    LV_TRY(c->uploader.upload_concat(c->stream, bw, ng, ws, boff.data()));
    return LV_OK;
}

// lv_batch_load / lv_batch_load_graphs: either one contiguous edge list
// (src/dst/w) or per-graph pointers (srcs/dsts/ws).
int batch_load(lv_ctx *c, int32_t ngraphs, const int32_t *n, const int64_t *e, const int32_t *src,
               const int32_t *dst, const void *w, const int32_t *const *srcs,
               const int32_t *const *dsts, const void *const *ws, int wtype, int32_t nres,
               const double *gamma)
{
    LV_TRY(check_ctx(c));
    if (ngraphs < 1 || nres < 1 || !n || !e || !gamma ||
        (wtype != LV_WEIGHT_INT32 && wtype != LV_WEIGHT_FLOAT32))
        return LV_ERR_ARG;
    if ((int64_t)ngraphs * nres > INT32_MAX) return LV_ERR_ARG;
    int64_t nb = 0, eb = 0;
    for (int32_t g = 0; g < ngraphs; ++g) {
        if (n[g] < 0 || e[g] < 0) return LV_ERR_ARG;
        nb += n[g];
        eb += e[g];
    }
    for (int32_t r = 0; r < nres; ++r)
        if (!(gamma[r] >= 0.0)) return LV_ERR_ARG;
    if (eb > 0 && !srcs && (!src || !dst || !w)) return LV_ERR_ARG;
    if (eb > 0 && srcs && (!dsts || !ws)) return LV_ERR_ARG;
    const int64_t nu = nb * nres, eu = eb * nres;
    if (nu > INT32_MAX) return LV_ERR_RANGE;
    LV_TRY(set_device(c));
    LV_HIP(hipStreamSynchronize(c->stream));
    reset_run(c);
    free_input(c);
    c->arena.release_all();  // also what an earlier failed call left behind
    c->state = ST_EMPTY;
    Arena &A = c->arena;
    hipStream_t s = c->stream;
    const int32_t nc = ngraphs * nres;
    apply_tuning(c, 2 * eu);
    if (!c->cfg.arena_bytes) {
        size_t want = (size_t)200 * (size_t)eu + (size_t)320 * (size_t)nu + (size_t)16 * (size_t)eb +
                      (size_t)96 * (size_t)nc + ((size_t)256 << 20);
        size_t freeb = 0, totalb = 0;
        if (hipMemGetInfo(&freeb, &totalb) == hipSuccess) {
            size_t avail = (size_t)((double)(freeb + A.capacity()) * 0.92);
            if (want > avail) want = avail;
        }
        if (A.capacity() < want) {
            A.release_all();
            int rc = A.reserve(want);
            if (rc != LV_OK) return rc;
        }
    }
    // offsets of the graphs (edges, vertices); component k is graph k / nres
    std::vector<int64_t> beoff((size_t)ngraphs + 1, 0), bvoff((size_t)ngraphs + 1, 0);
    for (int32_t g = 0; g < ngraphs; ++g) {
        beoff[g + 1] = beoff[g] + e[g];
        bvoff[g + 1] = bvoff[g] + n[g];
    }
    c->comp_n0.assign((size_t)nc, 0);
    c->comp_e0.assign((size_t)nc, 0);
    c->comp_gamma.assign((size_t)nc, 1.0);
    for (int32_t k = 0; k < nc; ++k) {
        const int32_t g = k / nres;
        c->comp_n0[k] = n[g];
        c->comp_e0[k] = e[g];
        c->comp_gamma[k] = gamma[k % nres];
    }
    int32_t *bsrc = nullptr, *bdst = nullptr, *bn = nullptr;
    uint32_t *bw = nullptr;
    int *gscale = nullptr;
    int64_t *dbeoff = nullptr;
    LV_TRY(arena_alloc(A, bsrc, eb));
    LV_TRY(arena_alloc(A, bdst, eb));
    LV_TRY(arena_alloc(A, bw, eb));
    LV_TRY(arena_alloc(A, bn, ngraphs));
    LV_TRY(arena_alloc(A, gscale, ngraphs));
    LV_TRY(arena_alloc(A, dbeoff, (int64_t)ngraphs + 1));
    LV_TRY(arena_alloc(A, c->d_bvoff, (int64_t)ngraphs + 1));
    LV_TRY(arena_alloc(A, c->d_src, eb));
    LV_TRY(arena_alloc(A, c->d_dst, eb));
    LV_TRY(arena_alloc(A, c->d_w, eb));
    if (srcs) {
        LV_TRY(upload_graphs(c, ngraphs, e, beoff, srcs, dsts, ws, bsrc, bdst, bw));
    } else {
        const int64_t boff[2] = {0, 4 * eb};
        const void *hs[1] = {src}, *hd[1] = {dst}, *hw[1] = {w};
        LV_TRY(c->uploader.upload_concat(s, bsrc, 1, hs, boff));
        LV_TRY(c->uploader.upload_concat(s, bdst, 1, hd, boff));
        LV_TRY(c->uploader.upload_concat(s, bw, 1, hw, boff));
    }
    LV_TRY(h2d(c, bn, n, ngraphs));
    LV_TRY(h2d(c, dbeoff, (const int64_t *)beoff.data(), (int64_t)ngraphs + 1));
    LV_TRY(h2d(c, c->d_bvoff, (const int64_t *)bvoff.data(), (int64_t)ngraphs + 1));
    LV_HIP(hipMemsetAsync(c->d_flags, 0, sizeof(DevFlags), s));
    fill_i32(s, gscale, ngraphs, 0);
    launch_b_graph_exponent(s, bw, wtype, dbeoff, ngraphs, eb, gscale, c->d_flags);
    // the graphs once, in their own units, with vertex ids offset by graph
    launch_b_union_edges(s, bsrc, bdst, bw, wtype, gscale, dbeoff, bn, dbeoff, c->d_bvoff,
                         ngraphs, 1, eb, c->d_src, c->d_dst, c->d_w, c->d_flags);
    LV_HIP(hipGetLastError());
    std::vector<int> gs((size_t)ngraphs, 0);
    LV_TRY(d2h(c, gs.data(), gscale, ngraphs));
    LV_TRY(d2h(c, c->h_flags, c->d_flags, 1));
    LV_TRY(sync(c));  // the host offset vectors are read by the copies above
    arena_free(A, bsrc);
    arena_free(A, bdst);
    arena_free(A, bw);
    arena_free(A, bn);
    arena_free(A, gscale);
    if (c->h_flags->err) {
        arena_free(A, dbeoff);
        free_input(c);
        return (c->h_flags->err & 1) ? LV_ERR_INVALID_INPUT : LV_ERR_RANGE;
    }
    int scale = 0;
    c->comp_scale.assign((size_t)nc, 0);
    for (int32_t k = 0; k < nc; ++k) {
        c->comp_scale[k] = gs[k / nres];
        scale = std::max(scale, c->comp_scale[k]);
    }
    // m of every graph (exact units), that of each of its components
    LV_TRY(pinned_cap(c, nc));
    ull *dcm = nullptr;
    LV_TRY(arena_alloc(A, dcm, ngraphs));
    fill_i64(s, (int64_t *)dcm, ngraphs, 0);
    launch_b_edge_sum(s, c->d_w, dbeoff, ngraphs, eb, dcm);
    LV_HIP(hipGetLastError());
    LV_TRY(d2h(c, c->h_ccnt, dcm, ngraphs));
    LV_TRY(sync(c));
    c->comp_m0.assign((size_t)nc, 0);
    for (int32_t k = 0; k < nc; ++k) c->comp_m0[k] = (int64_t)c->h_ccnt[k / nres];
    arena_free(A, dcm);
    arena_free(A, dbeoff);
    c->ngraphs = ngraphs;
    c->base_n0 = (int32_t)nb;
    c->base_e0 = eb;
    c->scale = scale;
    c->m0 = 0;
    for (int32_t k = 0; k < nc; ++k) c->m0 += c->comp_m0[k];
    c->max_units0 = c->h_flags->max_units;
    c->n0 = (int32_t)nu;
    c->e0 = eu;
    c->ncomp = nc;
    c->nres = nres;
    c->batch = true;
    c->loaded = true;
    c->load_mark = c->arena.mark();
    c->level = 0;
    c->state = ST_LOADED;
    return LV_OK;
}

}  // namespace

extern "C" {

int lv_batch_load(lv_ctx *c, int32_t ngraphs, const int32_t *n, const int64_t *e,
                  const int32_t *src, const int32_t *dst, const void *w, int wtype, int32_t nres,
                  const double *gamma)
{
    return batch_load(c, ngraphs, n, e, src, dst, w, nullptr, nullptr, nullptr, wtype, nres,
                      gamma);
}

int lv_batch_load_graphs(lv_ctx *c, int32_t ngraphs, const int32_t *n, const int64_t *e,
                         const int32_t *const *src, const int32_t *const *dst,
                         const void *const *w, int wtype, int32_t nres, const double *gamma)
{
    if (!src) return LV_ERR_ARG;
    return batch_load(c, ngraphs, n, e, nullptr, nullptr, nullptr, src, dst, w, wtype, nres,
                      gamma);
}

int lv_batch_run(lv_ctx *c)
{
    LV_TRY(check_ctx(c));
    if (!c->loaded || !c->batch) return LV_ERR_ARG;
    LV_TRY(set_device(c));
    auto t_start = std::chrono::steady_clock::now();
    LV_HIP(hipStreamSynchronize(c->stream));
    reset_run(c);
    if (c->e0 == 0) {
        // no component has edges: every one is its identity partition
        const int32_t nc = c->ncomp;
        c->comp_rec.assign((size_t)nc, {});
        std::vector<int32_t> fin((size_t)c->n0);
        int64_t off = 0;
        for (int32_t k = 0; k < nc; ++k) {
            lv_level_info r{};
            r.n = r.k = c->comp_n0[k];
            r.scale = c->comp_scale[k];
            c->comp_rec[k].push_back(r);
            for (int32_t i = 0; i < c->comp_n0[k]; ++i) fin[(size_t)(off + i)] = i;
            off += c->comp_n0[k];
        }
        LV_TRY(arena_alloc(c->arena, c->d_final, c->n0));
        LV_TRY(h2d(c, c->d_final, (const int32_t *)fin.data(), c->n0));
        LV_TRY(sync(c));
        c->final_valid = true;
    } else {
        LV_TRY(batch_run(c));
    }
    c->state = ST_DONE;
    c->last_run_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return LV_OK;
}

int lv_batch_num_components(lv_ctx *c, int32_t *ncomp)
{
    LV_TRY(check_ctx(c));
    if (!ncomp || !c->batch) return LV_ERR_ARG;
    *ncomp = c->ncomp;
    return LV_OK;
}

int lv_batch_num_levels(lv_ctx *c, int32_t comp, int32_t *levels)
{
    LV_TRY(check_ctx(c));
    if (!levels || !c->batch || !c->final_valid || comp < 0 || comp >= c->ncomp) return LV_ERR_ARG;
    *levels = (int32_t)c->comp_rec[comp].size();
    return LV_OK;
}

int lv_batch_level_info(lv_ctx *c, int32_t comp, int32_t level, lv_level_info *out)
{
    LV_TRY(check_ctx(c));
    if (!out || !c->batch || !c->final_valid || comp < 0 || comp >= c->ncomp) return LV_ERR_ARG;
    if (level < 0 || level >= (int32_t)c->comp_rec[comp].size()) return LV_ERR_ARG;
    *out = c->comp_rec[comp][level];
    return LV_OK;
}

int lv_batch_level_assignment(lv_ctx *c, int32_t comp, int32_t level, int32_t *comm)
{
    LV_TRY(check_ctx(c));
    if (!comm || !c->batch || !c->final_valid || comp < 0 || comp >= c->ncomp) return LV_ERR_ARG;
    if (level < 0 || level >= (int32_t)c->comp_rec[comp].size()) return LV_ERR_ARG;
    LV_TRY(set_device(c));
    const int32_t nv = c->comp_rec[comp][level].n;
    if (nv == 0) return LV_OK;
    if (c->levels.empty()) {  // edgeless set: identity
        for (int32_t i = 0; i < nv; ++i) comm[i] = i;
        return LV_OK;
    }
    const LevelRec &L = c->levels[level];
    LV_TRY(fetch(c, comm, L.newc + L.voff_h[comp], nv));
    const int32_t ko = (int32_t)L.koff_h[comp];
    for (int32_t i = 0; i < nv; ++i) comm[i] -= ko;
    return LV_OK;
}

int lv_batch_final_assignment(lv_ctx *c, int32_t comp, int32_t *comm)
{
    LV_TRY(check_ctx(c));
    if (!comm || !c->batch || !c->final_valid || comp < -1 || comp >= c->ncomp) return LV_ERR_ARG;
    LV_TRY(set_device(c));
    if (comp < 0) return fetch(c, comm, c->d_final, c->n0);
    int64_t off = 0;
    for (int32_t k = 0; k < comp; ++k) off += c->comp_n0[k];
    return fetch(c, comm, c->d_final + off, c->comp_n0[comp]);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// lv_run_static: the single-graph run with every device buffer carved out of
// one file-scope array, for callers that cannot allocate device memory.

// Each device that loads this library holds its own copy of the array until
// the process ends.  In host code `scratch` names a host-side stand-in: its
// device address comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) unsigned char scratch[LV_STATIC_SCRATCH_BYTES];

namespace {

constexpr int kStaticMaxDevices = 64;
// Head of the array: the context's device scalars; the arena takes the rest.
constexpr size_t kStaticFlagsOff = 0;
constexpr size_t kStaticQaccOff = 256;
constexpr size_t kStaticQpartOff = 512;
constexpr size_t kStaticSmallOff = kStaticQpartOff + sizeof(QAcc) * kQParts;
constexpr size_t kStaticHead = 64 << 10;
static_assert(kStaticSmallOff + 64 * sizeof(long long) <= kStaticHead, "static head");
static_assert(sizeof(DevFlags) <= kStaticQaccOff - kStaticFlagsOff, "static flags");

std::mutex g_static_mu[kStaticMaxDevices];  // one run at a time per device
void *g_static_ptr[kStaticMaxDevices];      // device address of `scratch` per device

// Device address of this device's copy of the array (the device is current).
int static_scratch(int32_t device, char **out)
{
    if (!g_static_ptr[device]) {
        void *p = nullptr;
        LV_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return LV_ERR_DEVICE;
        g_static_ptr[device] = p;
    }
    *out = (char *)g_static_ptr[device];
    return LV_OK;
}

int static_run(lv_ctx *c, int32_t n, const int32_t *ids, int64_t e, const int32_t *src,
               const int32_t *dst, const void *w, int wtype, int32_t *final_comm,
               int32_t *levels, double *modularity)
{
    int32_t nl = 0;
    LV_TRY(lv_load_edges(c, n, ids, e, src, dst, w, wtype));
    LV_TRY(lv_run(c, &nl));
    LV_TRY(lv_final_assignment(c, final_comm));
    if (modularity) LV_TRY(lv_final_modularity(c, modularity));
    if (levels) *levels = nl;
    return LV_OK;
}

}  // namespace

extern "C" {

int lv_run_static(const lv_config *cfg, void *stream, int32_t n, const int32_t *ids, int64_t e,
                  const int32_t *src, const int32_t *dst, const void *w, int wtype,
                  int32_t *final_comm, int32_t *levels, double *modularity)
{
    lv_config cf;
    if (cfg)
        cf = *cfg;
    else
        lv_config_default(&cf);
    if (n < 0 || e < 0 || (e > 0 && (!src || !dst || !w)) || (n > 0 && !final_comm) ||
        (wtype != LV_WEIGHT_INT32 && wtype != LV_WEIGHT_FLOAT32) || cf.flags != 0 ||
        !(cf.gamma >= 0.0) || !(cf.min_gain == cf.min_gain) || cf.max_passes < 0 ||
        cf.max_levels < 0 || cf.device < 0 || cf.device >= kStaticMaxDevices ||
        !tunables_valid(cf))
        return LV_ERR_ARG;
    // the footprint bound of the arena sizing (200 B per edge, 320 B per vertex)
    const size_t room = (size_t)LV_STATIC_SCRATCH_BYTES - kStaticHead;
    if ((double)e * 200.0 + (double)n * 320.0 > (double)room) return LV_ERR_TOO_LARGE;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || cf.device >= ndev) {
        (void)hipGetLastError();
        return LV_ERR_DEVICE;
    }
    std::lock_guard<std::mutex> lock(g_static_mu[cf.device]);
    LV_HIP(hipSetDevice(cf.device));
    char *base = nullptr;
    LV_TRY(static_scratch(cf.device, &base));
    lv_ctx *c = new (std::nothrow) lv_ctx();
    if (!c) return LV_ERR_NOMEM;
    c->cfg = cf;
    c->cfg.arena_bytes = room;  // fixed: lv_load_edges does not size the arena
    c->device = cf.device;
    c->stream = (hipStream_t)stream;
    c->borrowed = true;
    memset(&c->stats, 0, sizeof(c->stats));
    c->d_flags = (DevFlags *)(base + kStaticFlagsOff);
    c->d_qacc = (QAcc *)(base + kStaticQaccOff);
    c->d_qpart = (QAcc *)(base + kStaticQpartOff);
    c->d_small = (long long *)(base + kStaticSmallOff);
    c->h_flags = &c->hs_flags;
    c->h_qacc = &c->hs_qacc;
    c->h_small = c->hs_small;
    c->arena.adopt(base + kStaticHead, room);
    const int rc = static_run(c, n, ids, e, src, dst, w, wtype, final_comm, levels, modularity);
    // leave nothing in flight on the caller's stream that reads the array
    const hipError_t se = hipStreamSynchronize(c->stream);
    delete c;  // host bookkeeping only: the arena does not own its block
    if (rc != LV_OK) return rc;
    return se == hipSuccess ? LV_OK : hip_fail(se, "hipStreamSynchronize", __FILE__, __LINE__);
}

}  // extern "C"
