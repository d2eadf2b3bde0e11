// upload.hip - the edge-list upload of lv_load_edges and its device checks.
//
// The reference copies src, dst and w to the device as they are
// (LouvainCUDA/LouvainCPP.cpp:283-287): 12 bytes per edge over PCIe, which at
// R-MAT scale 22 is half of the end-to-end time.  Here the copy of dst starts
// at once while host threads encode the rest (encode.cpp): int weights are
// packed 2, 4 or 8 bits each (the others go to an escape list), and a
// source-sorted src becomes its row offsets (n + 1 entries).  The device
// decodes both inside the pass that checks the endpoints and weights, so
// everything after the load sees the caller's arrays.
#include <sched.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstring>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>

#include "device.hpp"
#include "encode.hpp"

namespace lv {

// ---------------------------------------------------------------------------
// host threads

namespace {

int usable_cpus()
{
    cpu_set_t set;
    CPU_ZERO(&set);
    int n = 0;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) n = CPU_COUNT(&set);
    if (n <= 0) n = (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(n, 64));
}

}  // namespace

// Fixed workers that run one job at a time: run(fn) calls fn(t) for every
// t in [0, size()), t = 0 on the calling thread, and returns when all are done.
class HostPool {
  public:
    explicit HostPool(int n) : n_(n)
    {
        for (int t = 1; t < n_; ++t) th_.emplace_back([this, t] { loop(t); });
    }
    ~HostPool()
    {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    int size() const { return n_; }
    void run(const std::function<void(int)> &fn)
    {
        if (n_ > 1) {
            std::lock_guard<std::mutex> g(m_);
            job_ = &fn;
            left_ = n_ - 1;
            ++gen_;
        }
        cv_.notify_all();
        fn(0);
        if (n_ > 1) {
            std::unique_lock<std::mutex> g(m_);
            done_.wait(g, [this] { return left_ == 0; });
            job_ = nullptr;
        }
    }

  private:
    void loop(int t)
    {
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(int)> *job;
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                job = job_;
            }
            (*job)(t);
            std::lock_guard<std::mutex> g(m_);
            if (--left_ == 0) done_.notify_one();
        }
    }
    int n_;
    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    const std::function<void(int)> *job_ = nullptr;
    uint64_t gen_ = 0;
    int left_ = 0;
    bool stop_ = false;
};

EdgeUploader::~EdgeUploader()
{
    if (pool_pid_ == getpid()) delete pool_;  // (a forked child's copy has no threads to join)
    if (h_wpk_) (void)hipHostFree(h_wpk_);
    if (h_off_) (void)hipHostFree(h_off_);
    if (h_esc_) (void)hipHostFree(h_esc_);
    for (int b = 0; b < 2; ++b) {
        if (ring_ev_[b]) {
            (void)hipEventSynchronize(ring_ev_[b]);  // no copy still reads the buffer
            (void)hipEventDestroy(ring_ev_[b]);
        }
        if (h_ring_[b]) (void)hipHostFree(h_ring_[b]);
    }
}

int EdgeUploader::ring()
{
    for (int b = 0; b < 2; ++b) {
        if (!ring_ev_[b] && hipEventCreateWithFlags(&ring_ev_[b], hipEventDisableTiming) != hipSuccess) {
            ring_ev_[b] = nullptr;
            (void)hipGetLastError();
            return LV_ERR_DEVICE;
        }
        if (!h_ring_[b] && hipHostMalloc(&h_ring_[b], (size_t)kRingBytes, hipHostMallocDefault) != hipSuccess) {
            h_ring_[b] = nullptr;
            (void)hipGetLastError();
            return LV_ERR_NOMEM;
        }
    }
    return LV_OK;
}

int EdgeUploader::upload_concat(hipStream_t s, void *dev, int32_t np, const void *const *host,
                                const int64_t *off)
{
    const int64_t total = off[np];
    if (total <= 0) return LV_OK;
    LV_TRY(ring());
    const int T = threads();
    int64_t k = 0;
    for (int64_t c0 = 0; c0 < total; c0 += kRingBytes, ++k) {
        const int64_t len = std::min(kRingBytes, total - c0);
        const int b = (int)(k & 1);
        LV_HIP(hipEventSynchronize(ring_ev_[b]));  // its previous copy has left the buffer
        char *buf = (char *)h_ring_[b];
        // bytes [a, z) of this chunk, piece by piece (off[i] <= pos < off[i + 1],
        // so i < np; empty pieces are stepped over)
        auto fill = [&](int64_t a, int64_t z) {
            int64_t pos = c0 + a;
            const int64_t end = c0 + z;
            int32_t i = (int32_t)(std::upper_bound(off, off + np + 1, pos) - off) - 1;
            while (pos < end) {
                const int64_t len_i = std::min(end, off[i + 1]) - pos;
                if (len_i > 0) {
                    memcpy(buf + (pos - c0), (const char *)host[i] + (pos - off[i]), (size_t)len_i);
                    pos += len_i;
                }
                ++i;
            }
        };
        // about 1 MiB per thread; small chunks on the calling thread
        const int use = (int)std::min<int64_t>(T, (len + (1 << 20) - 1) >> 20);
        if (use <= 1)
            fill(0, len);
        else
            pool_->run([&](int t) {
                if (t < use) fill(len * t / use, len * (t + 1) / use);
            });
        LV_HIP(hipMemcpyAsync((char *)dev + c0, buf, (size_t)len, hipMemcpyHostToDevice, s));
        LV_HIP(hipEventRecord(ring_ev_[b], s));
    }
    return LV_OK;
}

int EdgeUploader::threads()
{
    // a child forked after the pool started has the pool's state but not its
    // threads: it starts its own (the parent's copy is left alone)
    if (pool_ && pool_pid_ != getpid()) pool_ = nullptr;
    if (!pool_) {
        pool_ = new HostPool(usable_cpus());
        pool_pid_ = getpid();
    }
    return pool_->size();
}

// Page-locked staging, kept for the next upload (the copies from it have
// finished before lv_load_edges returns).
int EdgeUploader::stage(void **p, size_t *cap, size_t bytes)
{
    if (*cap >= bytes) return LV_OK;
    if (*p) (void)hipHostFree(*p);
    *p = nullptr;
    *cap = 0;
    if (hipHostMalloc(p, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return LV_ERR_NOMEM;
    }
    *cap = bytes;
    return LV_OK;
}


// True for page-locked host memory (hipHostMalloc), which the copy engines
// read directly; pageable memory (or a pointer the runtime does not know)
// goes through the staging ring instead of the runtime's pageable path,
// whose rate varies several-fold between hosts.
static bool host_pinned(const void *p)
{
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return at.type == hipMemoryTypeHost;
}

int EdgeUploader::download(hipStream_t s, void *host, const void *dev, size_t bytes)
{
    if (bytes < (size_t)kRingMin || host_pinned(host)) {
        LV_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, s));
        LV_HIP(hipStreamSynchronize(s));
        return LV_OK;
    }
    LV_TRY(ring());
    const int T = threads();
    const int64_t total = (int64_t)bytes;
    const int64_t nch = (total + kRingBytes - 1) / kRingBytes;
    // chunk k goes to buffer k & 1; chunk k + 1 is on the link while chunk k is
    // copied out (its buffer's last chunk, k - 1, was copied out already)
    auto issue = [&](int64_t k) -> int {
        const int b = (int)(k & 1);
        const int64_t c0 = k * kRingBytes, len = std::min(kRingBytes, total - c0);
        LV_HIP(hipEventSynchronize(ring_ev_[b]));  // an earlier upload's copy out of it
        LV_HIP(hipMemcpyAsync(h_ring_[b], (const char *)dev + c0, (size_t)len,
                              hipMemcpyDeviceToHost, s));
        LV_HIP(hipEventRecord(ring_ev_[b], s));
        return LV_OK;
    };
    LV_TRY(issue(0));
    for (int64_t k = 0; k < nch; ++k) {
        if (k + 1 < nch) LV_TRY(issue(k + 1));
        const int b = (int)(k & 1);
        const int64_t c0 = k * kRingBytes, len = std::min(kRingBytes, total - c0);
        LV_HIP(hipEventSynchronize(ring_ev_[b]));
        const char *buf = (const char *)h_ring_[b];
        char *out = (char *)host + c0;
        const int use = (int)std::min<int64_t>(T, (len + (1 << 20) - 1) >> 20);
        if (use <= 1)
            memcpy(out, buf, (size_t)len);
        else
            pool_->run([&](int t) {
                if (t < use) {
                    const int64_t a = len * t / use, z = len * (t + 1) / use;
                    memcpy(out + a, buf + a, (size_t)(z - a));
                }
            });
    }
    return LV_OK;
}

// One raw host array to the device: a direct copy from page-locked memory
// (or when the staging ring may not be allocated), else through the ring.
int EdgeUploader::raw(hipStream_t s, void *dev, const void *host, size_t bytes)
{
    if (!stage_ok_ || host_pinned(host)) {
        LV_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, s));
        return LV_OK;
    }
    const int64_t off[2] = {0, (int64_t)bytes};
    const void *h[1] = {host};
    return upload_concat(s, dev, 1, h, off);
}

int EdgeUploader::upload(hipStream_t s, int32_t n, int64_t e, const int32_t *src,
                         const int32_t *dst, const void *w, int wtype, bool encode, bool staging,
                         const UploadBufs &d, UploadPlan *plan)
{
    *plan = UploadPlan{};
    stage_ok_ = staging;
    if (e <= 0) return LV_OK;
    const size_t eb = 4 * (size_t)e;
    LV_TRY(raw(s, d.dst, dst, eb));
    plan->bytes = (int64_t)eb;
    if (!encode) {
        LV_TRY(raw(s, d.w, w, eb));
        LV_TRY(raw(s, d.src, src, eb));
        plan->bytes += 2 * (int64_t)eb;
        return LV_OK;
    }
    const int T = threads();
    auto part = [&](int t, int64_t &a, int64_t &b) {
        a = e * t / T;
        b = e * (t + 1) / T;
    };
    // weights first: their copy then queues behind dst while src is scanned
    bool packed = false;
    if (wtype == LV_WEIGHT_INT32) {
        LV_TRY(stage(&h_wpk_, &wpk_cap_, (size_t)e));
        esc_parts_.resize(T);
        std::atomic<bool> over{false};
        const uint32_t *x = (const uint32_t *)w;
        uint8_t *o = (uint8_t *)h_wpk_;
        const int64_t cap = d.esc_cap;
        const int wb = enc::pack_width_of(x, e);
        pool_->run([&](int t) {
            std::vector<int64_t> &esc = esc_parts_[t];
            esc.clear();
            // parts start on 64-edge boundaries, so no byte is shared
            const int64_t a = (e * t / T) & ~(int64_t)63, b = t + 1 == T ? e : (e * (t + 1) / T) & ~(int64_t)63;
            if (a < b && !enc::pack_part(x, a, b, o, esc, cap, wb)) over = true;
        });
        std::vector<int64_t> pre((size_t)T + 1, 0);
        for (int t = 0; t < T; ++t) pre[t + 1] = pre[t] + (int64_t)esc_parts_[t].size();
        const int64_t ne = pre[T];
        if (!over && ne <= cap) {
            LV_TRY(stage(&h_esc_, &esc_bytes_, 12 * (size_t)std::max<int64_t>(ne, 1)));
            int64_t *hi = (int64_t *)h_esc_;
            uint32_t *hw = (uint32_t *)(hi + ne);
            // the escape list in edge order, each thread its own part's
            pool_->run([&](int t) {
                int64_t j = pre[t];
                for (int64_t i : esc_parts_[t]) {
                    hi[j] = i;
                    hw[j] = x[i];
                    ++j;
                }
            });
            const size_t pb = (size_t)((e * wb + 7) / 8);
            LV_HIP(hipMemcpyAsync(d.wpk, h_wpk_, pb, hipMemcpyHostToDevice, s));
            if (ne > 0) {
                LV_HIP(hipMemcpyAsync(d.esc_i, hi, 8 * (size_t)ne, hipMemcpyHostToDevice, s));
                LV_HIP(hipMemcpyAsync(d.esc_w, hw, 4 * (size_t)ne, hipMemcpyHostToDevice, s));
            }
            plan->wbits = wb;
            plan->nesc = ne;
            plan->bytes += (int64_t)pb + 12 * ne;
            packed = true;
        }
    }
    if (!packed) {
        LV_TRY(raw(s, d.w, w, eb));
        plan->bytes += (int64_t)eb;
    }
    // sources: row offsets when sorted and in range
    const bool o64 = e > INT32_MAX;
    const size_t ob = (size_t)(n + 1) * (o64 ? 8 : 4);
    LV_TRY(stage(&h_off_, &off_cap_, ob));
    std::atomic<bool> bad{false};
    pool_->run([&](int t) {
        int64_t a, b;
        part(t, a, b);
        if (a >= b) return;
        const bool ok = o64 ? enc::offsets_part(src, a, b, e, n, (int64_t *)h_off_)
                            : enc::offsets_part(src, a, b, e, n, (int32_t *)h_off_);
        if (!ok) bad = true;
    });
    if (!bad) {
        LV_HIP(hipMemcpyAsync(d.off, h_off_, ob, hipMemcpyHostToDevice, s));
        plan->src_off = true;
        plan->off64 = o64;
        plan->bytes += (int64_t)ob;
    } else {
        LV_TRY(raw(s, d.src, src, eb));
        plan->bytes += (int64_t)eb;
    }
    return LV_OK;
}

// ---------------------------------------------------------------------------
// device: decode and check

constexpr int kDecSpan = 4096;  // vertices of a decode tile scattered in LDS
constexpr int kDecPer = kDecTile / 256;  // tile positions per thread in the scan

// Tile position p in LDS, one pad word per kDecPer so that the threads' runs of
// kDecPer consecutive positions fall in different banks.
__device__ __forceinline__ int dec_pad(int p) { return p + p / kDecPer; }

// Source of the first edge of every decode tile, and (entry ntiles) of edge
// e - 1: the largest v with off[v] <= t (off[0] = 0 <= t < e = off[n]).
template <typename O>
__global__ void k_dec_tiles(const O *off, int32_t n, int64_t e, int64_t ntiles, int32_t *tile_v)
{
    GRID_STRIDE(k, ntiles + 1)
    {
        const int64_t t = k < ntiles ? k * (int64_t)kDecTile : e - 1;
        int32_t lo = 0, hi = n - 1;
        while (lo < hi) {
            const int32_t mid = lo + ((hi - lo + 1) >> 1);
            if ((int64_t)off[mid] <= t)
                lo = mid;
            else
                hi = mid - 1;
        }
        tile_v[k] = lo;
    }
}

// Largest j in [0, hi] with o[j] <= i (o[0] <= i).
template <typename A>
__device__ __forceinline__ int32_t last_le(const A &o, int32_t hi, int64_t i)
{
    int32_t lo = 0;
    while (lo < hi) {
        const int32_t mid = lo + ((hi - lo + 1) >> 1);
        if ((int64_t)o[mid] <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// The weight checks and sums of one edge (x: its WB-bit value when packed,
// else its raw 32-bit pattern); an escape (all WB bits set) is left to
// k_load_escapes.
template <int WB>
__device__ __forceinline__ void check_weight(uint32_t x, int wtype, int &bad, int &smax,
                                             uint32_t &mb, ull &tot, ull &msum, ull &umax)
{
    if (WB) {
        if (x != (1u << WB) - 1u) {
            umax = max(umax, (ull)x);
            tot = sat_add(tot, (ull)(x == 0 ? 1 : x));
            msum = sat_add(msum, (ull)x);
        }
    } else if (wtype == LV_WEIGHT_INT32) {
        if ((int32_t)x < 0) bad |= 1;
        const int64_t u = weight_units(x, wtype, 0, bad);
        umax = max(umax, (ull)u);
        tot = sat_add(tot, (ull)(u == 0 ? 1 : u));
        msum = sat_add(msum, (ull)u);
    } else {
        const int sx = f32_unit_exponent(x);
        if (sx < 0) {
            bad |= 1;
        } else {
            smax = max(smax, sx);
            mb = max(mb, x & 0x7fffffffu);
        }
    }
}

// One pass over the edges in tiles of kDecTile: decodes the sources (SOFF: each
// row start in the tile marks its position with the row - the largest row
// when empty rows share a start - and a running maximum over the tile's
// positions, seeded with the row of the tile's first edge, gives every edge
// its row; a tile spanning more than kDecSpan rows binary-searches the
// offsets instead), checks the endpoints, and checks the weights as k_weight_exponent does (int32
// weights also summed as k_weight_units does with exponent 0: they are their
// own units, and packed weights (WB bits) are widened into wu32).  Escaped
// weights (all WB bits set) are left to k_load_escapes.
template <bool SOFF, int WB, typename O>
__global__ __launch_bounds__(256) void k_load_checks(LoadedEdges L, const O *off, int wtype,
                                                    int64_t e, int32_t n, uint32_t *wu32,
                                                    DevFlags *f)
{
    __shared__ int32_t s_src[SOFF ? kDecTile + kDecTile / kDecPer : 1];
    __shared__ int32_t s_wt[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int bad = 0, smax = 0;
    uint32_t mb = 0;
    ull tot = 0, msum = 0, umax = 0;
    const int64_t ntiles = (e + kDecTile - 1) / kDecTile;
    for (int64_t k = blockIdx.x; k < ntiles; k += gridDim.x) {
        const int64_t t0 = k * (int64_t)kDecTile, t1 = min(e, t0 + kDecTile);
        const int T = (int)(t1 - t0);
        int32_t lo = 0, span = 0;
        if (SOFF) {
            // the tile's edges have sources in [lo, tile_v[k + 1]]
            lo = L.tile_v[k];
            span = L.tile_v[k + 1] - lo + 2;  // off[lo .. tile_v[k + 1] + 1]
            if (span <= kDecSpan) {
                for (int p = tid; p < T; p += 256) s_src[dec_pad(p)] = -1;
                __syncthreads();
                for (int j = tid; j < span; j += 256) {
                    const int64_t p = (int64_t)off[lo + j] - t0;
                    if (p >= 0 && p < T) atomicMax(&s_src[dec_pad((int)p)], lo + j);
                }
                __syncthreads();
                // running maximum: kDecPer consecutive positions per thread, then
                // the threads' maxima scanned over the block
                int32_t run = tid == 0 ? lo : -1;
                const int p0 = tid * kDecPer;
#pragma unroll
                for (int q = 0; q < kDecPer; ++q)
                    if (p0 + q < T) run = max(run, s_src[dec_pad(p0 + q)]);
                int32_t inc = run;
#pragma unroll
                for (int off2 = 1; off2 < 64; off2 <<= 1) {
                    const int32_t y = __shfl_up(inc, off2, 64);
                    if (lane >= off2) inc = max(inc, y);
                }
                if (lane == 63) s_wt[wave] = inc;
                __syncthreads();
                int32_t carry = __shfl_up(inc, 1, 64);
                if (lane == 0) carry = -1;
                for (int w2 = 0; w2 < wave; ++w2) carry = max(carry, s_wt[w2]);
                if (tid == 0) carry = lo;
#pragma unroll
                for (int q = 0; q < kDecPer; ++q)
                    if (p0 + q < T) {
                        carry = max(carry, s_src[dec_pad(p0 + q)]);
                        s_src[dec_pad(p0 + q)] = carry;
                    }
            }
            __syncthreads();
        }
        if (T == kDecTile) {
            // a full tile: four consecutive edges per thread and round, every load
            // of the tile issued before the first store
            constexpr int R = kDecTile / 1024;
            int4 b4[R], a4[R];
            uint4 x4[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int64_t i = t0 + r * 1024 + tid * 4;
                b4[r] = *(const int4 *)(L.dst + i);
                if (!SOFF) a4[r] = *(const int4 *)(L.src + i);
                if (WB) {
                    // the four edges' 4 * WB bits, whole bytes since i is a multiple of 4
                    constexpr uint32_t M = (1u << WB) - 1u;
                    const uint32_t q = WB == 8   ? *(const uint32_t *)(L.wpk + i)
                                       : WB == 4 ? (uint32_t)*(const uint16_t *)(L.wpk + i / 2)
                                                 : (uint32_t)L.wpk[i / 4];
                    x4[r] = make_uint4(q & M, (q >> WB) & M, (q >> (2 * WB)) & M,
                                       (q >> (3 * WB)) & M);
                } else {
                    x4[r] = *(const uint4 *)(L.w + i);
                }
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int p = r * 1024 + tid * 4;
                const int64_t i = t0 + p;
                if (SOFF) {
                    if (span <= kDecSpan) {
                        a4[r] = make_int4(s_src[dec_pad(p)], s_src[dec_pad(p + 1)],
                                          s_src[dec_pad(p + 2)], s_src[dec_pad(p + 3)]);
                    } else {
                        a4[r].x = lo + last_le(off + lo, span - 2, i);
                        a4[r].y = lo + last_le(off + lo, span - 2, i + 1);
                        a4[r].z = lo + last_le(off + lo, span - 2, i + 2);
                        a4[r].w = lo + last_le(off + lo, span - 2, i + 3);
                    }
                    *(int4 *)(L.src_out + i) = a4[r];
                }
                const int32_t av[4] = {a4[r].x, a4[r].y, a4[r].z, a4[r].w};
                const int32_t bv[4] = {b4[r].x, b4[r].y, b4[r].z, b4[r].w};
                const uint32_t xv[4] = {x4[r].x, x4[r].y, x4[r].z, x4[r].w};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (av[q] < 0 || av[q] >= n || bv[q] < 0 || bv[q] >= n) bad |= 1;
                    check_weight<WB>(xv[q], wtype, bad, smax, mb, tot, msum, umax);
                }
                // an escape's all-ones value is written too; k_load_escapes overwrites it
                if (WB) *(uint4 *)(wu32 + i) = x4[r];
            }
        } else {
            for (int64_t i = t0 + tid; i < t1; i += 256) {
                int32_t a;
                if (SOFF) {
                    a = span <= kDecSpan ? s_src[dec_pad((int)(i - t0))]
                                         : lo + last_le(off + lo, span - 2, i);
                    L.src_out[i] = a;
                } else {
                    a = L.src[i];
                }
                const int32_t b = L.dst[i];
                if (a < 0 || a >= n || b < 0 || b >= n) bad |= 1;
                const uint32_t x =
                    WB ? ((uint32_t)L.wpk[(i * WB) >> 3] >> ((i * WB) & 7)) & ((1u << WB) - 1u)
                       : L.w[i];
                check_weight<WB>(x, wtype, bad, smax, mb, tot, msum, umax);
                if (WB) wu32[i] = x;
            }
        }
        if (SOFF) __syncthreads();  // s_src is refilled for the next tile
    }
    smax = wave_max_i(smax);
    mb = (uint32_t)wave_max_i((int)mb);
    if ((threadIdx.x & 63) == 0) {
        if (smax) atomicMax(&f->scale, smax);
        if (mb) atomicMax(&f->max_bits, mb);
    }
    block_units_flush(tot, msum, umax, bad, f);
}

// The escaped int weights (255 and up, negative ones included) of a byte upload.
__global__ __launch_bounds__(256) void k_load_escapes(const int64_t *ei, const uint32_t *ew,
                                                     int64_t nesc, uint32_t *wu32, DevFlags *f)
{
    int bad = 0;
    ull tot = 0, msum = 0, umax = 0;
    GRID_STRIDE(j, nesc)
    {
        const uint32_t x = ew[j];
        if ((int32_t)x < 0) bad |= 1;
        const int64_t u = weight_units(x, LV_WEIGHT_INT32, 0, bad);
        wu32[ei[j]] = x;
        umax = max(umax, (ull)u);
        tot = sat_add(tot, (ull)(u == 0 ? 1 : u));
        msum = sat_add(msum, (ull)u);
    }
    block_units_flush(tot, msum, umax, bad, f);
}

template <bool SOFF, typename O>
static void launch_checks_w(hipStream_t s, const LoadedEdges &L, const O *off, int wtype, int64_t e,
                            int32_t n, uint32_t *wu, DevFlags *f, int grid)
{
    switch (L.wpk ? L.wbits : 0) {
    case 2:
        hipLaunchKernelGGL((k_load_checks<SOFF, 2, O>), grid, 256, 0, s, L, off, wtype, e, n, wu, f);
        break;
    case 4:
        hipLaunchKernelGGL((k_load_checks<SOFF, 4, O>), grid, 256, 0, s, L, off, wtype, e, n, wu, f);
        break;
    case 8:
        hipLaunchKernelGGL((k_load_checks<SOFF, 8, O>), grid, 256, 0, s, L, off, wtype, e, n, wu, f);
        break;
    default:
        hipLaunchKernelGGL((k_load_checks<SOFF, 0, O>), grid, 256, 0, s, L, off, wtype, e, n, wu, f);
    }
}

void launch_load_checks(hipStream_t s, const LoadedEdges &L, int wtype, int64_t e, int32_t n,
                        uint32_t *wu, DevFlags *f)
{
    if (e <= 0) return;
    const int64_t ntiles = (e + kDecTile - 1) / kDecTile;
    const int grid = (int)std::min<int64_t>(ntiles, 8192);
    if (L.off32 || L.off64) {
        if (L.off32) {
            hipLaunchKernelGGL(k_dec_tiles<int32_t>, grid_for(ntiles + 1, 256), 256, 0, s, L.off32,
                               n, e, ntiles, L.tile_v);
            launch_checks_w<true>(s, L, L.off32, wtype, e, n, wu, f, grid);
        } else {
            hipLaunchKernelGGL(k_dec_tiles<int64_t>, grid_for(ntiles + 1, 256), 256, 0, s, L.off64,
                               n, e, ntiles, L.tile_v);
            launch_checks_w<true>(s, L, L.off64, wtype, e, n, wu, f, grid);
        }
    } else {
        launch_checks_w<false>(s, L, (const int32_t *)nullptr, wtype, e, n, wu, f, grid);
    }
    if (L.wpk && L.nesc > 0)
        hipLaunchKernelGGL(k_load_escapes, grid_for(L.nesc, 256, 1024), 256, 0, s, L.esc_i,
                           L.esc_w, L.nesc, wu, f);
}

}  // namespace lv
