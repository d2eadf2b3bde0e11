// graphio.cpp - graph files: the reference CLI's text loaders
// (LouvainCUDA/LouvainCPP.cpp:660-737) and a little-endian binary edge list
// for graphs too large for text (include/louvain_host_c.h).
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "louvain_host.hpp"
#include "louvain_host_c.h"

namespace {

// Whitespace-separated tokens of a file, with the one fact the reference's
// `stream >> token` loops depend on: whether a token runs to the end of the
// file (then the stream's eof bit is set right after reading it).
class Tokens {
public:
    explicit Tokens(const std::string &path)
    {
        f_ = std::fopen(path.c_str(), "rb");
        if (!f_) throw std::runtime_error("cannot open " + path);
    }
    ~Tokens()
    {
        if (f_) std::fclose(f_);
    }
    // Next token into tok; false at the end.  at_eof: the token ends the file.
    bool next(std::string &tok, bool &at_eof)
    {
        tok.clear();
        int ch;
        do {
            ch = get();
        } while (ch != EOF && is_space(ch));
        if (ch == EOF) return false;
        while (ch != EOF && !is_space(ch)) {
            tok.push_back((char)ch);
            ch = get();
        }
        at_eof = ch == EOF;
        return true;
    }

private:
    static bool is_space(int c)
    {
        return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f';
    }
    int get()
    {
        if (pos_ == len_) {
            len_ = std::fread(buf_, 1, sizeof(buf_), f_);
            pos_ = 0;
            if (len_ == 0) return EOF;
        }
        return (unsigned char)buf_[pos_++];
    }
    std::FILE *f_ = nullptr;
    char buf_[1 << 16];
    size_t pos_ = 0, len_ = 0;
};

// The count header: the reference allocates that many entries and never checks
// it, so more entries than the count would overrun its array; here that is an
// error.
long long read_count(Tokens &t, const std::string &path)
{
    std::string tok;
    bool at_eof = false;
    if (!t.next(tok, at_eof)) throw std::runtime_error("missing count in " + path);
    const long long c = std::atoll(tok.c_str());
    if (c < 0) throw std::runtime_error("negative count in " + path);
    return c;
}

struct FileGraph {
    std::vector<int32_t> ids, src, dst;
    std::vector<float> w;
};

// ---------------------------------------------------------------------------
// Binary edge list.  Little-endian, 32-byte header then the arrays:
//   0  "LVEB"            4  u32 version (1)     8  i32 n
//   12 u32 flags (bit 0: node ids present, bit 1: float32 weights, else int32)
//   16 i64 e              24 8 zero bytes
//   ids int32[n] (if flagged), src int32[e], dst int32[e], w 4-byte[e]
constexpr char kMagic[4] = {'L', 'V', 'E', 'B'};
constexpr uint32_t kVersion = 1;
constexpr size_t kHeader = 32;

bool little_endian()
{
    const uint32_t x = 1;
    unsigned char b;
    std::memcpy(&b, &x, 1);
    return b == 1;
}

void swap4(void *p, size_t count)
{
    unsigned char *b = (unsigned char *)p;
    for (size_t i = 0; i < count; ++i, b += 4) {
        std::swap(b[0], b[3]);
        std::swap(b[1], b[2]);
    }
}

template <typename T>
void put_le(unsigned char *p, T v)
{
    for (size_t i = 0; i < sizeof(T); ++i) p[i] = (unsigned char)((uint64_t)v >> (8 * i));
}

template <typename T>
T get_le(const unsigned char *p)
{
    uint64_t v = 0;
    for (size_t i = 0; i < sizeof(T); ++i) v |= (uint64_t)p[i] << (8 * i);
    return (T)v;
}

struct BinHeader {
    int32_t n = 0;
    int64_t e = 0;
    int wtype = LV_WEIGHT_INT32;
    bool has_ids = false;
};

// Reads and checks the header; the file size must match it exactly.
int read_header(std::FILE *f, BinHeader &h)
{
    unsigned char b[kHeader];
    if (std::fread(b, 1, kHeader, f) != kHeader) return LV_ERR_INVALID_INPUT;
    if (std::memcmp(b, kMagic, 4) != 0 || get_le<uint32_t>(b + 4) != kVersion)
        return LV_ERR_INVALID_INPUT;
    h.n = get_le<int32_t>(b + 8);
    const uint32_t flags = get_le<uint32_t>(b + 12);
    h.e = get_le<int64_t>(b + 16);
    if (h.n < 0 || h.e < 0 || (flags & ~3u) != 0 || get_le<uint64_t>(b + 24) != 0)
        return LV_ERR_INVALID_INPUT;
    h.has_ids = (flags & 1u) != 0;
    h.wtype = (flags & 2u) ? LV_WEIGHT_FLOAT32 : LV_WEIGHT_INT32;
    const long long want = (long long)kHeader + (h.has_ids ? 4ll * h.n : 0) + 12ll * h.e;
    if (std::fseek(f, 0, SEEK_END) != 0) return LV_ERR_INVALID_INPUT;
    const long long size = std::ftell(f);
    if (size != want || std::fseek(f, (long)kHeader, SEEK_SET) != 0) return LV_ERR_INVALID_INPUT;
    return LV_OK;
}

int read_array(std::FILE *f, void *dst, int64_t count)
{
    const size_t chunk = (size_t)1 << 26;  // 256 MiB of 4-byte items
    unsigned char *p = (unsigned char *)dst;
    for (int64_t done = 0; done < count;) {
        const size_t k = (size_t)std::min<int64_t>(count - done, (int64_t)chunk);
        if (std::fread(p + 4 * (size_t)done, 4, k, f) != k) return LV_ERR_INVALID_INPUT;
        if (!little_endian()) swap4(p + 4 * (size_t)done, k);
        done += (int64_t)k;
    }
    return LV_OK;
}

int write_array(std::FILE *f, const void *src, int64_t count)
{
    if (count <= 0) return LV_OK;
    if (little_endian()) {
        return std::fwrite(src, 4, (size_t)count, f) == (size_t)count ? LV_OK : LV_ERR_ARG;
    }
    std::vector<unsigned char> tmp((size_t)1 << 20);
    const unsigned char *p = (const unsigned char *)src;
    for (int64_t done = 0; done < count;) {
        const size_t k = (size_t)std::min<int64_t>(count - done, (int64_t)tmp.size() / 4);
        std::memcpy(tmp.data(), p + 4 * (size_t)done, 4 * k);
        swap4(tmp.data(), k);
        if (std::fwrite(tmp.data(), 4, k, f) != k) return LV_ERR_ARG;
        done += (int64_t)k;
    }
    return LV_OK;
}

}  // namespace

// loadNodes (LouvainCUDA/LouvainCPP.cpp:660-696): a count, then node ids.  The
// reference stores a token only when reading it did not reach the end of the
// file, so without trailing whitespace the last id is dropped; kept here.
std::vector<int> loadNodesFile(const std::string &path)
{
    Tokens t(path);
    const long long count = read_count(t, path);
    std::vector<int> ids;
    std::string tok;
    bool at_eof = false;
    while (t.next(tok, at_eof) && !at_eof) {
        if ((long long)ids.size() == count)
            throw std::runtime_error("more node ids than the count in " + path);
        ids.push_back(std::atoi(tok.c_str()));
    }
    return ids;
}

// loadEdges (LouvainCUDA/LouvainCPP.cpp:698-737): a count, then `source target
// weight` triples (atoi, atoi, atof).  A triple is kept when its source token
// does not end the file, which keeps every complete triple; a trailing
// incomplete triple, which the reference fills from a stale token, is an error
// here.
std::vector<Edge> loadEdgesFile(const std::string &path)
{
    Tokens t(path);
    const long long count = read_count(t, path);
    std::vector<Edge> edges;
    std::string a, b, c;
    bool at_eof = false;
    while (t.next(a, at_eof) && !at_eof) {
        bool eb = false, ec = false;
        if (!t.next(b, eb) || eb || !t.next(c, ec))
            throw std::runtime_error("incomplete edge triple in " + path);
        if ((long long)edges.size() == count)
            throw std::runtime_error("more edges than the count in " + path);
        Edge e;
        e.source = std::atoi(a.c_str());
        e.target = std::atoi(b.c_str());
        e.weight = (float)std::atof(c.c_str());
        edges.push_back(e);
        if (ec) break;
    }
    return edges;
}

struct lvh_file_graph : FileGraph {};

extern "C" {

int lvh_load_text(const char *nodes_path, const char *edges_path, lvh_file_graph **out)
{
    if (!nodes_path || !edges_path || !out) return LV_ERR_ARG;
    *out = nullptr;
    std::unique_ptr<lvh_file_graph> g(new (std::nothrow) lvh_file_graph());
    if (!g) return LV_ERR_NOMEM;
    try {
        std::vector<int> ids = loadNodesFile(nodes_path);
        std::vector<Edge> edges = loadEdgesFile(edges_path);
        g->ids.assign(ids.begin(), ids.end());
        g->src.resize(edges.size());
        g->dst.resize(edges.size());
        g->w.resize(edges.size());
        for (size_t i = 0; i < edges.size(); ++i) {
            g->src[i] = edges[i].source;
            g->dst[i] = edges[i].target;
            g->w[i] = edges[i].weight;
        }
    } catch (const std::bad_alloc &) {
        return LV_ERR_NOMEM;
    } catch (const std::runtime_error &ex) {
        return std::strncmp(ex.what(), "cannot open", 11) == 0 ? LV_ERR_ARG : LV_ERR_INVALID_INPUT;
    }
    *out = g.release();
    return LV_OK;
}

int32_t lvh_file_num_nodes(const lvh_file_graph *g) { return g ? (int32_t)g->ids.size() : -1; }

int64_t lvh_file_num_edges(const lvh_file_graph *g) { return g ? (int64_t)g->src.size() : -1; }

int lvh_file_copy(const lvh_file_graph *g, int32_t *ids, int32_t *src, int32_t *dst, float *w)
{
    if (!g) return LV_ERR_ARG;
    if (ids && !g->ids.empty()) std::memcpy(ids, g->ids.data(), 4 * g->ids.size());
    if (!g->src.empty()) {
        if (!src || !dst || !w) return LV_ERR_ARG;
        std::memcpy(src, g->src.data(), 4 * g->src.size());
        std::memcpy(dst, g->dst.data(), 4 * g->dst.size());
        std::memcpy(w, g->w.data(), 4 * g->w.size());
    }
    return LV_OK;
}

void lvh_file_free(lvh_file_graph *g) { delete g; }

int lvh_binary_header(const char *path, int32_t *n, int64_t *e, int *wtype, int *has_ids)
{
    if (!path) return LV_ERR_ARG;
    std::FILE *f = std::fopen(path, "rb");
    if (!f) return LV_ERR_ARG;
    BinHeader h;
    const int rc = read_header(f, h);
    std::fclose(f);
    if (rc != LV_OK) return rc;
    if (n) *n = h.n;
    if (e) *e = h.e;
    if (wtype) *wtype = h.wtype;
    if (has_ids) *has_ids = h.has_ids ? 1 : 0;
    return LV_OK;
}

int lvh_binary_read(const char *path, int32_t *ids, int32_t *src, int32_t *dst, void *w)
{
    if (!path) return LV_ERR_ARG;
    std::FILE *f = std::fopen(path, "rb");
    if (!f) return LV_ERR_ARG;
    BinHeader h;
    int rc = read_header(f, h);
    if (rc == LV_OK && h.e > 0 && (!src || !dst || !w)) rc = LV_ERR_ARG;
    if (rc == LV_OK && h.has_ids) {
        if (ids) {
            rc = read_array(f, ids, h.n);
        } else if (std::fseek(f, (long)(4ll * h.n), SEEK_CUR) != 0) {
            rc = LV_ERR_INVALID_INPUT;
        }
    }
    if (rc == LV_OK) rc = read_array(f, src, h.e);
    if (rc == LV_OK) rc = read_array(f, dst, h.e);
    if (rc == LV_OK) rc = read_array(f, w, h.e);
    std::fclose(f);
    return rc;
}

int lvh_binary_write(const char *path, int32_t n, const int32_t *ids, int64_t e,
                     const int32_t *src, const int32_t *dst, const void *w, int wtype)
{
    if (!path || n < 0 || e < 0 || (e > 0 && (!src || !dst || !w)) ||
        (wtype != LV_WEIGHT_INT32 && wtype != LV_WEIGHT_FLOAT32))
        return LV_ERR_ARG;
    std::FILE *f = std::fopen(path, "wb");
    if (!f) return LV_ERR_ARG;
    unsigned char b[kHeader] = {0};
    std::memcpy(b, kMagic, 4);
    put_le<uint32_t>(b + 4, kVersion);
    put_le<int32_t>(b + 8, n);
    put_le<uint32_t>(b + 12, (ids ? 1u : 0u) | (wtype == LV_WEIGHT_FLOAT32 ? 2u : 0u));
    put_le<int64_t>(b + 16, e);
    int rc = std::fwrite(b, 1, kHeader, f) == kHeader ? LV_OK : LV_ERR_ARG;
    if (rc == LV_OK && ids) rc = write_array(f, ids, n);
    if (rc == LV_OK) rc = write_array(f, src, e);
    if (rc == LV_OK) rc = write_array(f, dst, e);
    if (rc == LV_OK) rc = write_array(f, w, e);
    if (std::fclose(f) != 0 && rc == LV_OK) rc = LV_ERR_ARG;
    return rc;
}

}  // extern "C"
