// graphio.hip -- graph files on the host: graph::GraphStd's readers for
// MatrixMarket, DIMACS9 and the binary graph format, its COO -> CSR pass, and
// the binary / MatrixMarket writers (host code only; the result is handed to
// hn_graph_create_csr like HornetInit(graph.nV(), graph.nE(),
// graph.csr_out_offsets(), graph.csr_out_edges()) in the reference drivers).
//
//   format dispatch            xlib/src/Graph/GraphBase.cpp:130-206
//   MatrixMarket header/body   GraphBase.cpp:214-238, GraphStdRead.cpp:45-61
//   DIMACS9 header/body        GraphBase.cpp:252-262, GraphStdRead.cpp:87-105
//   direction resolution       xlib/src/Graph/GraphStd.cpp:85-101
//   COOtoCSR                   GraphStd.cpp:173-351
//   binary read / write        GraphStdRead.cpp:206-236, GraphStd.cpp:387-421
//   writeMarket                GraphStd.cpp:425-437
//   max_out_degree_id          xlib/include/Graph/GraphStd.i.hpp:296-299
//
// Text bodies are parsed in parallel: the mapped file is cut into chunks at
// line starts, each thread parses its chunk into its own list, and the lists
// are joined in file order.  Deviations from the reference, where its
// behaviour is undefined: ids outside [1, n], entry counts that differ from
// the header and a malformed binary file are errors; RM_SINGLETON drops the
// edges touching a removed vertex; an UNDIRECTED graph's in-CSR is its
// out-CSR; the binary reader's degree is offsets[i+1] - offsets[i] (the
// reference subtracts offsets[i-1]).  A file stored undirected read into a
// DIRECTED structure (random edge removal, GraphStd.cpp:96-101) is rejected.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"

struct hn_gstd {
    int32_t nv = 0;
    int64_t ne = 0;
    uint32_t structure = 0;
    std::vector<int64_t> out_off, in_off;   // in_* only for DIRECTED | ENABLE_INGOING
    std::vector<int32_t> out_adj, in_adj;
};

namespace hn {
namespace {

constexpr uint32_t S_DIRECTED = HN_GS_DIRECTED, S_UNDIRECTED = HN_GS_UNDIRECTED,
                   S_INGOING = HN_GS_ENABLE_INGOING;

struct Mapped {
    const char* p = nullptr;
    size_t n = 0;
    int fd = -1;
    ~Mapped() {
        if (p && n) munmap(const_cast<char*>(p), n);
        if (fd >= 0) close(fd);
    }
};

hn_status map_file(const char* path, Mapped& m) {
    m.fd = open(path, O_RDONLY);
    HN_ARG(m.fd >= 0, std::string("cannot open ") + path);
    struct stat st;
    HN_ARG(fstat(m.fd, &st) == 0 && S_ISREG(st.st_mode), std::string(path) + " is not a regular file");
    HN_ARG(st.st_size > 0, std::string(path) + " is empty");
    m.n = (size_t)st.st_size;
    void* p = mmap(nullptr, m.n, PROT_READ, MAP_PRIVATE, m.fd, 0);
    HN_ARG(p != MAP_FAILED, std::string("cannot map ") + path);
    m.p = static_cast<const char*>(p);
    return HN_OK;
}

inline bool is_blank(char c) { return c == ' ' || c == '\t' || c == '\r'; }
inline bool is_space(char c) { return is_blank(c) || c == '\n' || c == '\v' || c == '\f'; }

// cursor over [p, e)
struct Cur {
    const char* p;
    const char* e;
    bool end() const { return p >= e; }
    char peek() const { return p < e ? *p : '\0'; }
    void skip_line() {
        const void* nl = p < e ? memchr(p, '\n', (size_t)(e - p)) : nullptr;
        p = nl ? static_cast<const char*>(nl) + 1 : e;
    }
    void skip_space() {
        while (p < e && is_space(*p)) p++;
    }
    void skip_blank() {
        while (p < e && is_blank(*p)) p++;
    }
    std::string word() {
        skip_space();
        const char* b = p;
        while (p < e && !is_space(*p)) p++;
        return std::string(b, (size_t)(p - b));
    }
    // unsigned decimal, across newlines (operator>> semantics)
    bool uint(uint64_t& v) {
        skip_space();
        return digits(v);
    }
    // unsigned decimal within the current line
    bool uint_inline(uint64_t& v) {
        skip_blank();
        return digits(v);
    }
    bool digits(uint64_t& v) {
        if (p < e && *p == '+') p++;
        const char* b = p;
        uint64_t x = 0;
        while (p < e && *p >= '0' && *p <= '9') {
            if (x > (UINT64_MAX - 9) / 10) return false;
            x = x * 10 + (uint64_t)(*p - '0');
            p++;
        }
        v = x;
        return p > b;
    }
};

std::string extension(const char* path) {
    std::string s(path);
    size_t slash = s.find_last_of('/');
    size_t dot = s.find_last_of('.');
    if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return std::string();
    return s.substr(dot);
}

// One parsed entry line: (u, v) 0-based; a chunk's error stops its parse.
struct Chunk {
    std::vector<uint64_t> keys;   // (u << 32) | v
    std::string err;
};

inline uint64_t pack(uint32_t u, uint32_t v) { return ((uint64_t)u << 32) | v; }

// vector storage without value-initialisation on resize (the contents are
// written right after)
template <typename T>
struct NoInit : std::allocator<T> {
    template <typename U>
    struct rebind {
        using other = NoInit<U>;
    };
    NoInit() = default;
    template <typename U>
    NoInit(const NoInit<U>&) noexcept {}
    template <typename U>
    void construct(U* p) noexcept {}
    template <typename U, typename... A>
    void construct(U* p, A&&... a) {
        ::new ((void*)p) U(std::forward<A>(a)...);
    }
};
using Keys = std::vector<uint64_t, NoInit<uint64_t>>;

int workers(size_t items) {
    unsigned hw = std::thread::hardware_concurrency();
    return (int)std::min<size_t>(std::max(1u, std::min(hw, 16u)), std::max<size_t>(1, items >> 18));
}

template <typename F>
void parallel(int T, F f) {
    std::vector<std::thread> th;
    for (int t = 1; t < T; t++) th.emplace_back(f, t);
    f(0);
    for (auto& x : th) x.join();
}

// Parse the entry lines of [b, e) in parallel.  `entry(cur, chunk)` parses one
// line (the cursor at its first non-blank character) and returns false to
// stop; lines that are empty are skipped.
template <typename Entry>
void parse_lines(const char* b, const char* e, std::vector<Chunk>& chunks, Entry entry) {
    const size_t len = (size_t)(e - b);
    unsigned hw = std::thread::hardware_concurrency();
    int T = (int)std::min<size_t>(std::max(1u, std::min(hw, 16u)), std::max<size_t>(1, len >> 20));
    std::vector<const char*> cut((size_t)T + 1);
    cut[0] = b;
    cut[(size_t)T] = e;
    for (int t = 1; t < T; t++) {
        const char* c = b + len * (size_t)t / (size_t)T;
        if (c < cut[(size_t)t - 1]) c = cut[(size_t)t - 1];
        const void* nl = c < e ? memchr(c, '\n', (size_t)(e - c)) : nullptr;
        cut[(size_t)t] = nl ? static_cast<const char*>(nl) + 1 : e;
    }
    chunks.assign((size_t)T, Chunk());
    auto work = [&](int t) {
        Cur c{cut[(size_t)t], cut[(size_t)t + 1]};
        Chunk& ch = chunks[(size_t)t];
        ch.keys.reserve((size_t)(c.e - c.p) / 8 + 16);
        while (!c.end()) {
            c.skip_blank();
            if (c.peek() == '\n') {
                c.p++;
                continue;
            }
            if (c.end()) break;
            if (!entry(c, ch)) return;
            c.skip_line();
        }
    };
    parallel(T, work);
}

hn_status join_chunks(std::vector<Chunk>& chunks, size_t want, const char* what, Keys& out) {
    size_t total = 0;
    std::vector<size_t> at(chunks.size());
    for (size_t t = 0; t < chunks.size(); t++) {
        HN_ARG(chunks[t].err.empty(), std::string(what) + ": " + chunks[t].err);
        at[t] = total;
        total += chunks[t].keys.size();
    }
    HN_ARG(total == want, std::string(what) + ": the header announces " + std::to_string(want) +
                              " entries, the file holds " + std::to_string(total));
    out.resize(total);
    parallel((int)chunks.size(), [&](int t) {
        std::copy(chunks[(size_t)t].keys.begin(), chunks[(size_t)t].keys.end(), out.begin() + (ptrdiff_t)at[(size_t)t]);
        std::vector<uint64_t>().swap(chunks[(size_t)t].keys);
    });
    return HN_OK;
}

// parse "u v" (1-based, in [1, n]) into chunk.keys
inline bool pair_entry(Cur& c, Chunk& ch, uint64_t n) {
    uint64_t u, v;
    if (!c.uint_inline(u) || !c.uint_inline(v)) {
        ch.err = "malformed entry line";
        return false;
    }
    if (u < 1 || u > n || v < 1 || v > n) {
        ch.err = "vertex id " + std::to_string(u < 1 || u > n ? u : v) + " outside [1, " + std::to_string(n) + "]";
        return false;
    }
    ch.keys.push_back(pack((uint32_t)(u - 1), (uint32_t)(v - 1)));
    return true;
}

struct FileCoo {
    uint64_t n = 0;
    uint32_t direction = 0;       // S_DIRECTED or S_UNDIRECTED
    bool stored_undirected = false;
    Keys keys;
};

// GraphBase::getMarketHeader + GraphStd::readMarket
hn_status read_market(const Mapped& m, FileCoo& f) {
    Cur c{m.p, m.p + m.n};
    const char* l0 = c.p;
    c.skip_line();
    std::string header(l0, (size_t)(c.p - l0));
    HN_ARG(header.find("array") == std::string::npos, "MatrixMarket: the matrix is a vector (array format)");
    f.stored_undirected = header.find("symmetric") != std::string::npos;
    f.direction = f.stored_undirected ? S_UNDIRECTED : S_DIRECTED;
    while (c.peek() == '%') c.skip_line();
    uint64_t rows, cols, lines;
    HN_ARG(c.uint(rows) && c.uint(cols) && c.uint(lines), "MatrixMarket: malformed size line");
    c.skip_line();
    f.n = std::max(rows, cols);
    HN_ARG(f.n <= (uint64_t)INT32_MAX, "MatrixMarket: more vertices than int32 ids hold");
    std::vector<Chunk> chunks;
    const uint64_t n = f.n;
    parse_lines(c.p, c.e, chunks, [n](Cur& cc, Chunk& ch) {
        if (cc.peek() == '%') return true;   // comment line
        return pair_entry(cc, ch, n);
    });
    // the reference reads exactly `lines` entries: what follows them is ignored
    // (a chunk stops at its first bad line, so entries before it are complete)
    size_t total = 0;
    for (auto& ch : chunks) {
        const size_t room = (size_t)lines - std::min<size_t>(total, (size_t)lines);
        if (ch.keys.size() >= room) {
            ch.keys.resize(room);
            ch.err.clear();
        }
        total += ch.keys.size();
    }
    return join_chunks(chunks, (size_t)lines, "MatrixMarket", f.keys);
}

// GraphBase::getDimacs9Header + GraphStd::readDimacs9
hn_status read_dimacs9(const Mapped& m, FileCoo& f) {
    Cur c{m.p, m.p + m.n};
    while (c.peek() == 'c') c.skip_line();
    std::string p = c.word(), sp = c.word();
    HN_ARG(p == "p", "DIMACS9: missing 'p' line");
    uint64_t nv, na;
    HN_ARG(c.uint(nv) && c.uint(na), "DIMACS9: malformed 'p' line");
    c.skip_line();
    HN_ARG(nv <= (uint64_t)INT32_MAX, "DIMACS9: more vertices than int32 ids hold");
    f.n = nv;
    f.direction = S_UNDIRECTED;
    f.stored_undirected = false;
    std::vector<Chunk> chunks;
    parse_lines(c.p, c.e, chunks, [nv](Cur& cc, Chunk& ch) {
        if (cc.peek() != 'a') return true;
        cc.p++;
        return pair_entry(cc, ch, nv);
    });
    return join_chunks(chunks, (size_t)na, "DIMACS9", f.keys);
}

// LSD radix sort of 64-bit keys, 16-bit digits; digits equal across all keys
// are skipped
void radix_sort(Keys& a) {
    const size_t n = a.size();
    if (n < 2) return;
    Keys b(n);
    std::vector<size_t> cnt(1 << 16);
    for (int shift = 0; shift < 64; shift += 16) {
        std::fill(cnt.begin(), cnt.end(), 0);
        for (size_t i = 0; i < n; i++) cnt[(a[i] >> shift) & 0xFFFF]++;
        if (cnt[(a[0] >> shift) & 0xFFFF] == n) continue;
        size_t s = 0;
        for (auto& x : cnt) {
            size_t t = x;
            x = s;
            s += t;
        }
        for (size_t i = 0; i < n; i++) b[cnt[(a[i] >> shift) & 0xFFFF]++] = a[i];
        a.swap(b);
    }
}

// Row counts of keys: workers take contiguous slices of the keys and add into
// one counter array with relaxed atomic increments.
template <typename Row>
std::vector<int64_t> row_counts(int32_t nv, const Keys& keys, Row row) {
    std::vector<int64_t> cnt((size_t)nv, 0);
    const int T = workers(keys.size());
    const size_t n = keys.size();
    parallel(T, [&](int t) {
        for (size_t i = n * (size_t)t / (size_t)T, e = n * (size_t)(t + 1) / (size_t)T; i < e; i++)
            __atomic_fetch_add(&cnt[row(keys[i])], (int64_t)1, __ATOMIC_RELAXED);
    });
    return cnt;
}

// rows r(key), entries e(key), each row in key order; workers own row ranges
template <typename Row, typename Ent>
void stable_csr(int32_t nv, const Keys& keys, Row row, Ent ent, std::vector<int64_t>& off,
                std::vector<int32_t>& adj) {
    std::vector<int64_t> cnt = row_counts(nv, keys, row);
    off.resize((size_t)nv + 1);
    off[0] = 0;
    for (int32_t v = 0; v < nv; v++) off[(size_t)v + 1] = off[(size_t)v] + cnt[(size_t)v];
    std::vector<int64_t>& pos = cnt;
    std::copy(off.begin(), off.end() - 1, pos.begin());
    adj.resize(keys.size());
    const int T = std::min(workers(keys.size()), std::max(1, nv));
    parallel(T, [&](int t) {
        const uint32_t lo = (uint32_t)((int64_t)nv * t / T), hi = (uint32_t)((int64_t)nv * (t + 1) / T);
        for (uint64_t k : keys) {
            const uint32_t r = row(k);
            if (r >= lo && r < hi) adj[(size_t)pos[r]++] = (int32_t)ent(k);
        }
    });
}

inline uint32_t ksrc(uint64_t k) { return (uint32_t)(k >> 32); }
inline uint32_t kdst(uint64_t k) { return (uint32_t)k; }

// GraphStd::allocateAux direction rules + COOtoCSR
hn_status coo_to_csr(FileCoo& f, uint32_t structure, uint32_t parse, uint64_t seed, hn_gstd& g) {
    HN_ARG(!((structure & S_DIRECTED) && (structure & S_UNDIRECTED)), "structure is both DIRECTED and UNDIRECTED");
    if (!(structure & (S_DIRECTED | S_UNDIRECTED))) structure |= f.direction;
    HN_ARG(!(f.direction == S_UNDIRECTED && (structure & S_DIRECTED)),
           "a file stored undirected cannot be read into a DIRECTED structure (the reference removes random "
           "edges there)");
    const bool d2u = f.direction == S_DIRECTED && (structure & S_UNDIRECTED);
    const int32_t n0 = (int32_t)f.n;
    Keys& keys = f.keys;
    if (d2u || f.stored_undirected) {
        const size_t half = keys.size();
        size_t loops = 0;
        for (size_t i = 0; i < half; i++) loops += ksrc(keys[i]) == kdst(keys[i]);
        keys.reserve(2 * half - loops);
        for (size_t i = 0; i < half; i++) {
            const uint64_t k = keys[i];
            if (ksrc(k) != kdst(k)) keys.push_back(pack(kdst(k), ksrc(k)));
        }
    }
    if (d2u) {
        radix_sort(keys);
        keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    }
    if (parse & HN_PARSE_RANDOMIZE) {
        // Fisher-Yates over SplitMix64 stream `seed` (the reference uses a clock seed)
        std::vector<uint32_t> perm((size_t)n0);
        for (int32_t v = 0; v < n0; v++) perm[(size_t)v] = (uint32_t)v;
        const uint64_t key = stream_key(seed);
        for (int64_t i = (int64_t)n0 - 1; i > 0; i--) {
            const uint32_t j = uniform_below(draw(key, (uint64_t)i), (uint32_t)(i + 1));
            std::swap(perm[(size_t)i], perm[j]);
        }
        for (auto& k : keys) k = pack(perm[ksrc(k)], perm[kdst(k)]);
    }
    const bool directed = structure & S_DIRECTED;
    const bool ingoing = directed && (structure & S_INGOING);
    const bool reverse = structure & (S_INGOING | S_UNDIRECTED);
    const bool rm_singleton = (parse & HN_PARSE_RM_SINGLETON) && reverse;
    std::vector<int64_t> outd, ind;
    auto degrees = [&]() {
        outd = row_counts(n0, keys, ksrc);
        if (ingoing) ind = row_counts(n0, keys, kdst);
    };
    if (parse & HN_PARSE_DIRECTED_BY_DEGREE) {
        degrees();
        size_t w = 0;
        for (uint64_t k : keys) {
            const uint32_t u = ksrc(k), v = kdst(k);
            if (outd[u] < outd[v] || (outd[u] == outd[v] && u < v)) keys[w++] = k;
        }
        keys.resize(w);
    }
    if ((parse & HN_PARSE_SORT) && (!d2u || (parse & HN_PARSE_RANDOMIZE))) radix_sort(keys);
    int32_t n = n0;
    if (rm_singleton) {
        degrees();
        const std::vector<int64_t>& in = ingoing ? ind : outd;
        std::vector<int32_t> label((size_t)n0, -1);
        int32_t k = 0;
        for (int32_t v = 0; v < n0; v++)
            if (outd[(size_t)v] > 0 && in[(size_t)v] > 0) label[(size_t)v] = k++;
        size_t w = 0;
        for (uint64_t key : keys) {
            const int32_t a = label[ksrc(key)], b = label[kdst(key)];
            if (a >= 0 && b >= 0) keys[w++] = pack((uint32_t)a, (uint32_t)b);
        }
        keys.resize(w);
        n = k;
    }
    g.nv = n;
    g.ne = (int64_t)keys.size();
    g.structure = structure;
    stable_csr(n, keys, ksrc, kdst, g.out_off, g.out_adj);
    if (ingoing) stable_csr(n, keys, kdst, ksrc, g.in_off, g.in_adj);
    return HN_OK;
}

// class identifiers type_name<vid_t>() + type_name<eoff_t>() of GraphStd<int, int>
// and GraphStd<int, int64_t> as g++ demangles them
constexpr const char* ID_INT32 = "intint";
constexpr const char* ID_INT64 = "intlong";

template <typename T>
bool take(const char*& p, const char* e, T* out, size_t count) {
    const size_t bytes = sizeof(T) * count;
    if ((size_t)(e - p) < bytes) return false;
    memcpy(out, p, bytes);
    p += bytes;
    return true;
}

template <typename Off>
hn_status read_binary_arrays(const char* p, const char* e, hn_gstd& g) {
    int32_t nv;
    Off ne;
    uint64_t blob[2];   // StructureProp object: vtable pointer, then its uint64 state
    HN_ARG(take(p, e, &nv, 1) && take(p, e, &ne, 1) && take(p, e, blob, 2), "binary graph: truncated header");
    HN_ARG(nv >= 0 && ne >= 0, "binary graph: negative size");
    const uint64_t state = blob[1];
    const bool twice = (state & S_DIRECTED) && (state & S_INGOING);
    const size_t want = (twice ? 2 : 1) * ((size_t)nv + 1) * sizeof(Off) + (twice ? 2 : 1) * (size_t)ne * 4;
    HN_ARG((size_t)(e - p) == want, "binary graph: file size does not match its header");
    std::vector<Off> off((size_t)nv + 1), ioff;
    take(p, e, off.data(), off.size());
    if (twice) {
        ioff.resize((size_t)nv + 1);
        take(p, e, ioff.data(), ioff.size());
    }
    g.nv = nv;
    g.ne = (int64_t)ne;
    g.structure = (uint32_t)state;
    g.out_adj.resize((size_t)ne);
    take(p, e, g.out_adj.data(), g.out_adj.size());
    if (twice) {
        g.in_adj.resize((size_t)ne);
        take(p, e, g.in_adj.data(), g.in_adj.size());
    }
    auto check = [&](const std::vector<Off>& o, const std::vector<int32_t>& a, std::vector<int64_t>& dst) -> hn_status {
        HN_ARG(o[0] == 0 && (int64_t)o[(size_t)nv] == (int64_t)ne, "binary graph: offsets do not span the edges");
        for (int32_t v = 0; v < nv; v++) HN_ARG(o[(size_t)v] <= o[(size_t)v + 1], "binary graph: offsets decrease");
        for (int32_t x : a) HN_ARG(x >= 0 && x < nv, "binary graph: edge endpoint outside [0, nV)");
        dst.assign(o.begin(), o.end());
        return HN_OK;
    };
    HN_TRY(check(off, g.out_adj, g.out_off));
    if (twice) HN_TRY(check(ioff, g.in_adj, g.in_off));
    return HN_OK;
}

hn_status read_binary(const Mapped& m, hn_gstd& g) {
    const char* p = m.p;
    const char* e = m.p + m.n;
    const size_t l64 = strlen(ID_INT64), l32 = strlen(ID_INT32);
    if (m.n >= l64 && memcmp(p, ID_INT64, l64) == 0) return read_binary_arrays<int64_t>(p + l64, e, g);
    if (m.n >= l32 && memcmp(p, ID_INT32, l32) == 0) return read_binary_arrays<int32_t>(p + l32, e, g);
    HN_ARG(false, "binary graph: different class identifier (expected intint or intlong)");
    return HN_OK;
}

struct OutFile {
    FILE* f = nullptr;
    ~OutFile() {
        if (f) fclose(f);
    }
};

template <typename Off>
hn_status write_binary_arrays(const hn_gstd& g, const char* cid, FILE* f) {
    const bool twice = (g.structure & S_DIRECTED) && (g.structure & S_INGOING) && !g.in_off.empty();
    const uint64_t blob[2] = {0, twice ? (uint64_t)(S_DIRECTED | S_INGOING) : (uint64_t)S_DIRECTED};
    const Off ne = (Off)g.ne;
    bool ok = fwrite(cid, 1, strlen(cid), f) == strlen(cid) && fwrite(&g.nv, 4, 1, f) == 1 &&
              fwrite(&ne, sizeof(Off), 1, f) == 1 && fwrite(blob, 8, 2, f) == 2;
    auto put_off = [&](const std::vector<int64_t>& o) {
        std::vector<Off> t(o.begin(), o.end());
        return fwrite(t.data(), sizeof(Off), t.size(), f) == t.size();
    };
    auto put_adj = [&](const std::vector<int32_t>& a) {
        return a.empty() || fwrite(a.data(), 4, a.size(), f) == a.size();
    };
    ok = ok && put_off(g.out_off) && (!twice || put_off(g.in_off)) && put_adj(g.out_adj) &&
         (!twice || put_adj(g.in_adj));
    HN_ARG(ok, "binary graph: write failed");
    return HN_OK;
}

char* put_uint(char* b, uint64_t x) {
    char t[24];
    int k = 0;
    do {
        t[k++] = (char)('0' + x % 10);
        x /= 10;
    } while (x);
    while (k) *b++ = t[--k];
    return b;
}

}  // namespace
}  // namespace hn

using namespace hn;

extern "C" {

hn_status hn_gstd_read(const char* path, uint32_t structure, uint32_t parse, uint64_t seed, hn_gstd** out) {
    HN_ARG(out != nullptr && path != nullptr, "null argument");
    *out = nullptr;
    HN_ARG((structure & ~(uint32_t)(S_DIRECTED | S_UNDIRECTED | S_INGOING)) == 0, "unknown structure flag");
    Mapped m;
    HN_TRY(map_file(path, m));
    std::string ext = extension(path);
    hn_gstd* g = new hn_gstd();
    hn_status s = HN_OK;
    if (ext == ".bin") {
        s = read_binary(m, *g);
    } else {
        Cur c{m.p, m.p + m.n};
        std::string first = c.word();
        FileCoo f;
        if (ext == ".mtx" && first == "%%MatrixMarket")
            s = read_market(m, f);
        else if (ext == ".gr" && (first == "c" || first == "p"))
            s = read_dimacs9(m, f);
        else {
            set_error(std::string("graph type not recognized: ") + path +
                      " (MatrixMarket .mtx, DIMACS9 .gr and binary .bin are read)");
            s = HN_E_INVALID_ARG;
        }
        if (s == HN_OK) s = coo_to_csr(f, structure, parse, seed, *g);
    }
    if (s != HN_OK) {
        delete g;
        return s;
    }
    *out = g;
    return HN_OK;
}

hn_status hn_gstd_from_csr(int32_t nv, int64_t ne, const int64_t* offsets, const int32_t* edges, hn_gstd** out) {
    HN_ARG(out != nullptr && offsets != nullptr && (ne == 0 || edges != nullptr), "null argument");
    *out = nullptr;
    HN_ARG(nv >= 0 && ne >= 0, "negative size");
    HN_ARG(offsets[0] == 0 && offsets[nv] == ne, "offsets[0] must be 0 and offsets[nv] == ne");
    for (int32_t v = 0; v < nv; v++) HN_ARG(offsets[v] <= offsets[v + 1], "offsets not monotone");
    for (int64_t j = 0; j < ne; j++) HN_ARG(edges[j] >= 0 && edges[j] < nv, "edge endpoint outside [0, nv)");
    hn_gstd* g = new hn_gstd();
    g->nv = nv;
    g->ne = ne;
    g->structure = S_UNDIRECTED;
    g->out_off.assign(offsets, offsets + nv + 1);
    g->out_adj.assign(edges, edges + ne);
    *out = g;
    return HN_OK;
}

hn_status hn_gstd_destroy(hn_gstd* g) {
    delete g;
    return HN_OK;
}

hn_status hn_gstd_info(const hn_gstd* g, int32_t* nv, int64_t* ne, uint32_t* structure) {
    HN_ARG(g != nullptr, "null graph");
    if (nv) *nv = g->nv;
    if (ne) *ne = g->ne;
    if (structure) *structure = g->structure;
    return HN_OK;
}

hn_status hn_gstd_csr_out(const hn_gstd* g, const int64_t** offsets, const int32_t** edges) {
    HN_ARG(g && offsets && edges, "null argument");
    *offsets = g->out_off.data();
    *edges = g->out_adj.data();
    return HN_OK;
}

hn_status hn_gstd_csr_in(const hn_gstd* g, const int64_t** offsets, const int32_t** edges) {
    HN_ARG(g && offsets && edges, "null argument");
    if (g->structure & S_DIRECTED) {
        HN_ARG(!g->in_off.empty(), "the graph was read without ENABLE_INGOING");
        *offsets = g->in_off.data();
        *edges = g->in_adj.data();
    } else {
        *offsets = g->out_off.data();
        *edges = g->out_adj.data();
    }
    return HN_OK;
}

hn_status hn_gstd_max_out_degree_id(const hn_gstd* g, int32_t* id) {
    HN_ARG(g && id, "null argument");
    int32_t best = 0;
    int64_t bd = -1;
    for (int32_t v = 0; v < g->nv; v++) {
        const int64_t d = g->out_off[(size_t)v + 1] - g->out_off[(size_t)v];
        if (d > bd) bd = d, best = v;
    }
    *id = best;
    return HN_OK;
}

hn_status hn_gstd_write_binary(const hn_gstd* g, const char* path, int32_t offset_bytes) {
    HN_ARG(g && path, "null argument");
    HN_ARG(offset_bytes == 4 || offset_bytes == 8, "offset_bytes must be 4 or 8");
    HN_ARG(offset_bytes == 8 || g->ne <= (int64_t)INT32_MAX, "more edges than int32 offsets hold");
    OutFile o;
    o.f = fopen(path, "wb");
    HN_ARG(o.f != nullptr, std::string("cannot create ") + path);
    if (offset_bytes == 4) return write_binary_arrays<int32_t>(*g, ID_INT32, o.f);
    return write_binary_arrays<int64_t>(*g, ID_INT64, o.f);
}

hn_status hn_gstd_write_market(const hn_gstd* g, const char* path) {
    HN_ARG(g && path, "null argument");
    OutFile o;
    o.f = fopen(path, "wb");
    HN_ARG(o.f != nullptr, std::string("cannot create ") + path);
    std::string head = "%%MatrixMarket matrix coordinate pattern general\n" + std::to_string(g->nv) + " " +
                       std::to_string(g->nv) + " " + std::to_string(g->ne) + "\n";
    bool ok = fwrite(head.data(), 1, head.size(), o.f) == head.size();
    std::vector<char> buf(1 << 20);
    size_t at = 0;
    for (int32_t v = 0; v < g->nv && ok; v++) {
        for (int64_t j = g->out_off[(size_t)v]; j < g->out_off[(size_t)v + 1]; j++) {
            if (at + 32 > buf.size()) {
                ok = ok && fwrite(buf.data(), 1, at, o.f) == at;
                at = 0;
            }
            char* b = buf.data() + at;
            char* q = put_uint(b, (uint64_t)v + 1);
            *q++ = ' ';
            q = put_uint(q, (uint64_t)g->out_adj[(size_t)j] + 1);
            *q++ = '\n';
            at += (size_t)(q - b);
        }
    }
    ok = ok && fwrite(buf.data(), 1, at, o.f) == at;
    HN_ARG(ok, "MatrixMarket: write failed");
    return HN_OK;
}

}  // extern "C"
