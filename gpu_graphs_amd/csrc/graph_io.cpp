// Graph file formats: the reference edge-list text and a binary CSR cache.
//
// Text format (1D Partition/src/csr.cpp:16-29): first token `num_nodes`,
// second `num_edges`, then whitespace-separated `from to` pairs.  The
// reference reads it with `ifstream >>` one token at a time and trusts the
// header's edge count for its malloc (csr.cpp:22).  Here the file is read in
// one block and parsed by hand, in parallel: the body is cut into slices at
// line starts, each worker scans its slice into validated int32 tokens, and
// tokens are paired by their global index, so a pair may span lines exactly
// as with `>>`.  The header's m is only a capacity hint, ids are
// bounds-checked, and malformed text is reported (with its line) instead of
// silently ending the read.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <thread>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "common.hpp"
#include "host_graph.hpp"

namespace gg {

namespace {

struct FileCloser {
    void operator()(FILE* f) const {
        if (f) std::fclose(f);
    }
};
using FilePtr = std::unique_ptr<FILE, FileCloser>;

// Read-only private mapping of a whole file (pages come from the page cache;
// no copy of the text is made).
class MappedFile {
   public:
    explicit MappedFile(const std::string& path) {
        const int fd = ::open(path.c_str(), O_RDONLY);
        if (fd < 0) fail(GB_EIO, "unable to open file " + path);
        struct stat st;
        if (::fstat(fd, &st) != 0) {
            ::close(fd);
            fail(GB_EIO, "unable to stat file " + path);
        }
        size_ = static_cast<size_t>(st.st_size);
        if (size_ > 0) {
            void* p = ::mmap(nullptr, size_, PROT_READ, MAP_PRIVATE | MAP_POPULATE, fd, 0);
            if (p == MAP_FAILED) {
                ::close(fd);
                fail(GB_EIO, "unable to map file " + path);
            }
            data_ = static_cast<const char*>(p);
        }
        ::close(fd);
    }
    ~MappedFile() {
        if (data_) ::munmap(const_cast<char*>(data_), size_);
    }
    MappedFile(const MappedFile&) = delete;
    MappedFile& operator=(const MappedFile&) = delete;
    const char* begin() const { return data_ ? data_ : ""; }
    const char* end() const { return begin() + size_; }

   private:
    const char* data_ = nullptr;
    size_t size_ = 0;
};

// Anonymous array in transparent huge pages: the parser's edge arrays are
// first touched by several threads at once, and 2 MB pages take 512x fewer
// page faults than 4 KB ones.
template <class T>
class HugeArray {
   public:
    explicit HugeArray(int64_t count) {
        constexpr size_t kHuge = size_t(2) << 20;
        bytes_ = (static_cast<size_t>(count) * sizeof(T) + kHuge - 1) / kHuge * kHuge;
        if (bytes_ == 0) return;
        void* p = ::mmap(nullptr, bytes_, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
        if (p == MAP_FAILED) fail(GB_ENOMEM, "out of host memory for the edge list");
        (void)::madvise(p, bytes_, MADV_HUGEPAGE);  // advice only: 4 KB pages also work
        data_ = static_cast<T*>(p);
    }
    ~HugeArray() {
        if (data_) ::munmap(data_, bytes_);
    }
    HugeArray(const HugeArray&) = delete;
    HugeArray& operator=(const HugeArray&) = delete;
    T* data() { return data_; }
    T& operator[](int64_t i) { return data_[i]; }

   private:
    T* data_ = nullptr;
    size_t bytes_ = 0;
};

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n'; }

// Token scanner over [p, end).  Lines starting with '#' or '%' (SNAP /
// MatrixMarket comments) are skipped, as is the rest of a line after a token
// position holding one of them.
struct Scanner {
    const char* p;
    const char* end;
    int64_t line = 1;

    bool next(int64_t& out, bool& malformed) {
        malformed = false;
        for (;;) {
            while (p < end && is_space(*p)) {
                if (*p == '\n') ++line;
                ++p;
            }
            if (p < end && (*p == '#' || *p == '%')) {
                while (p < end && *p != '\n') ++p;
                continue;
            }
            break;
        }
        if (p >= end) return false;
        bool neg = false;
        if (*p == '-' || *p == '+') {
            neg = *p == '-';
            ++p;
        }
        if (p >= end || *p < '0' || *p > '9') {
            malformed = true;
            return false;
        }
        int64_t v = 0;
        while (p < end && *p >= '0' && *p <= '9') {
            v = v * 10 + (*p - '0');
            if (v > (int64_t(1) << 40)) {
                malformed = true;
                return false;
            }
            ++p;
        }
        if (p < end && !is_space(*p)) {
            malformed = true;
            return false;
        }
        out = neg ? -v : v;
        return true;
    }
};

// Serial scan of the body; throws the first error in file order.  Also the
// reference for the parallel scan: any slice that meets an error hands the
// whole body to this function, so messages and line numbers never depend on
// the number of workers.
HostGraph parse_serial(const std::string& path, Scanner sc, int64_t n, int64_t m_hint, uint32_t flags) {
    const int64_t shift = (flags & GB_ONE_BASED) ? 1 : 0;
    const bool auto_n = flags & GB_AUTO_N;
    std::vector<int32_t> src, dst;
    if (m_hint > 0 && m_hint < (int64_t(1) << 32)) {
        src.reserve(m_hint);
        dst.reserve(m_hint);
    }
    int64_t a = 0, b = 0;
    bool bad = false;
    for (;;) {
        if (!sc.next(a, bad)) {
            if (bad) fail(GB_EIO, path + ": malformed token on line " + std::to_string(sc.line));
            break;
        }
        int64_t line = sc.line;
        if (!sc.next(b, bad)) {
            if (bad) fail(GB_EIO, path + ": malformed token on line " + std::to_string(sc.line));
            fail(GB_EIO, path + ": dangling vertex id on line " + std::to_string(line));
        }
        a -= shift;
        b -= shift;
        if (auto_n && a >= 0 && b >= 0 && a < INT32_MAX && b < INT32_MAX) n = std::max(n, std::max(a, b) + 1);
        if (a < 0 || a >= n || b < 0 || b >= n)
            fail(GB_ERANGE, path + ":" + std::to_string(line) + ": edge (" + std::to_string(a + shift) + ", " +
                                std::to_string(b + shift) + ") outside the " + std::to_string(n) + "-vertex id range" +
                                (shift ? " (one-based)" : " (zero-based)"));
        src.push_back(static_cast<int32_t>(a));
        dst.push_back(static_cast<int32_t>(b));
    }
    return build_csr(n, src.data(), dst.data(), static_cast<int64_t>(src.size()), flags);
}

// Bodies below this size are scanned by one thread.
constexpr int64_t kParallelParseBytes = int64_t(1) << 20;

}  // namespace

HostGraph load_edgelist(const std::string& path, uint32_t flags) {
    MappedFile file(path);
    const char* end = file.end();
    Scanner sc{file.begin(), end};
    int64_t n = 0, m_hint = 0;
    bool bad = false;
    if (!sc.next(n, bad) || !sc.next(m_hint, bad))
        fail(GB_EIO, path + ": missing `num_nodes num_edges` header");
    if (n < 0 || n > INT32_MAX) fail(GB_EIO, path + ": num_nodes out of range");
    const int64_t body_bytes = end - sc.p;
    const int workers = host_threads();
    if (body_bytes < kParallelParseBytes || workers == 1) return parse_serial(path, sc, n, m_hint, flags);

    // Slices start at line starts; a token never spans a newline.
    std::vector<const char*> cut(workers + 1);
    cut[0] = sc.p;
    cut[workers] = end;
    for (int w = 1; w < workers; ++w) {
        const char* q = sc.p + body_bytes * w / workers;
        while (q < end && q[-1] != '\n') ++q;
        cut[w] = std::max(q, cut[w - 1]);
    }
    const int64_t shift = (flags & GB_ONE_BASED) ? 1 : 0;
    const int64_t limit = (flags & GB_AUTO_N) ? int64_t(INT32_MAX) : n;
    auto run = [workers](auto&& fn) {
        std::vector<std::thread> pool;
        pool.reserve(workers);
        for (int w = 0; w < workers; ++w) pool.emplace_back(fn, w);
        for (auto& t : pool) t.join();
    };

    // Pass 1: count and validate each slice's tokens (no stores).
    std::vector<int64_t> count(workers, 0), max_id(workers, -1);
    std::atomic<bool> any_bad{false};
    run([&](int w) {
        Scanner s{cut[w], cut[w + 1]};
        int64_t t = 0, c = 0, hi = -1;
        bool malformed = false;
        while (s.next(t, malformed)) {
            t -= shift;
            if (t < 0 || t >= limit) {
                any_bad = true;
                return;
            }
            hi = std::max(hi, t);
            if ((++c & 0xffff) == 0 && any_bad.load(std::memory_order_relaxed)) return;
        }
        if (malformed) any_bad = true;
        count[w] = c;
        max_id[w] = hi;
    });
    std::vector<int64_t> base(workers + 1, 0);
    for (int w = 0; w < workers; ++w) base[w + 1] = base[w] + count[w];
    if (any_bad || (base[workers] & 1)) return parse_serial(path, sc, n, m_hint, flags);  // first error in file order
    if (flags & GB_AUTO_N)
        for (int64_t hi : max_id) n = std::max(n, hi + 1);

    // Pass 2: token g of the body is the source (g even) or target (g odd)
    // of pair g/2, so every slice writes its pairs in place.
    const int64_t m = base[workers] / 2;
    HugeArray<int32_t> src(m), dst(m);
    run([&](int w) {
        Scanner s{cut[w], cut[w + 1]};
        int64_t t = 0, g = base[w];
        bool malformed = false;
        while (s.next(t, malformed)) {
            (g & 1 ? dst : src)[g >> 1] = static_cast<int32_t>(t - shift);
            ++g;
        }
    });
    return build_csr(n, src.data(), dst.data(), m, flags);
}

void write_edgelist(const HostGraph& g, const std::string& path) {
    FilePtr f(std::fopen(path.c_str(), "wb"));
    if (!f) fail(GB_EIO, "unable to create file " + path);
    std::string out;
    out.reserve(1 << 22);
    out += std::to_string(g.n) + "\n" + std::to_string(g.m()) + "\n";
    char tmp[32];
    for (int64_t v = 0; v < g.n; ++v) {
        for (int64_t j = g.offsets[v]; j < g.offsets[v + 1]; ++j) {
            int k = std::snprintf(tmp, sizeof(tmp), "%lld %d\n", static_cast<long long>(v), g.cols[j]);
            out.append(tmp, k);
        }
        if (out.size() > (1u << 22)) {
            if (std::fwrite(out.data(), 1, out.size(), f.get()) != out.size()) fail(GB_EIO, "write failed: " + path);
            out.clear();
        }
    }
    if (!out.empty() && std::fwrite(out.data(), 1, out.size(), f.get()) != out.size())
        fail(GB_EIO, "write failed: " + path);
}

namespace {
const char kMagic[8] = {'G', 'B', 'C', 'S', 'R', '0', '0', '1'};
}

void save_binary(const HostGraph& g, const std::string& path) {
    FilePtr f(std::fopen(path.c_str(), "wb"));
    if (!f) fail(GB_EIO, "unable to create file " + path);
    int64_t hdr[2] = {g.n, g.m()};
    uint32_t fl[2] = {g.flags, g.symmetric ? 1u : 0u};
    bool ok = std::fwrite(kMagic, 1, 8, f.get()) == 8 && std::fwrite(hdr, sizeof(hdr), 1, f.get()) == 1 &&
              std::fwrite(fl, sizeof(fl), 1, f.get()) == 1 &&
              std::fwrite(g.offsets.data(), sizeof(int64_t), g.offsets.size(), f.get()) == g.offsets.size() &&
              std::fwrite(g.cols.data(), sizeof(int32_t), g.cols.size(), f.get()) == g.cols.size();
    if (!ok) fail(GB_EIO, "write failed: " + path);
}

HostGraph load_binary(const std::string& path) {
    FilePtr f(std::fopen(path.c_str(), "rb"));
    if (!f) fail(GB_EIO, "unable to open file " + path);
    char magic[8];
    int64_t hdr[2];
    uint32_t fl[2];
    if (std::fread(magic, 1, 8, f.get()) != 8 || std::memcmp(magic, kMagic, 8) != 0)
        fail(GB_EIO, path + ": not a binary CSR file");
    if (std::fread(hdr, sizeof(hdr), 1, f.get()) != 1 || std::fread(fl, sizeof(fl), 1, f.get()) != 1)
        fail(GB_EIO, path + ": truncated header");
    if (hdr[0] < 0 || hdr[0] > INT32_MAX || hdr[1] < 0) fail(GB_EIO, path + ": bad header");
    HostGraph g;
    g.n = hdr[0];
    g.flags = fl[0];
    g.symmetric = fl[1] != 0;
    g.offsets.resize(g.n + 1);
    g.cols.resize(hdr[1]);
    if (std::fread(g.offsets.data(), sizeof(int64_t), g.offsets.size(), f.get()) != g.offsets.size() ||
        std::fread(g.cols.data(), sizeof(int32_t), g.cols.size(), f.get()) != g.cols.size())
        fail(GB_EIO, path + ": truncated body");
    if (g.offsets[0] != 0 || g.offsets[g.n] != hdr[1]) fail(GB_EIO, path + ": inconsistent offsets");
    for (int64_t v = 0; v < g.n; ++v)
        if (g.offsets[v + 1] < g.offsets[v]) fail(GB_EIO, path + ": offsets not monotone");
    for (int32_t c : g.cols)
        if (c < 0 || c >= g.n) fail(GB_EIO, path + ": column id out of range");
    return g;
}

}  // namespace gg
