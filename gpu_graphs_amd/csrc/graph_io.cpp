// Graph file formats: the reference edge-list text and a binary CSR cache.
//
// Text format (1D Partition/src/csr.cpp:16-29): first token `num_nodes`,
// second `num_edges`, then whitespace-separated `from to` pairs.  The
// reference reads it with `ifstream >>` one token at a time and trusts the
// header's edge count for its malloc (csr.cpp:22).  Here the file is read in
// one block and parsed by hand; the header's m is only a capacity hint, ids
// are bounds-checked, and malformed text is reported instead of silently
// ending the read.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>

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

std::vector<char> read_all(const std::string& path) {
    FilePtr f(std::fopen(path.c_str(), "rb"));
    if (!f) fail(GB_EIO, "unable to open file " + path);
    std::vector<char> buf;
    if (std::fseek(f.get(), 0, SEEK_END) == 0) {
        long sz = std::ftell(f.get());
        if (sz > 0) {
            buf.resize(static_cast<size_t>(sz) + 1);
            std::rewind(f.get());
            size_t got = std::fread(buf.data(), 1, static_cast<size_t>(sz), f.get());
            buf.resize(got);
        }
    }
    buf.push_back('\0');
    return buf;
}

// Token scanner over a NUL-terminated buffer.  Lines starting with '#' or
// '%' (SNAP / MatrixMarket comments) are skipped.
struct Scanner {
    const char* p;
    int64_t line = 1;

    bool next(int64_t& out, bool& malformed) {
        malformed = false;
        for (;;) {
            while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n') {
                if (*p == '\n') ++line;
                ++p;
            }
            if (*p == '#' || *p == '%') {
                while (*p && *p != '\n') ++p;
                continue;
            }
            break;
        }
        if (!*p) return false;
        bool neg = false;
        if (*p == '-' || *p == '+') {
            neg = *p == '-';
            ++p;
        }
        if (*p < '0' || *p > '9') {
            malformed = true;
            return false;
        }
        int64_t v = 0;
        while (*p >= '0' && *p <= '9') {
            v = v * 10 + (*p - '0');
            if (v > (int64_t(1) << 40)) {
                malformed = true;
                return false;
            }
            ++p;
        }
        if (*p && *p != ' ' && *p != '\t' && *p != '\r' && *p != '\n') {
            malformed = true;
            return false;
        }
        out = neg ? -v : v;
        return true;
    }
};

}  // namespace

HostGraph load_edgelist(const std::string& path, uint32_t flags) {
    std::vector<char> buf = read_all(path);
    Scanner sc{buf.data()};
    int64_t n = 0, m_hint = 0;
    bool bad = false;
    if (!sc.next(n, bad) || !sc.next(m_hint, bad))
        fail(GB_EIO, path + ": missing `num_nodes num_edges` header");
    if (n < 0 || n > INT32_MAX) fail(GB_EIO, path + ": num_nodes out of range");
    const int64_t shift = (flags & GB_ONE_BASED) ? 1 : 0;
    const bool auto_n = flags & GB_AUTO_N;

    std::vector<int32_t> src, dst;
    if (m_hint > 0 && m_hint < (int64_t(1) << 32)) {
        src.reserve(m_hint);
        dst.reserve(m_hint);
    }
    int64_t a = 0, b = 0;
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
