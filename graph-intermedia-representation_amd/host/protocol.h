// Readers and the result-row writer of the streaming protocol's files: the
// text formats of experiments/baselines/common/protocol.h (loaded.el "u v",
// *.ops "a u v" / "d u v", sources.txt, meta.txt "key value", CSV rows
// dataset,system,metric,mode,value,unit,count,threads).
#pragma once
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace stream_proto {

struct Op {
    bool insert;
    int u, v;
};

// Whole file into memory.
inline std::string slurp(const std::string& path) {
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + path);
    std::string s;
    char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, n);
    std::fclose(f);
    return s;
}

// Unsigned decimal numbers of a text, in order (any other byte separates).
struct Numbers {
    const char* p;
    const char* e;
    bool next(uint32_t& x) {
        while (p < e && (*p < '0' || *p > '9')) ++p;
        if (p >= e) return false;
        uint64_t v = 0;
        while (p < e && *p >= '0' && *p <= '9') v = v * 10 + (uint64_t)(*p++ - '0');
        x = (uint32_t)v;
        return true;
    }
};

inline std::vector<std::pair<int, int>> readEdges(const std::string& path) {
    const std::string t = slurp(path);
    Numbers r{t.data(), t.data() + t.size()};
    std::vector<std::pair<int, int>> out;
    out.reserve(t.size() / 12);
    uint32_t a, b;
    while (r.next(a) && r.next(b)) out.push_back({(int)a, (int)b});
    return out;
}

inline std::vector<Op> readOps(const std::string& path) {
    const std::string t = slurp(path);
    std::vector<Op> out;
    const char* p = t.data();
    const char* e = p + t.size();
    while (p < e) {
        while (p < e && *p != 'a' && *p != 'd') ++p;
        if (p >= e) break;
        const bool ins = *p++ == 'a';
        Numbers r{p, e};
        uint32_t a, b;
        if (!r.next(a) || !r.next(b)) break;
        p = r.p;
        out.push_back({ins, (int)a, (int)b});
    }
    return out;
}

inline std::vector<int> readInts(const std::string& path) {
    const std::string t = slurp(path);
    Numbers r{t.data(), t.data() + t.size()};
    std::vector<int> out;
    uint32_t a;
    while (r.next(a)) out.push_back((int)a);
    return out;
}

inline long long metaValue(const std::string& dir, const std::string& key) {
    std::ifstream f(dir + "/meta.txt");
    std::string k;
    long long v;
    while (f >> k >> v)
        if (k == key) return v;
    throw std::runtime_error("meta.txt: no key " + key);
}

inline void emitRow(const std::string& csv, const std::string& dataset, const std::string& system,
                    const std::string& metric, const std::string& mode, double value, const char* unit,
                    long long count, int threads) {
    if (std::FILE* f = std::fopen(csv.c_str(), "a")) {
        std::fprintf(f, "%s,%s,%s,%s,%.6g,%s,%lld,%d\n", dataset.c_str(), system.c_str(), metric.c_str(),
                     mode.c_str(), value, unit, count, threads);
        std::fclose(f);
    }
    std::printf("  %-26s %-16s %.6g %s (n=%lld)\n", metric.c_str(), mode.c_str(), value, unit, count);
}

}  // namespace stream_proto
