// yz_state.h -- the restart state file of a run (field*.state).
//
// Format (what the reference's reader and writer agree on,
// yuzen/src/IO/RestartState.h:24-49; the byte layout below is what its
// writer produces for the same contents):
//
//   header   ASCII lines ending in '\n':
//     1. the magic line "# Yuzen restart state 1"
//     2. one line per scalar key, keys in ascending byte order:
//          key SP v1 SP v2 ...   each value as printf "%.17g"
//        (17 significant digits: a double reads back to the same bits)
//     3. one line per array, in the order the arrays were first stored:
//          "vector" SP name SP count
//     4. one line per text key, keys in ascending byte order:
//          "text" SP key SP word
//     5. the line "binary"
//   payload  the arrays' values as raw little-endian IEEE doubles, array
//            after array in the order of their "vector" lines, nothing
//            between or after them.
//
// A reader skips blank lines, takes every other header line by its first
// word, keeps keys it does not know (the caller ignores them), and refuses a
// file without the magic line or the "binary" line, with a scalar value that
// is not a whole number token, or whose payload is shorter than its
// "vector" lines announce.
//
// Host code only. The file is assembled in memory and written in one piece
// to path + ".tmp", which is then renamed over path.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace yz {

class StateFile {
public:
    void set(const std::string &k, double v) { scal_[k] = std::vector<double>(1, v); }
    void set(const std::string &k, const std::vector<double> &v) { scal_[k] = v; }
    void set_vector(const std::string &k, const std::vector<double> &v) {
        for (auto &a : arrays_)
            if (a.name == k) { a.values = v; return; }
        arrays_.push_back(Array{k, v});
    }
    void set_text(const std::string &k, const std::string &v) { text_[k] = v; }

    double get(const std::string &k, size_t i = 0, double dflt = 0) const {
        const auto it = scal_.find(k);
        return (it != scal_.end() && i < it->second.size()) ? it->second[i] : dflt;
    }
    bool has(const std::string &k) const { return scal_.count(k) != 0; }
    bool has_vector(const std::string &k) const { return find(k) != nullptr; }
    const std::vector<double> &vec(const std::string &k) const {
        static const std::vector<double> none;
        const Array *a = find(k);
        return a ? a->values : none;
    }
    std::string text(const std::string &k) const {
        const auto it = text_.find(k);
        return it == text_.end() ? std::string() : it->second;
    }

    bool write(const std::string &path) const {
        std::string out = kMagic;
        out += " 1\n";
        char num[40];
        for (const auto &kv : scal_) {
            out += kv.first;
            for (double v : kv.second) {
                std::snprintf(num, sizeof(num), " %.17g", v);
                out += num;
            }
            out += '\n';
        }
        for (const Array &a : arrays_) {
            std::snprintf(num, sizeof(num), " %zu\n", a.values.size());
            out += "vector " + a.name + num;
        }
        for (const auto &kv : text_) out += "text " + kv.first + " " + kv.second + "\n";
        out += "binary\n";
        for (const Array &a : arrays_)
            out.append(reinterpret_cast<const char *>(a.values.data()), a.values.size() * sizeof(double));

        const std::string tmp = path + ".tmp";
        FILE *f = std::fopen(tmp.c_str(), "wb");
        if (!f) return false;
        const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
        if (std::fclose(f) != 0 || !ok || std::rename(tmp.c_str(), path.c_str()) != 0) {
            std::remove(tmp.c_str());
            return false;
        }
        return true;
    }

    // false when the file is missing, not a state file, or cut short
    bool read(const std::string &path) {
        scal_.clear(); arrays_.clear(); text_.clear();
        std::vector<char> buf;
        if (!slurp(path, buf)) return false;
        const char *p = buf.data(), *end = p + buf.size();
        std::string line;
        if (!next_line(p, end, line) || line.compare(0, std::strlen(kMagic), kMagic) != 0) return false;
        std::vector<std::pair<std::string, size_t>> announced;
        bool payload = false;
        while (next_line(p, end, line)) {
            if (line == "binary") { payload = true; break; }
            std::vector<std::string> w = words(line);
            if (w.empty()) continue;
            if (w[0] == "text") {
                text_[w.size() > 1 ? w[1] : std::string()] = w.size() > 2 ? w[2] : std::string();
            } else if (w[0] == "vector") {
                if (w.size() < 3) return false;
                char *e = nullptr;
                const unsigned long long n = std::strtoull(w[2].c_str(), &e, 10);
                if (e == w[2].c_str() || *e != '\0') return false;
                announced.emplace_back(w[1], (size_t)n);
            } else {
                std::vector<double> v;
                for (size_t i = 1; i < w.size(); ++i) {
                    char *e = nullptr;
                    v.push_back(std::strtod(w[i].c_str(), &e));
                    if (e == w[i].c_str() || *e != '\0') return false;
                }
                scal_[w[0]] = v;
            }
        }
        if (!payload) return false;
        size_t avail = (size_t)(end - p) / sizeof(double);
        for (const auto &a : announced) {
            if (a.second > avail) return false;
            avail -= a.second;
        }
        for (const auto &a : announced) {
            std::vector<double> v(a.second);
            if (a.second) std::memcpy(v.data(), p, a.second * sizeof(double));
            p += a.second * sizeof(double);
            set_vector(a.first, v);
        }
        return true;
    }

private:
    static constexpr const char *kMagic = "# Yuzen restart state";   // then " <version>"
    struct Array {
        std::string name;
        std::vector<double> values;
    };
    std::map<std::string, std::vector<double>> scal_;   // ascending key order
    std::vector<Array> arrays_;                           // order first stored
    std::map<std::string, std::string> text_;

    const Array *find(const std::string &k) const {
        for (const Array &a : arrays_)
            if (a.name == k) return &a;
        return nullptr;
    }
    static bool slurp(const std::string &path, std::vector<char> &buf) {
        FILE *f = std::fopen(path.c_str(), "rb");
        if (!f) return false;
        char chunk[1 << 16];
        size_t got;
        while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
        const bool ok = !std::ferror(f);
        std::fclose(f);
        return ok;
    }
    // the next '\n'-terminated line (without it); false at the end
    static bool next_line(const char *&p, const char *end, std::string &line) {
        if (p >= end) return false;
        const char *q = static_cast<const char *>(std::memchr(p, '\n', (size_t)(end - p)));
        const char *stop = q ? q : end;
        line.assign(p, stop);
        p = q ? q + 1 : end;
        return true;
    }
    static std::vector<std::string> words(const std::string &s) {
        std::vector<std::string> w;
        size_t i = 0;
        while (i < s.size()) {
            while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\r')) ++i;
            size_t j = i;
            while (j < s.size() && s[j] != ' ' && s[j] != '\t' && s[j] != '\r') ++j;
            if (j > i) w.emplace_back(s, i, j - i);
            i = j;
        }
        return w;
    }
};

}  // namespace yz
