// yz_state.h -- the restart state file of a run (field*.state), in the
// reference's layout (yuzen/src/IO/RestartState.h:24-49, RestartState.cpp
// :43-146): a magic line, text lines "key v1 v2 ..." of scalars (17
// significant digits, which a double survives exactly; keys in sorted
// order), "vector name n" lines announcing the arrays in the order they were
// set, "text key word" lines, a line "binary", then the arrays as raw
// little-endian doubles. Unknown keys are ignored by a reader, so a file can
// carry more than the reference's keys.
//
// Host code only.
#pragma once
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

namespace yz {

struct StateFile {
    std::map<std::string, std::vector<double>> scalars, vectors;
    std::vector<std::string> order;
    std::map<std::string, std::string> texts;

    void set(const std::string &k, double v) { scalars[k] = std::vector<double>(1, v); }
    void set(const std::string &k, const std::vector<double> &v) { scalars[k] = v; }
    void set_vector(const std::string &k, const std::vector<double> &v) {
        if (!vectors.count(k)) order.push_back(k);
        vectors[k] = v;
    }
    void set_text(const std::string &k, const std::string &v) { texts[k] = v; }
    double get(const std::string &k, size_t i = 0, double dflt = 0) const {
        auto it = scalars.find(k);
        if (it == scalars.end() || it->second.size() <= i) return dflt;
        return it->second[i];
    }
    bool has_vector(const std::string &k) const { return vectors.count(k) != 0; }
    const std::vector<double> &vec(const std::string &k) const {
        static const std::vector<double> empty;
        auto it = vectors.find(k);
        return it == vectors.end() ? empty : it->second;
    }
    std::string text(const std::string &k) const {
        auto it = texts.find(k);
        return it == texts.end() ? std::string() : it->second;
    }

    static const char *magic() { return "# Yuzen restart state"; }

    // to path + ".tmp", then renamed over path
    bool write(const std::string &path) const {
        const std::string tmp = path + ".tmp";
        {
            std::ofstream f(tmp.c_str(), std::ios::binary);
            if (!f.good()) return false;
            f << magic() << " 1\n";
            char buf[64];
            for (const auto &kv : scalars) {
                f << kv.first;
                for (double v : kv.second) {
                    std::snprintf(buf, sizeof(buf), " %.17g", v);
                    f << buf;
                }
                f << "\n";
            }
            for (const auto &k : order) f << "vector " << k << " " << vectors.find(k)->second.size() << "\n";
            for (const auto &kv : texts) f << "text " << kv.first << " " << kv.second << "\n";
            f << "binary\n";
            for (const auto &k : order) {
                const std::vector<double> &v = vectors.find(k)->second;
                if (!v.empty()) f.write((const char *)v.data(), (std::streamsize)(v.size() * sizeof(double)));
            }
            f.close();
            if (f.fail()) { std::remove(tmp.c_str()); return false; }
        }
        if (std::rename(tmp.c_str(), path.c_str()) != 0) { std::remove(tmp.c_str()); return false; }
        return true;
    }

    // false when the file is missing, not a state file, or cut short
    bool read(const std::string &path) {
        scalars.clear(); vectors.clear(); order.clear(); texts.clear();
        std::ifstream f(path.c_str(), std::ios::binary);
        if (!f.good()) return false;
        std::string line;
        const std::string m = magic();
        if (!std::getline(f, line) || line.compare(0, m.size(), m) != 0) return false;
        std::vector<std::pair<std::string, size_t>> arrays;
        bool binary = false;
        while (std::getline(f, line)) {
            if (line == "binary") { binary = true; break; }
            std::istringstream is(line);
            std::string key;
            if (!(is >> key)) continue;
            if (key == "text") {
                std::string name, val;
                is >> name >> val;
                texts[name] = val;
                continue;
            }
            if (key == "vector") {
                std::string name;
                size_t n = 0;
                if (!(is >> name >> n)) return false;
                arrays.push_back(std::make_pair(name, n));
                continue;
            }
            std::vector<double> v;
            std::string tok;
            while (is >> tok) {
                char *end = nullptr;
                v.push_back(std::strtod(tok.c_str(), &end));
                if (end == tok.c_str() || *end != '\0') return false;
            }
            scalars[key] = v;
        }
        if (!binary) return false;
        const std::streampos here = f.tellg();
        f.seekg(0, std::ios::end);
        const std::streampos end = f.tellg();
        f.seekg(here);
        if (here < 0 || end < here) return false;
        size_t left = (size_t)(end - here);
        for (const auto &a : arrays) {
            if (a.second > left / sizeof(double)) return false;
            left -= a.second * sizeof(double);
        }
        for (const auto &a : arrays) {
            std::vector<double> v(a.second);
            if (!v.empty()) f.read((char *)v.data(), (std::streamsize)(v.size() * sizeof(double)));
            if (!f) return false;
            set_vector(a.first, v);
        }
        return true;
    }
};

}  // namespace yz
