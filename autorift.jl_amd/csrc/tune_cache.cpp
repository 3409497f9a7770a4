// Persisted tuner choices: one small text file per device model and library version, read and
// written without ever failing the caller. The reference keeps its FFTW wisdom the same way
// (src/plans.jl:133-257): the directory is resolved on every call (so a caller may redirect it
// at any time), a missing, unreadable or corrupt file reads as empty, and a write goes to a
// private temporary that is renamed over the file, so two processes writing at once never leave
// a half-written file behind. A write that fails (read-only directory, full disk) is dropped:
// tuning is an optimisation, and the next process tunes again.
//
// File format: a header line, then one "key<TAB>value" line per entry (value a decimal int32).
// Lines that do not parse are skipped; a file whose header differs is treated as empty.
#include <cerrno>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include <sys/stat.h>
#include <sys/types.h>
#include <unistd.h>

#include "../../include/autorift_hip.h"

namespace {

const char kHeader[] = "autorift_hip tune cache 1";
const char kVersion[] = "0.1.0";
std::mutex g_mu;

bool usable_name(const char* s) {
    if (!s || !*s) return false;
    for (const char* p = s; *p; p++)
        if (*p == '\t' || *p == '\n' || *p == '\r') return false;
    return true;
}

// mkdir -p; true when the directory exists afterwards
bool make_dirs(const std::string& d) {
    if (d.empty()) return false;
    std::string cur;
    size_t i = 0;
    while (i <= d.size()) {
        const size_t j = d.find('/', i);
        cur = d.substr(0, j == std::string::npos ? d.size() : j);
        if (!cur.empty()) {
            struct stat st;
            if (stat(cur.c_str(), &st) != 0) {
                if (mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) return false;
            } else if (!S_ISDIR(st.st_mode)) {
                return false;
            }
        }
        if (j == std::string::npos) break;
        i = j + 1;
    }
    return true;
}

// The cache directory: $AR_TUNE_CACHE_DIR, else $XDG_CACHE_HOME/autorift_hip, else
// $HOME/.cache/autorift_hip; "" when caching is off (AR_TUNE_CACHE=off) or no directory is usable.
std::string cache_dir() {
    const char* off = std::getenv("AR_TUNE_CACHE");
    if (off && std::strcmp(off, "off") == 0) return "";
    std::string d;
    if (const char* e = std::getenv("AR_TUNE_CACHE_DIR"); e && *e) d = e;
    else if (const char* x = std::getenv("XDG_CACHE_HOME"); x && *x) d = std::string(x) + "/autorift_hip";
    else if (const char* h = std::getenv("HOME"); h && *h) d = std::string(h) + "/.cache/autorift_hip";
    if (d.empty() || !make_dirs(d)) return "";
    return d;
}

// File for a device string: characters outside [A-Za-z0-9._-] become '_'.
std::string cache_file(const char* device) {
    const std::string d = cache_dir();
    if (d.empty() || !usable_name(device)) return "";
    std::string name;
    for (const char* p = device; *p; p++) {
        const char c = *p;
        const bool ok = (c >= 'A' && c <= 'Z') || (c >= 'a' && c <= 'z') || (c >= '0' && c <= '9') ||
                        c == '.' || c == '_' || c == '-';
        name += ok ? c : '_';
    }
    return d + "/" + name + "-autorift_hip" + kVersion + ".tune";
}

// Every entry of the file that parses; empty for a missing, unreadable or foreign file.
std::map<std::string, int32_t> read_entries(const std::string& path) {
    std::map<std::string, int32_t> m;
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) return m;
    char line[1024];
    bool header = false;
    if (std::fgets(line, sizeof line, f)) {
        line[std::strcspn(line, "\r\n")] = 0;
        header = std::strcmp(line, kHeader) == 0;
    }
    while (header && std::fgets(line, sizeof line, f)) {
        if (!std::strchr(line, '\n') && !std::feof(f)) {     // over-long line: skip its rest
            int c;
            while ((c = std::fgetc(f)) != EOF && c != '\n') {}
            continue;
        }
        line[std::strcspn(line, "\r\n")] = 0;
        char* tab = std::strchr(line, '\t');
        if (!tab || tab == line) continue;
        *tab = 0;
        const char* v = tab + 1;
        if (!*v) continue;
        char* end = nullptr;
        errno = 0;
        const long x = std::strtol(v, &end, 10);
        if (errno != 0 || *end != 0 || x < INT32_MIN || x > INT32_MAX) continue;
        m[line] = (int32_t)x;
    }
    std::fclose(f);
    return m;
}

bool write_entries(const std::string& path, const std::map<std::string, int32_t>& m) {
    const std::string tmp = path + "." + std::to_string((long)getpid()) + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "w");
    if (!f) return false;
    bool ok = std::fprintf(f, "%s\n", kHeader) > 0;
    for (const auto& kv : m)
        ok = ok && std::fprintf(f, "%s\t%d\n", kv.first.c_str(), (int)kv.second) > 0;
    ok = (std::fflush(f) == 0) && ok;
    ok = (std::fclose(f) == 0) && ok;
    if (ok && std::rename(tmp.c_str(), path.c_str()) != 0) ok = false;
    if (!ok) std::remove(tmp.c_str());
    return ok;
}

}  // namespace

extern "C" {

int64_t ar_tune_cache_path(const char* device, char* buf, int64_t len) {
    std::lock_guard<std::mutex> lk(g_mu);
    const std::string p = cache_file(device);
    if (p.empty()) return 0;
    if (buf && len > 0) {
        const size_t n = p.size() < (size_t)(len - 1) ? p.size() : (size_t)(len - 1);
        std::memcpy(buf, p.data(), n);
        buf[n] = 0;
    }
    return (int64_t)p.size();
}

int32_t ar_tune_cache_lookup(const char* device, const char* key, int32_t* value) {
    if (!usable_name(key) || !value) return 0;
    std::lock_guard<std::mutex> lk(g_mu);
    const std::string p = cache_file(device);
    if (p.empty()) return 0;
    const auto m = read_entries(p);
    const auto it = m.find(key);
    if (it == m.end()) return 0;
    *value = it->second;
    return 1;
}

int32_t ar_tune_cache_record(const char* device, const char* key, int32_t value) {
    if (!usable_name(key) || std::strlen(key) > 512) return 0;
    std::lock_guard<std::mutex> lk(g_mu);
    const std::string p = cache_file(device);
    if (p.empty()) return 0;
    auto m = read_entries(p);         // a corrupt file is replaced by what parses plus this entry
    m[key] = value;
    return write_entries(p, m) ? 1 : 0;
}

}  // extern "C"
