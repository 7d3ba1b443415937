// Common pieces of the C++ host layer over libpasc_hip.so: the HipVector backend tag, status
// checking in the reference's print-and-continue style, options with the reference's flag names,
// timers and PETSc binary Vec I/O.
#pragma once

#include <pasc_hip.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace pascinference {

// Backend tag for the class templates (the reference instantiates them with PetscVector).  The
// vectors of the gamma QP live in device memory owned by the library's pasc_problem.
struct HipVector {
    static std::string get_name() { return "HipVector"; }
};

// TRYCXX (include/external/petscvector/petscvector.h:51) prints a failed call and continues;
// PASC_TRY does the same for the C ABI and evaluates to true on success.
inline bool pasc_try(int rc, const char *what, const char *file, int line) {
    if (rc == PASC_OK) return true;
    std::cerr << "pasc error " << rc << " in " << what << " (" << file << ":" << line
              << "): " << pasc_last_error() << std::endl;
    return false;
}
#define PASC_TRY(expr) ::pascinference::pasc_try((expr), #expr, __FILE__, __LINE__)

// Console options: "--name=value" or "--name value", repeated names build lists
// (the names follow src/boost/options.cpp).
class Options {
  public:
    Options() {}
    Options(int argc, char **argv) {
        for (int i = 1; i < argc; i++) {
            std::string a = argv[i];
            if (a.rfind("--", 0) != 0) continue;
            a = a.substr(2);
            std::string v;
            const size_t eq = a.find('=');
            if (eq != std::string::npos) {
                v = a.substr(eq + 1);
                a = a.substr(0, eq);
            } else if (i + 1 < argc && std::string(argv[i + 1]).rfind("--", 0) != 0) {
                v = argv[++i];
            } else {
                v = "1";
            }
            vals_[a].push_back(v);
        }
    }
    bool has(const std::string &n) const { return vals_.count(n) > 0; }
    template <typename T> bool set_option_value(const std::string &n, T *out, T def) const {
        auto it = vals_.find(n);
        if (it == vals_.end()) {
            *out = def;
            return false;
        }
        std::istringstream ss(it->second.back());
        ss >> *out;
        return true;
    }
    bool set_option_value(const std::string &n, std::string *out, const std::string &def) const {
        auto it = vals_.find(n);
        *out = it == vals_.end() ? def : it->second.back();
        return it != vals_.end();
    }
    bool set_option_value(const std::string &n, bool *out, bool def) const {
        auto it = vals_.find(n);
        if (it == vals_.end()) {
            *out = def;
            return false;
        }
        const std::string &v = it->second.back();
        *out = !(v == "0" || v == "false" || v == "False");
        return true;
    }
    bool set_option_list(const std::string &n, std::vector<double> *out) const {
        auto it = vals_.find(n);
        if (it == vals_.end()) return false;
        out->clear();
        for (auto &v : it->second) {
            std::istringstream ss(v);
            double d;
            while (ss >> d) out->push_back(d);
        }
        return true;
    }

  private:
    std::map<std::string, std::vector<std::string>> vals_;
};

// Timer with the reference's start/stop/get_value_sum/get_value_last (common/timer.h).
class Timer {
  public:
    void restart() { sum_ = last_ = 0.0; running_ = false; }
    void start() { t0_ = std::chrono::steady_clock::now(); running_ = true; }
    void stop() {
        if (!running_) return;
        last_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
        sum_ += last_;
        running_ = false;
    }
    double get_value_sum() const { return sum_; }
    double get_value_last() const { return last_; }

  private:
    std::chrono::steady_clock::time_point t0_;
    double sum_ = 0.0, last_ = 0.0;
    bool running_ = false;
};

// PETSc binary Vec files (big-endian int32 class id 1211214, int32 n, n big-endian float64),
// the reference's on-disk format (GeneralVector::load_global / save_binary).
namespace petscio {
constexpr int32_t VEC_FILE_CLASSID = 1211214;

inline uint32_t bswap32(uint32_t v) {
    return (v >> 24) | ((v >> 8) & 0xFF00u) | ((v << 8) & 0xFF0000u) | (v << 24);
}
inline uint64_t bswap64(uint64_t v) {
    return ((uint64_t)bswap32((uint32_t)v) << 32) | bswap32((uint32_t)(v >> 32));
}

inline std::vector<double> read_vec(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    int32_t head[2];
    f.read(reinterpret_cast<char *>(head), 8);
    const int32_t cid = (int32_t)bswap32((uint32_t)head[0]);
    const int32_t n = (int32_t)bswap32((uint32_t)head[1]);
    if (!f || cid != VEC_FILE_CLASSID || n < 0) throw std::runtime_error(path + ": not a PETSc Vec");
    std::vector<uint64_t> raw((size_t)n);
    f.read(reinterpret_cast<char *>(raw.data()), 8 * (std::streamsize)n);
    if (!f) throw std::runtime_error(path + ": truncated");
    std::vector<double> out((size_t)n);
    for (int32_t i = 0; i < n; i++) {
        const uint64_t b = bswap64(raw[(size_t)i]);
        std::memcpy(&out[(size_t)i], &b, 8);
    }
    return out;
}

inline void write_vec(const std::string &path, const std::vector<double> &v) {
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + path);
    const int32_t head[2] = {(int32_t)bswap32((uint32_t)VEC_FILE_CLASSID),
                             (int32_t)bswap32((uint32_t)v.size())};
    f.write(reinterpret_cast<const char *>(head), 8);
    for (double d : v) {
        uint64_t b;
        std::memcpy(&b, &d, 8);
        b = bswap64(b);
        f.write(reinterpret_cast<const char *>(&b), 8);
    }
}
}  // namespace petscio

}  // namespace pascinference
