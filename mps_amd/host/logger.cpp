#include "logger.hpp"

#include <ctime>
#include <fstream>
#include <iostream>
#include <mutex>

namespace mpslog {

static std::ofstream g_file;
static std::mutex g_mu;

static std::string stamp() {
    char buf[32];
    std::time_t t = std::time(nullptr);
    std::tm tmv;
    localtime_r(&t, &tmv);
    std::strftime(buf, sizeof buf, "%Y-%m-%d %H:%M:%S", &tmv);
    return std::string("[") + buf + "]";
}

static void emit(int level, const char* name, const std::string& msg) {
    std::lock_guard<std::mutex> lk(g_mu);
    const std::string line = stamp() + " [" + name + "] " + msg;
    if (level >= 1) std::cout << line << std::endl;  // console: info and above
    if (g_file.is_open()) {
        g_file << line << '\n';
        if (level >= 2) g_file.flush();  // flush on error
    }
}

void set_file(const std::filesystem::path& path) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_file.is_open()) g_file.close();
    g_file.open(path, std::ios::trunc);
}

void close_file() {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_file.is_open()) g_file.close();
}

void debug(const std::string& msg) { emit(0, "debug", msg); }
void info(const std::string& msg) { emit(1, "info", msg); }
void error(const std::string& msg) { emit(2, "error", msg); }

}  // namespace mpslog
