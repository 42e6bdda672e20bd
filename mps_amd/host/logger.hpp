// logger.hpp — the logging the reference does through spdlog
// (src/simulation.cpp:79-110): a console sink at info and, once a file is
// attached, a file sink at debug, both with the pattern
// "[%Y-%m-%d %H:%M:%S] [%l] %v".
#pragma once

#include <filesystem>
#include <string>

namespace mpslog {

void set_file(const std::filesystem::path& path);  // attaches (truncates) the debug file sink
void close_file();
void debug(const std::string& msg);
void info(const std::string& msg);
void error(const std::string& msg);

}  // namespace mpslog
