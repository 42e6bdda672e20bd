// async_export.hpp — the reference's CSV/VTU output (src/exporter.cpp:10-181)
// written from an asynchronous device snapshot: formatting runs on several
// threads over particle ranges whose texts are joined in order, and the
// files are written on a background thread while the time loop steps on.
// The bytes are those of Exporter::toCsv / Exporter::toVtu for the same
// values.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <filesystem>
#include <functional>
#include <mutex>
#include <string>
#include <thread>

// Export fields in id order (structure of arrays), as mps_snapshot_end
// returns them.
struct ExportView {
    size_t n = 0;
    const int8_t* type = nullptr;
    const double* pos = nullptr;  // n x 3
    const double* vel = nullptr;  // n x 3
    const double* p = nullptr;
    const double* ndr = nullptr;
    const int8_t* bc = nullptr;
};

// Texts of the two writers, built by `threads` threads (>= 1).
std::string csvText(const ExportView& v, double time, int threads);
std::string vtuText(const ExportView& v, double time, int threads);

class AsyncExporter {
public:
    explicit AsyncExporter(int formatThreads = 0);  // 0: hardware concurrency, at most 16
    ~AsyncExporter();                               // finishes every queued write
    AsyncExporter(const AsyncExporter&) = delete;
    AsyncExporter& operator=(const AsyncExporter&) = delete;

    // Queues one output pair.  `wait` blocks until the snapshot is on the host
    // and returns it; it runs on the writer thread.  Jobs run in order.
    void submit(int slot, std::function<ExportView()> wait, std::filesystem::path csv, std::filesystem::path vtu,
                double time);
    // Blocks until no queued or running job uses `slot`.
    void waitSlot(int slot);
    void drain();

private:
    struct Job {
        int slot;
        std::function<ExportView()> wait;
        std::filesystem::path csv, vtu;
        double time;
    };
    int threads_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<Job> q_;
    int busy_[2] = {0, 0};
    bool stop_ = false;
    std::thread worker_;
    void loop();
};
