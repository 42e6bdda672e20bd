#include "async_export.hpp"

#include <algorithm>
#include <fstream>
#include <vector>

#include "exporter.hpp"
#include "logger.hpp"

namespace {

// Runs body(begin, end, out) over `threads` contiguous ranges of [0, n) and
// joins the texts in range order.
template <typename F>
std::string parallel_text(size_t n, int threads, F&& body) {
    const size_t t = (size_t) std::max(1, threads);
    const size_t per = (n + t - 1) / t;
    std::vector<std::string> parts(t);
    std::vector<std::thread> pool;
    for (size_t k = 0; k < t; k++) {
        const size_t b = std::min(n, k * per), e = std::min(n, b + per);
        if (b >= e) continue;
        pool.emplace_back([&, k, b, e] { body(b, e, parts[k]); });
    }
    for (auto& th : pool) th.join();
    size_t total = 0;
    for (auto& s : parts) total += s.size();
    std::string out;
    out.reserve(total);
    for (auto& s : parts) out += s;
    return out;
}

}  // namespace

// src/exporter.cpp:10-55 (ghosts skipped in the rows, counted in the header)
std::string csvText(const ExportView& v, double time, int threads) {
    std::string head;
    head += "Time (s)," + fmt_fixed6(time) + "\n";
    head += "Number of Particles," + std::to_string(v.n) + "\n";
    head += "ID,Type,Position.x (m),Position.y (m),Position.z (m),Velocity.x (m/s),Velocity.y (m/s),"
            "Velocity.z (m/s),Pressure (Pa),Number Density Ratio\n";
    return head + parallel_text(v.n, threads, [&](size_t b, size_t e, std::string& out) {
               out.reserve(96 * (e - b));
               for (size_t i = b; i < e; i++) {
                   if (v.type[i] == -1) continue;
                   out += std::to_string(i);
                   out += ',';
                   out += std::to_string((int) v.type[i]);
                   const double vals[8] = {v.pos[3 * i], v.pos[3 * i + 1], v.pos[3 * i + 2], v.vel[3 * i],
                                           v.vel[3 * i + 1], v.vel[3 * i + 2], v.p[i], v.ndr[i]};
                   for (double x : vals) {
                       out += ',';
                       out += fmt_fixed6(x);
                   }
                   out += '\n';
               }
           });
}

// src/exporter.cpp:57-181 (every particle, ghosts included; z velocity 0)
std::string vtuText(const ExportView& v, double time, int threads) {
    const size_t n = v.n;
    std::string b;
    auto begin = [&](const char* comps, const char* type, const char* name) {
        b += std::string("<DataArray NumberOfComponents='") + comps + "' type='" + type + "' Name='" + name +
             "' format='ascii'>\n";
    };
    auto section = [&](auto&& line) {
        b += parallel_text(n, threads, [&](size_t s, size_t e, std::string& out) {
            for (size_t i = s; i < e; i++) line(i, out);
        });
    };
    b += "<?xml version='1.0' encoding='UTF-8'?>\n";
    b += "<VTKFile byte_order='LittleEndian' version='0.1' type = 'UnstructuredGrid'>\n";
    b += "<UnstructuredGrid>\n";
    b += "<FieldData>\n";
    b += "<DataArray type='Float64' Name='Time' NumberOfComponents='1' NumberOfTuples='1' format='ascii'>\n";
    b += fmt_general(time) + "\n";
    b += "</DataArray>\n";
    b += "</FieldData>\n";
    b += "<Piece NumberOfCells='" + std::to_string(n) + "' NumberOfPoints='" + std::to_string(n) + "'>\n";
    b += "<Points>\n";
    b += "<DataArray NumberOfComponents='3' type='Float64' Name='position' format='ascii'>\n";
    section([&](size_t i, std::string& out) {
        out += fmt_general(v.pos[3 * i]) + " " + fmt_general(v.pos[3 * i + 1]) + " " + fmt_general(v.pos[3 * i + 2]) +
               "\n";
    });
    b += "</DataArray>\n";
    b += "</Points>\n";
    b += "<PointData>\n";
    begin("1", "Int32", "Type");
    section([&](size_t i, std::string& out) { out += std::to_string((int) v.type[i]) + "\n"; });
    b += "</DataArray>\n";
    begin("3", "Float64", "Velocity");
    section([&](size_t i, std::string& out) {
        out += fmt_general(v.vel[3 * i]) + " " + fmt_general(v.vel[3 * i + 1]) + " " + fmt_general(0.0) + "\n";
    });
    b += "</DataArray>\n";
    begin("1", "Float64", "Pressure");
    section([&](size_t i, std::string& out) { out += fmt_general(v.p[i]) + "\n"; });
    b += "</DataArray>\n";
    begin("1", "Float64", "Number Density Ratio");
    section([&](size_t i, std::string& out) { out += fmt_general(v.ndr[i]) + "\n"; });
    b += "</DataArray>\n";
    begin("1", "Int32", "Boundary Condition");
    section([&](size_t i, std::string& out) { out += std::to_string((int) v.bc[i]) + "\n"; });
    b += "</DataArray>\n";
    b += "</PointData>\n";
    b += "<Cells>\n";
    b += "<DataArray type='Int32' Name='connectivity' format='ascii'>\n";
    section([&](size_t i, std::string& out) { out += std::to_string(i) + "\n"; });
    b += "</DataArray>\n";
    b += "<DataArray type='Int32' Name='offsets' format='ascii'>\n";
    section([&](size_t i, std::string& out) { out += std::to_string(i + 1) + "\n"; });
    b += "</DataArray>\n";
    b += "<DataArray type='UInt8' Name='types' format='ascii'>\n";
    section([&](size_t, std::string& out) { out += "1\n"; });
    b += "</DataArray>\n";
    b += "</Cells>\n";
    b += "</Piece>\n";
    b += "</UnstructuredGrid>\n";
    b += "</VTKFile>\n";
    return b;
}

AsyncExporter::AsyncExporter(int formatThreads) {
    int t = formatThreads;
    if (t <= 0) t = (int) std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    threads_ = t;
    worker_ = std::thread([this] { loop(); });
}

AsyncExporter::~AsyncExporter() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    worker_.join();
}

void AsyncExporter::submit(int slot, std::function<ExportView()> wait, std::filesystem::path csv,
                           std::filesystem::path vtu, double time) {
    {
        std::lock_guard<std::mutex> lk(mu_);
        q_.push_back(Job{slot, std::move(wait), std::move(csv), std::move(vtu), time});
        busy_[slot & 1]++;
    }
    cv_.notify_all();
}

void AsyncExporter::waitSlot(int slot) {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return busy_[slot & 1] == 0; });
}

void AsyncExporter::drain() {
    waitSlot(0);
    waitSlot(1);
}

void AsyncExporter::loop() {
    for (;;) {
        Job job;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_ || !q_.empty(); });
            if (q_.empty()) return;  // stop requested and nothing left
            job = std::move(q_.front());
            q_.pop_front();
        }
        try {
            const ExportView v = job.wait();
            {
                std::ofstream f(job.csv, std::ios::binary);
                if (!f.is_open()) mpslog::error("Could not open target csv file: " + job.csv.string());
                f << csvText(v, job.time, threads_);
            }
            {
                std::ofstream f(job.vtu, std::ios::binary);
                if (!f.is_open()) mpslog::error("Could not open target vtu file: " + job.vtu.string());
                f << vtuText(v, job.time, threads_);
            }
        } catch (const std::exception& e) {
            mpslog::error(std::string("export failed: ") + e.what());
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            busy_[job.slot & 1]--;
        }
        cv_.notify_all();
    }
}
