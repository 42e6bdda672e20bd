// mps.hpp — the reference's MPS class surface (src/mps.hpp:22-71) over the
// GPU time step of libmps_amd.so.  The particle state stays in HBM; the
// `particles` vector is refreshed on export steps (and after
// loadInitialState), which is when the reference's callers read it
// (src/simulation.cpp:47-49, 86-88).
#pragma once

#include <filesystem>
#include <vector>

#include "../../include/mps_amd.h"
#include "async_export.hpp"
#include "bucket.hpp"
#include "particle.hpp"
#include "settings.hpp"

class MPS {
public:
    Settings settings;
    std::vector<Particle> particles;
    int debugLogCount = 0;

    MPS() = default;
    MPS(const std::filesystem::path& inputYamlPath);
    MPS(const MPS&) = delete;
    MPS& operator=(const MPS&) = delete;
    MPS(MPS&& o) noexcept;
    MPS& operator=(MPS&& o) noexcept;
    ~MPS();

    double loadInitialState();  // Initialize particles and return initial time
    double stepForward(const bool isTimeToExport);

    // ---- beyond the reference ----
    // Pressure-solver options (default: FP64, machine-epsilon tolerance, cold
    // start, as the reference's Eigen solver).  May be called before or after
    // loadInitialState.
    void setSolver(const mps_solver_options& o);
    mps_solver_stats solverStats() const;
    mps_handle* handle() const { return h; }
    // Copies the device state into `particles`.
    void syncParticles();
    // false: stepForward(true) leaves `particles` as they are (the caller
    // exports from snapshots instead).
    void setSyncOnExport(bool on) { syncOnExport = on; }
    // Asynchronous snapshot of the export fields (mps_snapshot_begin/_end);
    // snapshotEnd may run on another thread.
    void snapshotBegin(int slot);
    ExportView snapshotEnd(int slot) const;

private:
    mps_handle* h = nullptr;
    mps_solver_options opt{};
    bool haveOpt = false;
    double lastCourant = 0.0;
    bool syncOnExport = true;

    double importInitialCondition();
    void reportEvents(double courant);
    void check(int rc, const char* where) const;
};
