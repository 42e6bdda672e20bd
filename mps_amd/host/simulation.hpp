// simulation.hpp — the reference's driver (src/simulation.hpp:11-39): result
// directory, logging, time loop, per-step report and export cadence.
#pragma once

#include <chrono>
#include <filesystem>
#include <fstream>
#include <memory>
#include <vector>

#include "mps.hpp"
#include "particle.hpp"

class Simulation {
public:
    void run(const std::filesystem::path& inputYamlPath);

    // beyond the reference: pressure-solver options applied to the MPS object
    void setSolver(const mps_solver_options& o) {
        opt = o;
        haveOpt = true;
    }

private:
    MPS mps;
    mps_solver_options opt{};
    bool haveOpt = false;

    int timeStep = 0;
    double time = 0.0;
    std::chrono::system_clock::time_point simulationStartTime, simulationEndTime;

    int outFileNum = 0;
    std::ofstream timeStepReportFile;
    std::filesystem::path resultDirectory;

    void createResultDirectory(const std::filesystem::path& inputYamlPath);
    void prepareLogFile();
    void timeStepReport(const std::chrono::system_clock::time_point& timeStepStartTime,
                        const std::chrono::system_clock::time_point& timeStepEndTime, const double& courantNumber,
                        const bool& isDebugLogAdded);
    bool isTimeToExport();
    void exportParticles(const std::vector<Particle>& particles);
    // beyond the reference: the same files from an asynchronous device
    // snapshot, formatted and written on a background thread
    std::unique_ptr<AsyncExporter> asyncExporter;
    void exportAsync();
    void endSimulation();
};
