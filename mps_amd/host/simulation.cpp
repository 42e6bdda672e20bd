#include "simulation.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "exporter.hpp"
#include "logger.hpp"

using std::chrono::duration_cast;
using std::chrono::milliseconds;
using std::chrono::seconds;
using std::chrono::system_clock;
namespace fs = std::filesystem;

namespace {

// std::format("{}", double): the shortest text that reads back as the value
std::string fmt_shortest(double v) {
    char b[64];
    for (int p = 1; p <= 17; p++) {
        std::snprintf(b, sizeof b, "%.*g", p, v);
        if (std::strtod(b, nullptr) == v) break;
    }
    return b;
}

std::string fmt_fixed(double v, int digits) {
    char b[64];
    std::snprintf(b, sizeof b, "%.*f", digits, v);
    return b;
}

// std::format("{:%T}", seconds): HH:MM:SS
std::string fmt_hms(long long s) {
    char b[64];
    const bool neg = s < 0;
    if (neg) s = -s;
    std::snprintf(b, sizeof b, "%s%02lld:%02lld:%02lld", neg ? "-" : "", s / 3600, (s / 60) % 60, s % 60);
    return b;
}

// std::format("{:%S}", milliseconds): SS.mmm
std::string fmt_sec_ms(long long ms) {
    char b[64];
    std::snprintf(b, sizeof b, "%02lld.%03lld", (ms / 1000) % 60, ms % 1000);
    return b;
}

}  // namespace

// src/simulation.cpp:29-64
void Simulation::run(const fs::path& inputYamlPath) {
    createResultDirectory(inputYamlPath);
    prepareLogFile();

    mps = MPS(inputYamlPath);
    if (haveOpt) mps.setSolver(opt);
    time = mps.loadInitialState();
    // asynchronous export unless MPS_AMD_SYNC_EXPORT=1 (same files either way)
    const char* se = std::getenv("MPS_AMD_SYNC_EXPORT");
    if (!(se && std::atoi(se))) {
        asyncExporter = std::make_unique<AsyncExporter>();
        mps.setSyncOnExport(false);
    }
    if (asyncExporter)
        exportAsync();
    else
        exportParticles(mps.particles);

    mpslog::info("START SIMULATION");
    simulationStartTime = system_clock::now();
    while (time + mps.settings.dt <= mps.settings.finishTime) {
        auto timeStepStartTime = system_clock::now();
        int initialDebugLogCount = mps.debugLogCount;

        time += mps.settings.dt;

        double courantNumber = mps.stepForward(isTimeToExport());

        if (isTimeToExport()) {
            if (asyncExporter)
                exportAsync();
            else
                exportParticles(mps.particles);
        }

        auto timeStepEndTime = system_clock::now();
        timeStepReport(timeStepStartTime, timeStepEndTime, courantNumber, mps.debugLogCount > initialDebugLogCount);

        timeStep++;
    }
    if (asyncExporter) asyncExporter->drain();
    simulationEndTime = system_clock::now();

    endSimulation();
}

// src/simulation.cpp:66-77
void Simulation::createResultDirectory(const fs::path& inputYamlPath) {
    resultDirectory = inputYamlPath.parent_path() / "result";
    if (fs::exists(resultDirectory)) fs::remove_all(resultDirectory);
    fs::create_directory(resultDirectory);
    fs::create_directory(resultDirectory / "csv");
    fs::create_directory(resultDirectory / "vtu");
    fs::copy(inputYamlPath, resultDirectory / inputYamlPath.filename());
}

// src/simulation.cpp:79-133
void Simulation::prepareLogFile() {
    mpslog::set_file(resultDirectory / "execution.log");
    fs::path path = resultDirectory / "time_step_report.csv";
    timeStepReportFile.open(path);
    if (!timeStepReportFile.is_open()) mpslog::error("Could not open the log file: " + path.string());
    timeStepReportFile << "Time Step,Dt (s),Time Before (s),Time After (s),Finish Time (s),Elapsed Time (h:m:s),"
                          "Elapsed Time (s),Remaining Time (h:m:s),Average Time per Time Step (s),"
                          "Time for This Time Step (s),Number of Output Files,Courant Number\n";
}

// src/simulation.cpp:135-151
void Simulation::endSimulation() {
    auto total = duration_cast<seconds>(simulationEndTime - simulationStartTime);
    mpslog::info("END SIMULATION");
    if (mps.debugLogCount > 0)
        mpslog::info("There were " + std::to_string(mps.debugLogCount) + " debug-level logs. Check 'execution.log'.");
    else
        mpslog::info("There were no debug-level logs.");
    mpslog::info("Total Simulation Time = " + fmt_hms(total.count()));
    timeStepReportFile.close();
    mpslog::close_file();
}

// src/simulation.cpp:153-222
void Simulation::timeStepReport(const system_clock::time_point& timeStepStartTime,
                                const system_clock::time_point& timeStepEndTime, const double& courantNumber,
                                const bool& isDebugLogAdded) {
    auto elapsed = duration_cast<seconds>(timeStepEndTime - simulationStartTime);
    double average = 0;
    if (timeStep != 0)
        average = duration_cast<milliseconds>(timeStepEndTime - simulationStartTime).count() / (double) (1000 * timeStep);
    seconds remain{int(((mps.settings.finishTime - time) / time) * average * timeStep)};
    auto last = duration_cast<milliseconds>(timeStepEndTime - timeStepStartTime);

    const std::string before = fmt_fixed(time - mps.settings.dt, 3);
    const std::string after = fmt_fixed(time, 3);
    const std::string fElapsed = fmt_hms(elapsed.count());
    const std::string fAverage = fmt_fixed(average, 3);
    const std::string fRemain = fmt_hms(remain.count());
    const std::string fLast = fmt_sec_ms(last.count());
    const std::string fCourant = fmt_fixed(courantNumber, 2);

    std::cout << timeStep << ": dt=" << fmt_shortest(mps.settings.dt) << "s  t=" << before << "s->" << after
              << "s  fin=" << fmt_shortest(mps.settings.finishTime) << "s  elapsed=" << fElapsed
              << "  remain=" << fRemain << "  ave=" << fAverage << "s/step  last=" << fLast
              << "s/step  out=" << outFileNum << "files  Courant=" << fCourant << std::endl;

    timeStepReportFile << timeStep << "," << fmt_fixed6(mps.settings.dt) << "," << before << "," << after << ","
                       << fmt_fixed6(mps.settings.finishTime) << "," << fElapsed << "," << elapsed.count() << ","
                       << fRemain << "," << fAverage << "," << fLast << "," << outFileNum << "," << fCourant << "\n";

    if (isDebugLogAdded)
        mpslog::debug("Time Step:" + std::to_string(timeStep) + " Time:" + before + "s->" + after + "s");
}

// src/simulation.cpp:224-229
bool Simulation::isTimeToExport() {
    double outputInterval = mps.settings.outputInterval;
    double remainder = std::fmod(time, outputInterval);
    remainder = std::min(remainder, outputInterval - remainder);
    return remainder < mps.settings.dt / 10.0;
}

// src/simulation.cpp:231-245
void Simulation::exportParticles(const std::vector<Particle>& particles) {
    Exporter exporter;
    char name[64];
    std::snprintf(name, sizeof name, "output_%04d", outFileNum);
    exporter.toCsv(resultDirectory / "csv" / (std::string(name) + ".csv"), particles, time);
    exporter.toVtu(resultDirectory / "vtu" / (std::string(name) + ".vtu"), particles, time);
    outFileNum++;
}

// exportParticles, from an asynchronous snapshot: slot k of the library holds
// one snapshot until its files are written, two alternate
void Simulation::exportAsync() {
    const int slot = outFileNum & 1;
    asyncExporter->waitSlot(slot);
    mps.snapshotBegin(slot);
    char name[64];
    std::snprintf(name, sizeof name, "output_%04d", outFileNum);
    MPS* m = &mps;
    asyncExporter->submit(
        slot, [m, slot] { return m->snapshotEnd(slot); }, resultDirectory / "csv" / (std::string(name) + ".csv"),
        resultDirectory / "vtu" / (std::string(name) + ".vtu"), time);
    outFileNum++;
}
