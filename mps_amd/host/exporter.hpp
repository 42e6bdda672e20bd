// exporter.hpp — CSV and ASCII VTU writers with the reference's byte layout
// (src/exporter.cpp:10-181), and the CSV reader of MPS::importInitialCondition
// (src/mps.cpp:38-94).
#pragma once

#include <filesystem>
#include <string>
#include <vector>

#include "particle.hpp"

class Exporter {
public:
    void toCsv(const std::filesystem::path& outputFilePath, const std::vector<Particle>& particles,
               const double& time);
    void toVtu(const std::filesystem::path& outputFilePath, const std::vector<Particle>& particles,
               const double& time);
};

// Reads an input/output CSV: time from row 1, header at row 3, particles
// after it (density set to `density` for every particle, src/mps.cpp:84-90).
// Returns false (and logs) when the file cannot be read.
bool importCsv(const std::filesystem::path& path, double density, std::vector<Particle>& particles, double& time);

// Number formats of the two writers: std::to_string (printf "%f") as the
// reference's CSV writer prints doubles under clang, and an ostream's default
// (printf "%g", 6 significant digits) in the VTU writer.
std::string fmt_fixed6(double v);
std::string fmt_general(double v);
