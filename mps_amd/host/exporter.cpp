#include "exporter.hpp"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>

#include "logger.hpp"

std::string fmt_fixed6(double v) {
    char b[512];
    std::snprintf(b, sizeof b, "%f", v);
    return b;
}

std::string fmt_general(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%g", v);
    return b;
}

namespace {

// A CSV field as the reference's CSV writer emits it: quoted only when it
// holds the delimiter, a quote or a line break.
std::string csv_field(const std::string& s) {
    if (s.find_first_of(",\"\n\r") == std::string::npos) return s;
    std::string q = "\"";
    for (char c : s) {
        if (c == '"') q += '"';
        q += c;
    }
    return q + "\"";
}

std::vector<std::string> csv_split(const std::string& line) {
    std::vector<std::string> out;
    std::string cur;
    bool inq = false;
    for (size_t k = 0; k < line.size(); k++) {
        const char c = line[k];
        if (inq) {
            if (c == '"') {
                if (k + 1 < line.size() && line[k + 1] == '"') {
                    cur += '"';
                    k++;
                } else {
                    inq = false;
                }
            } else {
                cur += c;
            }
        } else if (c == '"') {
            inq = true;
        } else if (c == ',') {
            out.push_back(cur);
            cur.clear();
        } else if (c != '\r') {
            cur += c;
        }
    }
    out.push_back(cur);
    return out;
}

}  // namespace

void Exporter::toCsv(const std::filesystem::path& outFilePath, const std::vector<Particle>& particles,
                     const double& time) {
    std::ofstream outFile(outFilePath, std::ios::binary);
    if (!outFile.is_open()) mpslog::error("Could not open target csv file: " + outFilePath.string());
    std::string buf;
    buf.reserve(96 * particles.size() + 512);
    buf += csv_field("Time (s)") + "," + fmt_fixed6(time) + "\n";
    buf += csv_field("Number of Particles") + "," + std::to_string(particles.size()) + "\n";
    buf += "ID,Type,Position.x (m),Position.y (m),Position.z (m),Velocity.x (m/s),Velocity.y (m/s),"
           "Velocity.z (m/s),Pressure (Pa),Number Density Ratio\n";
    for (const auto& pi : particles) {
        if (pi.type == ParticleType::Ghost) continue;  // ghosts are skipped, but counted above
        buf += std::to_string(pi.id);
        buf += ',';
        buf += std::to_string(static_cast<int>(pi.type));
        const double vals[8] = {pi.position.x(), pi.position.y(), pi.position.z(), pi.velocity.x(),
                                pi.velocity.y(), pi.velocity.z(), pi.pressure,     pi.numberDensityRatio};
        for (double v : vals) {
            buf += ',';
            buf += fmt_fixed6(v);
        }
        buf += '\n';
    }
    outFile << buf;
}

void Exporter::toVtu(const std::filesystem::path& outFilePath, const std::vector<Particle>& particles,
                     const double& time) {
    std::ofstream outFile(outFilePath, std::ios::binary);
    if (!outFile.is_open()) mpslog::error("Could not open target vtu file: " + outFilePath.string());
    const size_t n = particles.size();
    std::string b;
    b.reserve(160 * n + 2048);
    auto begin = [&](const char* comps, const char* type, const char* name) {
        b += std::string("<DataArray NumberOfComponents='") + comps + "' type='" + type + "' Name='" + name +
             "' format='ascii'>\n";
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
    for (const auto& pi : particles)
        b += fmt_general(pi.position.x()) + " " + fmt_general(pi.position.y()) + " " + fmt_general(pi.position.z()) +
             "\n";
    b += "</DataArray>\n";
    b += "</Points>\n";
    b += "<PointData>\n";
    begin("1", "Int32", "Type");
    for (const auto& pi : particles) b += std::to_string(static_cast<int>(pi.type)) + "\n";
    b += "</DataArray>\n";
    begin("3", "Float64", "Velocity");
    for (const auto& pi : particles)
        b += fmt_general(pi.velocity.x()) + " " + fmt_general(pi.velocity.y()) + " " + fmt_general(0.0) + "\n";
    b += "</DataArray>\n";
    begin("1", "Float64", "Pressure");
    for (const auto& pi : particles) b += fmt_general(pi.pressure) + "\n";
    b += "</DataArray>\n";
    begin("1", "Float64", "Number Density Ratio");
    for (const auto& pi : particles) b += fmt_general(pi.numberDensityRatio) + "\n";
    b += "</DataArray>\n";
    begin("1", "Int32", "Boundary Condition");
    for (const auto& pi : particles) b += std::to_string(static_cast<int>(pi.boundaryCondition)) + "\n";
    b += "</DataArray>\n";
    b += "</PointData>\n";
    b += "<Cells>\n";
    b += "<DataArray type='Int32' Name='connectivity' format='ascii'>\n";
    for (size_t i = 0; i < n; i++) b += std::to_string(i) + "\n";
    b += "</DataArray>\n";
    b += "<DataArray type='Int32' Name='offsets' format='ascii'>\n";
    for (size_t i = 0; i < n; i++) b += std::to_string(i + 1) + "\n";
    b += "</DataArray>\n";
    b += "<DataArray type='UInt8' Name='types' format='ascii'>\n";
    for (size_t i = 0; i < n; i++) b += "1\n";
    b += "</DataArray>\n";
    b += "</Cells>\n";
    b += "</Piece>\n";
    b += "</UnstructuredGrid>\n";
    b += "</VTKFile>\n";
    outFile << b;
}

bool importCsv(const std::filesystem::path& path, double density, std::vector<Particle>& particles, double& time) {
    if (!std::filesystem::exists(path)) {
        mpslog::error("Input file does not exist in the specified path: " + path.string());
        return false;
    }
    std::ifstream in(path);
    std::string line;
    std::vector<std::string> rows;
    for (int k = 0; k < 3 && std::getline(in, line); k++) rows.push_back(line);
    if (rows.size() < 3) {
        mpslog::error("Input file is too short: " + path.string());
        return false;
    }
    const auto meta = csv_split(rows[0]);
    if (meta.size() < 2) {
        mpslog::error("Input file has no time in its first row: " + path.string());
        return false;
    }
    time = std::strtod(meta[1].c_str(), nullptr);
    const auto header = csv_split(rows[2]);
    std::map<std::string, size_t> col;
    for (size_t k = 0; k < header.size(); k++) col[header[k]] = k;
    const char* need[] = {"ID", "Type", "Position.x (m)", "Position.y (m)", "Position.z (m)",
                          "Velocity.x (m/s)", "Velocity.y (m/s)", "Velocity.z (m/s)"};
    size_t idx[8];
    for (int k = 0; k < 8; k++) {
        auto it = col.find(need[k]);
        if (it == col.end()) {
            mpslog::error(std::string("Input file lacks the column ") + need[k]);
            return false;
        }
        idx[k] = it->second;
    }
    particles.clear();
    while (std::getline(in, line)) {
        if (line.empty() || line == "\r") continue;
        const auto f = csv_split(line);
        if (f.size() < header.size()) continue;
        const int id = std::atoi(f[idx[0]].c_str());
        const int type = std::atoi(f[idx[1]].c_str());
        double v[6];
        for (int k = 0; k < 6; k++) v[k] = std::strtod(f[idx[2 + k]].c_str(), nullptr);
        particles.emplace_back(id, static_cast<ParticleType>(type), Vec3(v[0], v[1], v[2]), Vec3(v[3], v[4], v[5]),
                               density);
    }
    return true;
}
