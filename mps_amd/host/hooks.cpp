// hooks.cpp — extern "C" entry points of libmps_host.so that exercise the
// host-side formats without a GPU (used by the CPU test suite).
#include <cstring>

#include "exporter.hpp"
#include "settings.hpp"

extern "C" {

// Settings::load of a YAML file into the C-ABI POD; returns 0 on success.
int mps_host_load_settings(const char* yaml, mps_settings* out, char* input_csv, int cap) {
    Settings s;
    if (!s.load(yaml)) return 1;
    *out = s.toC();
    if (input_csv && cap > 0) {
        std::strncpy(input_csv, s.inputCsvPath.string().c_str(), (size_t) cap - 1);
        input_csv[cap - 1] = '\0';
    }
    return 0;
}

// Reads a particle CSV and writes it back through the CSV and VTU writers.
int mps_host_convert_csv(const char* in_csv, double density, const char* out_csv, const char* out_vtu) {
    std::vector<Particle> ps;
    double t = 0.0;
    if (!importCsv(in_csv, density, ps, t)) return 1;
    Exporter ex;
    if (out_csv) ex.toCsv(out_csv, ps, t);
    if (out_vtu) ex.toVtu(out_vtu, ps, t);
    return 0;
}

}

#include "bucket.hpp"

extern "C" {

// Bucket(reMax, domain, N).storeParticles(particles of a CSV): dims4 gets
// num/numX/numY/numZ; first/last (num) and next (N) are filled when not NULL.
// The lists come from a second call: the first stores the particles shifted
// by a third of a bucket along x, so the second reuses the device grid with
// new positions.
int mps_host_bucket_from_csv(const char* csv, double re_max, const double* dmin, const double* dmax, int* dims4,
                             int* first, int* last, int* next) {
    std::vector<Particle> ps;
    double t = 0.0;
    if (!importCsv(csv, 1000.0, ps, t)) return 1;
    Domain d(dmin[0], dmax[0], dmin[1], dmax[1], dmin[2], dmax[2]);
    Bucket b(re_max, d, (int) ps.size());
    dims4[0] = b.num;
    dims4[1] = b.numX;
    dims4[2] = b.numY;
    dims4[3] = b.numZ;
    if (!first) return 0;
    try {
        std::vector<Particle> shifted = ps;
        for (auto& p : shifted) p.position = Vec3(p.position.x() + re_max / 3, p.position.y(), p.position.z());
        b.storeParticles(shifted);
        b.storeParticles(ps);
    } catch (const std::exception&) {
        return 2;
    }
    std::copy(b.first.begin(), b.first.end(), first);
    std::copy(b.last.begin(), b.last.end(), last);
    std::copy(b.next.begin(), b.next.end(), next);
    return 0;
}

}

#include <fstream>

#include "async_export.hpp"

extern "C" {

// Writes one particle table (id order; type/bc int8, pos/vel n x 3) through
// the synchronous writers (Exporter on a Particle vector) and through the
// threaded snapshot writers (csvText / vtuText with `threads` threads):
// out_prefix + "sync.csv", "sync.vtu", "async.csv", "async.vtu".
int mps_host_export_both(int n, const int8_t* type, const double* pos, const double* vel, const double* p,
                         const double* ndr, const int8_t* bc, double time, const char* out_prefix, int threads) {
    const std::string pre(out_prefix);
    std::vector<Particle> ps;
    ps.reserve((size_t) n);
    for (int i = 0; i < n; i++) {
        ps.emplace_back(i, static_cast<ParticleType>(type[i]), Vec3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]),
                        Vec3(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]), 1000.0);
        ps.back().pressure = p[i];
        ps.back().numberDensityRatio = ndr[i];
        ps.back().boundaryCondition = static_cast<BoundaryCondition>(bc[i]);
    }
    Exporter ex;
    ex.toCsv(pre + "sync.csv", ps, time);
    ex.toVtu(pre + "sync.vtu", ps, time);
    ExportView v;
    v.n = (size_t) n;
    v.type = type;
    v.pos = pos;
    v.vel = vel;
    v.p = p;
    v.ndr = ndr;
    v.bc = bc;
    std::ofstream(pre + "async.csv", std::ios::binary) << csvText(v, time, threads);
    std::ofstream(pre + "async.vtu", std::ios::binary) << vtuText(v, time, threads);
    return 0;
}

}
