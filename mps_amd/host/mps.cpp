#include "mps.hpp"

#include <algorithm>
#include <stdexcept>
#include <string>

#include "exporter.hpp"
#include "logger.hpp"

MPS::MPS(const std::filesystem::path& inputYamlPath) {
    // src/mps.cpp:18-36; the reference values are computed by the library
    settings.load(inputYamlPath);
}

MPS::MPS(MPS&& o) noexcept { *this = std::move(o); }

MPS& MPS::operator=(MPS&& o) noexcept {
    if (this != &o) {
        if (h) mps_destroy(h);
        settings = o.settings;
        particles = std::move(o.particles);
        debugLogCount = o.debugLogCount;
        h = o.h;
        opt = o.opt;
        haveOpt = o.haveOpt;
        lastCourant = o.lastCourant;
        o.h = nullptr;
    }
    return *this;
}

MPS::~MPS() {
    if (h) mps_destroy(h);
}

void MPS::check(int rc, const char* where) const {
    if (rc) {
        const std::string m = std::string(where) + ": " + mps_last_error(h);
        mpslog::error(m);
        throw std::runtime_error(m);
    }
}

// src/mps.cpp:38-94
double MPS::importInitialCondition() {
    double t = 0.0;
    importCsv(settings.inputCsvPath, settings.density, particles, t);
    return t;
}

void MPS::setSolver(const mps_solver_options& o) {
    opt = o;
    haveOpt = true;
    if (h) check(mps_set_solver(h, &opt), "mps_set_solver");
}

mps_solver_stats MPS::solverStats() const {
    mps_solver_stats s{};
    if (h) mps_get_solver_stats(h, &s);
    return s;
}

// src/mps.cpp:96-109
double MPS::loadInitialState() {
    const double initialTime = importInitialCondition();
    const size_t n = particles.size();
    std::vector<int32_t> id(n);
    std::vector<int8_t> type(n);
    std::vector<double> pos(3 * n), vel(3 * n);
    for (size_t i = 0; i < n; i++) {
        id[i] = particles[i].id;
        type[i] = (int8_t) static_cast<int>(particles[i].type);
        for (int d = 0; d < 3; d++) {
            pos[3 * i + d] = particles[i].position[d];
            vel[3 * i + d] = particles[i].velocity[d];
        }
    }
    const mps_settings cs = settings.toC();
    if (h) mps_destroy(h);
    h = nullptr;
    const int rc = mps_create(&cs, (int32_t) n, id.data(), type.data(), pos.data(), vel.data(), &h);
    if (rc) {
        const std::string m = std::string("mps_create: ") + mps_last_error(nullptr);
        mpslog::error(m);
        throw std::runtime_error(m);
    }
    if (haveOpt) check(mps_set_solver(h, &opt), "mps_set_solver");
    check(mps_load_initial_state(h), "mps_load_initial_state");
    syncParticles();
    return initialTime;
}

void MPS::syncParticles() {
    const size_t n = particles.size();
    std::vector<int8_t> type(n), bc(n);
    std::vector<double> pos(3 * n), vel(3 * n), p(n), ndr(n), pmin(n);
    check(mps_download(h, type.data(), pos.data(), vel.data(), p.data(), ndr.data(), bc.data()), "mps_download");
    check(mps_download_min_pressure(h, pmin.data()), "mps_download_min_pressure");
    for (size_t i = 0; i < n; i++) {
        Particle& pi = particles[i];
        pi.type = static_cast<ParticleType>(type[i]);
        pi.position = Vec3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        pi.velocity = Vec3(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
        pi.pressure = p[i];
        pi.minimumPressure = pmin[i];
        pi.numberDensityRatio = ndr[i];
        pi.boundaryCondition = static_cast<BoundaryCondition>(bc[i]);
    }
}

// The debug-level logs of src/mps.cpp:405-410, 559-568 and 655-663, from the
// library's event record of the step.
void MPS::reportEvents(double courant) {
    int64_t counts[3];
    check(mps_debug_events(h, counts, nullptr, nullptr, 0), "mps_debug_events");
    if (counts[0] || counts[1]) {
        const int32_t cap = (int32_t) std::max<int64_t>(counts[0], counts[1]);
        std::vector<int32_t> g(cap > 0 ? cap : 1), u(cap > 0 ? cap : 1);
        check(mps_debug_events(h, counts, g.data(), u.data(), cap), "mps_debug_events");
        if (counts[0]) {
            std::vector<double> pos(3 * particles.size());
            check(mps_download(h, nullptr, pos.data(), nullptr, nullptr, nullptr, nullptr), "mps_download");
            for (int64_t k = 0; k < counts[0] && k < cap; k++) {
                const size_t i = (size_t) g[k];
                mpslog::debug("Particle (id=" + std::to_string(i) + ") is out of domain. (x, y, z)=(" +
                              std::to_string(pos[3 * i]) + ", " + std::to_string(pos[3 * i + 1]) + ", " +
                              std::to_string(pos[3 * i + 2]) + ") ");
            }
        }
        for (int64_t k = 0; k < counts[1] && k < cap; k++)
            mpslog::debug("There is no Dirichlet boundary condition connected to the particle (id=" +
                          std::to_string(u[k]) + ").");
    }
    if (counts[2]) mpslog::debug("Courant number is larger than CFL condition. Courant=" + std::to_string(courant));
    debugLogCount += (int) (counts[0] + counts[1] + counts[2]);
    if (courant > 1.0) mpslog::error("Courant number is larger than 1.0. Exiting Simulation.");
    check(mps_clear_debug_events(h), "mps_clear_debug_events");
}

// src/mps.cpp:111-130
double MPS::stepForward(const bool isTimeToExport) {
    double courant = 0.0;
    const int rc = mps_step(h, isTimeToExport ? 1 : 0, &courant);
    // the reference ignores its solver's status (src/mps.cpp:457-460): a
    // solve that stopped above its tolerance is logged and the run goes on
    if (rc == MPS_E_SOLVER)
        mpslog::debug(mps_last_error(h));
    else
        check(rc, "mps_step");
    reportEvents(courant);
    if (isTimeToExport && syncOnExport) syncParticles();
    lastCourant = courant;
    return courant;
}

void MPS::snapshotBegin(int slot) { check(mps_snapshot_begin(h, slot), "mps_snapshot_begin"); }

ExportView MPS::snapshotEnd(int slot) const {
    ExportView v;
    v.n = particles.size();
    const int rc = mps_snapshot_end(h, slot, &v.type, &v.pos, &v.vel, &v.p, &v.ndr, &v.bc);
    if (rc) throw std::runtime_error("mps_snapshot_end failed with status " + std::to_string(rc));
    return v;
}
