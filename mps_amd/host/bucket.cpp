#include "bucket.hpp"

#include <stdexcept>
#include <string>

#include "../../include/mps_amd.h"

// src/bucket.cpp:10-22
Bucket::Bucket(const double& reMax, const Domain& domain, const int& particleSize) {
    this->length = reMax;
    this->domain = domain;
    this->numX = (int) (domain.x.length / length) + 3;
    this->numY = (int) (domain.y.length / length) + 3;
    this->numZ = (int) (domain.z.length / length) + 3;
    this->num = numX * numY * numZ;
    this->first.resize(num);
    this->last.resize(num);
    this->next.resize(particleSize);
}

// src/bucket.cpp:24-51, on the GPU: a device handle sorts the particles by
// cell key and reports the reference's first/last/next view.  The handle is
// created on the first call and reused while the particle count stays the
// same: later calls upload the positions and types only.
void Bucket::storeParticles(std::vector<Particle>& particles) {
    const int n = (int) particles.size();
    std::vector<int8_t> type(n);
    std::vector<double> pos(3 * (size_t) n);
    for (int i = 0; i < n; i++) {
        type[i] = (int8_t) static_cast<int>(particles[i].type);
        for (int d = 0; d < 3; d++) pos[3 * (size_t) i + d] = particles[i].position[d];
    }
    int rc = MPS_OK;
    if (!dev || dev_n != n) {
        mps_settings s{};
        s.dim = 3;  // keys over all three axes, whatever the z extent
        s.particle_distance = length;
        s.dt = 1.0;
        s.density = 1.0;
        s.re_pressure = s.re_viscosity = s.re_surface = s.re_max = length;
        s.domain_min[0] = domain.x.min;
        s.domain_max[0] = domain.x.max;
        s.domain_min[1] = domain.y.min;
        s.domain_max[1] = domain.y.max;
        s.domain_min[2] = domain.z.min;
        s.domain_max[2] = domain.z.max;
        std::vector<int32_t> id(n);
        std::vector<double> vel(3 * (size_t) n, 0.0);
        for (int i = 0; i < n; i++) id[i] = particles[i].id;
        mps_handle* h = nullptr;
        rc = mps_create(&s, n, id.data(), type.data(), pos.data(), vel.data(), &h);
        if (rc) throw std::runtime_error(std::string("Bucket::storeParticles: ") + mps_last_error(nullptr));
        dev.reset(h, [](mps_handle* p) { mps_destroy(p); });
        dev_n = n;
        int32_t dims[4];
        double len;
        rc = mps_bucket_shape(h, dims, &len);
        if (!rc && dims[0] != num) rc = MPS_E_INVALID;
    } else {
        rc = mps_upload(dev.get(), type.data(), pos.data(), nullptr, nullptr, nullptr, nullptr);
    }
    first.assign(num, -1);
    last.assign(num, -1);
    next.assign(n, -1);
    if (!rc) rc = mps_dump_bucket(dev.get(), first.data(), last.data(), next.data());
    if (rc) {
        const std::string msg = mps_last_error(dev.get());
        dev.reset();
        dev_n = -1;
        throw std::runtime_error("Bucket::storeParticles: " + msg);
    }
}
