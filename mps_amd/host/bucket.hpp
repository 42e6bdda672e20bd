// bucket.hpp — the reference's uniform bucket grid (src/bucket.hpp:8-19).
// Geometry as Bucket::Bucket computes it; storeParticles fills the linked
// lists from the GPU counting sort (libmps_amd.so), bit-identical to the
// reference's serial insertion.
#pragma once

#include <memory>
#include <vector>

#include "particle.hpp"
#include "settings.hpp"

struct mps_handle;

class Bucket {
public:
    int num = 0, numX = 0, numY = 0, numZ = 0;
    double length = 0;
    Domain domain;
    std::vector<int> next, first, last;

    Bucket() = default;
    Bucket(const double& reMax, const Domain& domain, const int& particleSize);
    // Throws std::runtime_error if the GPU library fails.
    void storeParticles(std::vector<Particle>& particles);

private:
    // The device grid of the last call, reused while the particle count stays
    // the same (later calls only upload positions and types).  Copies of a
    // Bucket share it, so one Bucket and its copies are driven by one thread.
    std::shared_ptr<mps_handle> dev;
    int dev_n = -1;
};
