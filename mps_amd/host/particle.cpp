#include "particle.hpp"

// src/particle.cpp:5-17
Particle::Particle(const int& id, const ParticleType& type, const Vec3& position, const Vec3& velocity,
                   const double& density)
    : id(id), type(type), position(position), velocity(velocity), density(density) {}

// src/particle.cpp:19-29
double Particle::inverseDensity() const {
    switch (type) {
        case ParticleType::Fluid:
            return 1 / density;
        case ParticleType::Wall:
        case ParticleType::DummyWall:
        case ParticleType::Ghost:
            return 0;
    }
    return 0;
}
