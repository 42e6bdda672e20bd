// settings.hpp — Domain and Settings as the reference defines them
// (src/settings.hpp:7-86), loaded from the same YAML key set
// (src/settings.cpp:21-84) by a project-owned reader for that subset of YAML.
#pragma once

#include <filesystem>

#include "../../include/mps_amd.h"
#include "vec3.hpp"

class Domain {
public:
    struct {
        double min, max, length;
    } x{}, y{}, z{};

    Domain() = default;
    Domain(const double& xMin, const double& xMax, const double& yMin, const double& yMax, const double& zMin,
           const double& zMax);
};

class Settings {
public:
    // computational conditions
    double dim = 0;
    double particleDistance = 0;
    double dt = 0;
    double finishTime = 0;
    double outputInterval = 0;
    double cflCondition = 0;

    Domain domain;

    double density = 0;
    double kinematicViscosity = 0;

    Vec3 gravity;

    struct {
        double pressure, viscosity, surfaceDetection, max;
    } effectiveRadius{};

    double thresholdForSurfaceDetection = 0;

    struct {
        bool on;
        double compressibility;
    } quasiCompressibility{};
    struct {
        bool on;
        double gamma;
    } higherOrderSourceTerm{};
    double relaxationCoefficientForPressure = 0;

    double collisionDistance = 0;
    double coefficientOfRestitution = 0;

    std::filesystem::path inputCsvPath;

    // Reads the YAML file; like the reference, a parse or type error is
    // logged and loading stops there (src/settings.cpp:86-88).  Returns false
    // on such an error.
    bool load(const std::filesystem::path& inputYamlPath);

    // POD view handed to the C ABI.
    mps_settings toC() const;
};
