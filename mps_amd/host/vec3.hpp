// vec3.hpp — the slice of Eigen::Vector3d the reference's class surface uses
// (x()/y()/z(), norm(), dot(), normalized(), arithmetic).  Reductions sum as
// (a0 + a1) + a2, the order Eigen's vectorised redux gives a Vector3d.
#pragma once

#include <cmath>

class Vec3 {
public:
    double v[3] = {0.0, 0.0, 0.0};

    Vec3() = default;
    Vec3(double x, double y, double z) : v{x, y, z} {}

    static Vec3 Zero() { return Vec3(); }
    void setZero() { v[0] = v[1] = v[2] = 0.0; }

    double& x() { return v[0]; }
    double& y() { return v[1]; }
    double& z() { return v[2]; }
    double x() const { return v[0]; }
    double y() const { return v[1]; }
    double z() const { return v[2]; }
    double& operator[](int k) { return v[k]; }
    double operator[](int k) const { return v[k]; }

    double squaredNorm() const { return (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]; }
    double norm() const { return std::sqrt(squaredNorm()); }
    double dot(const Vec3& o) const { return (v[0] * o.v[0] + v[1] * o.v[1]) + v[2] * o.v[2]; }
    Vec3 normalized() const {
        const double z2 = squaredNorm();
        if (z2 > 0.0) {
            const double s = std::sqrt(z2);
            return Vec3(v[0] / s, v[1] / s, v[2] / s);
        }
        return *this;
    }

    Vec3 operator+(const Vec3& o) const { return Vec3(v[0] + o.v[0], v[1] + o.v[1], v[2] + o.v[2]); }
    Vec3 operator-(const Vec3& o) const { return Vec3(v[0] - o.v[0], v[1] - o.v[1], v[2] - o.v[2]); }
    Vec3 operator*(double s) const { return Vec3(v[0] * s, v[1] * s, v[2] * s); }
    Vec3 operator/(double s) const { return Vec3(v[0] / s, v[1] / s, v[2] / s); }
    Vec3& operator+=(const Vec3& o) {
        for (int k = 0; k < 3; k++) v[k] += o.v[k];
        return *this;
    }
    Vec3& operator-=(const Vec3& o) {
        for (int k = 0; k < 3; k++) v[k] -= o.v[k];
        return *this;
    }
    Vec3& operator*=(double s) {
        for (int k = 0; k < 3; k++) v[k] *= s;
        return *this;
    }
};

inline Vec3 operator*(double s, const Vec3& a) { return a * s; }
