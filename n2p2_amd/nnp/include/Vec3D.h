// nnp::Vec3D for callers of libnnp (Vec3D.h of the reference): a 3-vector
// of doubles with the arithmetic the callers use.
#ifndef NNP_VEC3D_H
#define NNP_VEC3D_H

#include <cmath>
#include <cstddef>

namespace nnp
{

struct Vec3D
{
    double r[3];

    Vec3D() : r{0.0, 0.0, 0.0} {}
    Vec3D(double x, double y, double z) : r{x, y, z} {}

    double&       operator[](std::size_t i)       { return r[i]; }
    double const& operator[](std::size_t i) const { return r[i]; }

    Vec3D& operator+=(Vec3D const& v)
    {
        r[0] += v.r[0]; r[1] += v.r[1]; r[2] += v.r[2];
        return *this;
    }
    Vec3D& operator-=(Vec3D const& v)
    {
        r[0] -= v.r[0]; r[1] -= v.r[1]; r[2] -= v.r[2];
        return *this;
    }
    Vec3D& operator*=(double a) { r[0] *= a; r[1] *= a; r[2] *= a; return *this; }
    Vec3D& operator/=(double a) { return *this *= 1.0 / a; }
    double operator*(Vec3D const& v) const
    {
        return r[0] * v.r[0] + r[1] * v.r[1] + r[2] * v.r[2];
    }
    bool operator==(Vec3D const& v) const
    {
        return r[0] == v.r[0] && r[1] == v.r[1] && r[2] == v.r[2];
    }
    bool operator!=(Vec3D const& v) const { return !(*this == v); }
    double norm2() const { return r[0] * r[0] + r[1] * r[1] + r[2] * r[2]; }
    double norm() const { return std::sqrt(norm2()); }
    Vec3D& normalize() { return *this /= norm(); }
    Vec3D cross(Vec3D const& v) const
    {
        return Vec3D(r[1] * v.r[2] - r[2] * v.r[1],
                     r[2] * v.r[0] - r[0] * v.r[2],
                     r[0] * v.r[1] - r[1] * v.r[0]);
    }
};

inline Vec3D operator+(Vec3D a, Vec3D const& b) { return a += b; }
inline Vec3D operator-(Vec3D a, Vec3D const& b) { return a -= b; }
inline Vec3D operator-(Vec3D a) { return a *= -1.0; }
inline Vec3D operator*(Vec3D a, double s) { return a *= s; }
inline Vec3D operator*(double s, Vec3D a) { return a *= s; }
inline Vec3D operator/(Vec3D a, double s) { return a /= s; }

}  // namespace nnp

#endif
