// renzoku_iface.h — the slice of Renzoku's C++ interface the drop-in adapters
// implement, re-declared with the reference's signatures (aggregate.h:15-26,
// ray.h:14-47, hitrecord.h:24-60) so the adapters compile and run here; in
// the reference tree the adapters include the reference's own headers
// instead and nothing else changes.
#pragma once

#include <cmath>

namespace Renzoku {

typedef float Float;                                   // common.h:33

class Vec3 {
public:
    Vec3() : e{0.f, 0.f, 0.f} {}
    Vec3(Float x, Float y, Float z) : e{x, y, z} {}
    Float x() const { return e[0]; }
    Float y() const { return e[1]; }
    Float z() const { return e[2]; }
    Float operator[](int i) const { return e[i]; }
    Float e[3];
};

// Ray keeps the normalised direction (ray.h:17-36); the engine derives
// inv_d and org_inv_d from it exactly as the constructor does
class Ray {
public:
    Ray() {}
    Ray(const Vec3& o, const Vec3& d) : o(o) {
        Float len = std::sqrt(d.x() * d.x() + d.y() * d.y() + d.z() * d.z());
        Float inv = 1.0f / len;
        this->d = Vec3(d.x() * inv, d.y() * inv, d.z() * inv);
        if (std::isnan(this->d.x()) || std::isnan(this->d.y()) || std::isnan(this->d.z())) this->d = Vec3();
    }
    Vec3 org() const { return o; }
    Vec3 dir() const { return d; }

private:
    Vec3 o, d;
};

// the fields of HitRecord a traversal fills (hitrecord.h:24-60); the adapter
// adds the triangle index the engine returns
class HitRecord {
public:
    bool hit = false;
    Float t = 1e6f, beta = 0.f, gamma = 0.f;
    int triangle = -1;                                 // position in the flattened surface order
};

class Aggregate {
public:
    virtual ~Aggregate() {}
    virtual bool hit(const Ray& r, Float tmin, Float tmax, Float time, HitRecord& record) const = 0;
    virtual bool hit(const Ray& r, Float tmin, Float tmax, Float time) const = 0;
    // visibility between p and q: true = blocked (aggregate.cc:84-90)
    virtual bool hit(const Vec3& p, const Vec3& q, Float tmin, Float time) const = 0;
};

}  // namespace Renzoku
