// renzoku_iface.h — the slice of Renzoku's C++ interface the drop-in adapters
// implement or read, re-declared with the reference's names, signatures and
// fields so the adapters compile and run here: Aggregate (aggregate.h:15-26),
// Ray (ray.h:14-47), HitRecord (hitrecord.h:24-60), Integrator /
// MonteCarloIntegrator (integrator.h:54-89, monte_carlo_integrator.h:28-54),
// the sensor samples and primary-ray bundles (sensor.h:9-15, :124-149), the
// frame buffer's current image (frame.h:55-68), Receiver / LocalGeometry
// (local_geometry.h:11-60), BrdfPointLight (brdf_point_light.h:30-290), the
// VPL evaluator (vpl_interface.h:126-139) and the transport matrices
// (light_transport_matrix.h:16-80, matrix_c.h).  Scene objects the adapters
// only pass around (Material, AreaLight, Surface, Shape) stay opaque.
// In the reference tree the adapters include the reference's own headers
// instead and nothing else changes.
#pragma once

#include <cmath>
#include <string>
#include <vector>

namespace Renzoku {

typedef float Float;                                   // common.h:33
const Float RAY_TMAX = 1e6f;                           // common.h:50

class Vec2 {
public:
    Vec2() : e{0.f, 0.f} {}
    Vec2(Float x, Float y) : e{x, y} {}
    Float x() const { return e[0]; }
    Float y() const { return e[1]; }
    Float e[2];
};

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

class Rgb {
public:
    Rgb() : e{0.f, 0.f, 0.f} {}
    Rgb(Float r, Float g, Float b) : e{r, g, b} {}
    Float red() const { return e[0]; }
    Float green() const { return e[1]; }
    Float blue() const { return e[2]; }
    Float e[3];
};

// Onb (onb.h): the local frame u, v, n of a hit
class Onb {
public:
    Onb() {}
    Onb(const Vec3& u, const Vec3& v, const Vec3& n) : U(u), V(v), N(n) {}
    Vec3 u() const { return U; }
    Vec3 v() const { return V; }
    Vec3 n() const { return N; }

private:
    Vec3 U, V, N;
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

class Material;
class AreaLight;
class EnvLight;
class Shape;
class Surface;

// hitrecord.h:24-60, every field of the reference's HitRecord
class HitRecord {
public:
    bool hit;
    Float t, beta, gamma;
    Vec3 tangent, normal;
    Vec3 shading_normal;
    Onb uvn;
    Material* material;
    AreaLight* light;
    EnvLight* env_light;
    Shape* shape;
    Surface* surface;
    Vec2 uv;

    HitRecord()
        : hit(false), t(RAY_TMAX), beta(0.0f), gamma(0.0f), material(nullptr), light(nullptr), env_light(nullptr),
          shape(nullptr), surface(nullptr) {}
};

class Aggregate {
public:
    virtual ~Aggregate() {}
    virtual bool hit(const Ray& r, Float tmin, Float tmax, Float time, HitRecord& record) const = 0;
    virtual bool hit(const Ray& r, Float tmin, Float tmax, Float time) const = 0;
    // visibility between p and q: true = blocked (aggregate.cc:84-90)
    virtual bool hit(const Vec3& p, const Vec3& q, Float tmin, Float time) const = 0;
};

// local_geometry.h:11-33
struct Receiver {
    Vec3 p;
    Vec3 n;                  // geometric normal
    Vec3 wo;
    Material* m;
    AreaLight* light;
    Vec2 uv;
    Vec3 tangent;
    Vec3 shading_n;
    Onb shading_uvn;
    Receiver() : m(nullptr), light(nullptr) {}
    Receiver(const Ray& r, const HitRecord& h)
        : p(Vec3(r.org().x() + h.t * r.dir().x(), r.org().y() + h.t * r.dir().y(), r.org().z() + h.t * r.dir().z())),
          n(h.normal), wo(Vec3(-r.dir().x(), -r.dir().y(), -r.dir().z())), m(h.material), light(h.light), uv(h.uv),
          tangent(h.tangent), shading_n(h.shading_normal), shading_uvn(h.uvn) {}
};

// local_geometry.h:38-60
struct LocalGeometry {
    Vec3 p;
    Vec3 n;                  // shading normal
    Vec3 ng;                 // geometric normal
    Vec3 tangent;
    Vec2 uv;
    Onb uvn;
    LocalGeometry() {}
    LocalGeometry(const Vec3& p, const Vec3& n, const Vec3& ng, const Onb& uvn, const Vec3& tangent, const Vec2 uv)
        : p(p), n(n), ng(ng), tangent(tangent), uv(uv), uvn(uvn) {}
    explicit LocalGeometry(const Receiver& r)
        : p(r.p), n(r.shading_n), ng(r.n), tangent(r.tangent), uv(r.uv), uvn(r.shading_uvn) {}
};

// brdf_point_light.h: the parts a transport matrix reads.  A VPL on a light
// (no material) has a white BRDF (brdf_point_light.h:235-240).
class BrdfPointLight {
public:
    BrdfPointLight() : m(nullptr) {}
    BrdfPointLight(const LocalGeometry& dg, Rgb power, Material* m, const Vec3& wi)
        : dg(dg), contrib(power), m(m), wi(wi) {}
    Vec3 org() const { return dg.p; }
    Vec3 normal() const { return dg.n; }
    LocalGeometry get_dg() const { return dg; }
    Rgb power() const { return contrib; }
    Rgb contribution() const { return contrib; }
    Vec3 get_wi() const { return wi; }
    Material* get_material() const { return m; }
    bool is_on_light() const { return m == nullptr; }

private:
    LocalGeometry dg;
    Rgb contrib;
    Material* m;
    Vec3 wi;
};
typedef std::vector<BrdfPointLight> BrdfPointLights;   // common.h:123
typedef std::vector<Vec2> Pixels;                       // image_sampler.h:8

class Scene;

// vpl_interface.h:126-139
class IVirtualPointLightEvaluator {
public:
    virtual ~IVirtualPointLightEvaluator() {}
    virtual Rgb radiance(Scene* scene, const Receiver& r, const BrdfPointLight& light) {
        bool visible;
        return radiance(scene, r, light, visible);
    }
    virtual Rgb radiance(Scene* scene, const Receiver& r, const BrdfPointLight& light, bool& visible) = 0;
    virtual Rgb incident_radiance(Scene* scene, const Receiver& r, const BrdfPointLight& light) = 0;
};

// matrix_c.h: a column-major rows x cols matrix, element (i, j) = row i, column j
template <typename T>
class MatrixC {
public:
    MatrixC(int rows = 0, int cols = 0) { resize(rows, cols); }
    void resize(int rows, int cols) {
        this->rows = rows;
        this->cols = cols;
        data.assign((size_t)rows * cols, T());
    }
    T& operator()(int i, int j) { return data[(size_t)j * rows + i]; }
    const T& operator()(int i, int j) const { return data[(size_t)j * rows + i]; }
    int get_rows() const { return rows; }
    int get_cols() const { return cols; }
    T* get_data() { return data.data(); }

private:
    int rows = 0, cols = 0;
    std::vector<T> data;
};

// light_transport_matrix.h:16-55
class TransportMatrix {
public:
    virtual ~TransportMatrix() {}
    virtual void fill(Scene* scene, IVirtualPointLightEvaluator* renderer, BrdfPointLights& vpls, Pixels& pixels) = 0;
    virtual int get_cols() const = 0;
    virtual int get_rows() const = 0;
};

// light_transport_matrix.h:57-80: rows = pixel samples, columns = VPLs
class LightTransportMatrix : public TransportMatrix {
public:
    LightTransportMatrix(int rows = 0, int cols = 0) : m(rows, cols) {}
    int get_rows() const override { return m.get_rows(); }
    int get_cols() const override { return m.get_cols(); }
    virtual MatrixC<Rgb>& get_matrix() { return m; }

protected:
    MatrixC<Rgb> m;
};

// sensor.h:9-15
struct SensorSample {
    int x, y;
    Vec2 jitter;
    SensorSample() : x(0), y(0), jitter(Vec2(0.5f, 0.5f)) {}
};

enum SensorType { SENSOR_RGB };

// sensor.h:124-133
struct PrimaryRay {
    Ray r;
    SensorSample s;
    SensorType type;
    Rgb val;
    PrimaryRay(const Ray& r, const SensorSample& s) : r(r), s(s), type(SENSOR_RGB), val() {}
};
typedef std::vector<PrimaryRay> PrimaryRays;

// sensor.h:137-149
struct PrimaryRayBundle {
    enum RayBundleEnum { BUNDLE_DEFAULT, BUNDLE_BEGIN_FRAME, BUNDLE_END_FRAME } tag;
    PrimaryRays rays;
    int x0, y0, x1, y1;                                // image block to update
    PrimaryRayBundle() : tag(BUNDLE_DEFAULT), x0(0), y0(0), x1(0), y1(0) {}
};

struct Size2 {
    int height, width;
    Size2(int h = 0, int w = 0) : height(h), width(w) {}
};

// ImageFloat (image.h): height x width x channels floats, element (i, j, c) = row i, column j
class ImageFloat {
public:
    ImageFloat(Float init, int height, int width, int channels)
        : h(height), w(width), ch(channels), data((size_t)height * width * channels, init) {}
    Float& operator()(int i, int j, int c) { return data[((size_t)i * w + j) * ch + c]; }
    Float operator()(int i, int j, int c) const { return data[((size_t)i * w + j) * ch + c]; }
    int get_height() const { return h; }
    int get_width() const { return w; }
    const Float* get_data() const { return data.data(); }

private:
    int h, w, ch;
    std::vector<Float> data;
};

// frame.h:55-68: the latest progressive result
class FrameBuffer {
public:
    explicit FrameBuffer(Size2 img_size) : buffer(new ImageFloat(0.f, img_size.height, img_size.width, 3)) {}
    ~FrameBuffer() { delete buffer; }
    FrameBuffer(const FrameBuffer&) = delete;
    FrameBuffer& operator=(const FrameBuffer&) = delete;
    ImageFloat* get_current_buffer() const { return buffer; }

private:
    ImageFloat* buffer;
};

// the parts of Scene (scene.h) an integrator reads here
class Scene {
public:
    explicit Scene(Size2 img_size) : img_size(img_size) {}
    virtual ~Scene() {}
    Size2 get_image_size() const { return img_size; }

private:
    Size2 img_size;
};

// integrator.h:54-89
class Integrator {
public:
    Integrator() : scene(nullptr), max_bounce(1) {}
    virtual ~Integrator() {}
    virtual void initialize(Scene* scene) { this->scene = scene; }
    virtual void integrate() {}
    int get_max_bounce() const { return max_bounce; }
    void set_max_bounce(int max_bounce) { this->max_bounce = max_bounce; }
    Scene* get_scene() const { return scene; }

protected:
    Scene* scene;
    int max_bounce;
};

// monte_carlo_integrator.h:12-24
class IMonteCarloIntegrator {
public:
    IMonteCarloIntegrator() : num_samples(1) {}
    virtual ~IMonteCarloIntegrator() {}
    int get_num_samples() const { return num_samples; }
    void set_num_samples(int num_samples) { this->num_samples = num_samples; }

protected:
    int num_samples;
};

// monte_carlo_integrator.h:28-54: integrate(bundle, frame) evaluates every
// primary ray of the bundle (radiance(PrimaryRay&)) and accumulates it into
// target / weight and the frame's current image
class MonteCarloIntegrator : public Integrator, public IMonteCarloIntegrator {
public:
    MonteCarloIntegrator() : target(nullptr), weight(nullptr) {}
    ~MonteCarloIntegrator() override {
        delete target;
        delete weight;
    }
    void initialize(Scene* scene) override {                     // monte_carlo_integrator.cc:15-21
        Integrator::initialize(scene);
        Size2 img_size = scene->get_image_size();
        delete target;
        delete weight;
        target = new ImageFloat(0.0f, img_size.height, img_size.width, 3);
        weight = new ImageFloat(0.0f, img_size.height, img_size.width, 1);
    }
    // monte_carlo_integrator.cc:45-70
    virtual void integrate(PrimaryRayBundle& b, FrameBuffer* frame) {
        for (size_t i = 0; i < b.rays.size(); ++i) radiance(b.rays[i]);
        Float filter_weight = 0.25f;                // 2x2 non-overlap box filter at each pixel
        ImageFloat* img = frame->get_current_buffer();
        for (size_t k = 0; k < b.rays.size(); ++k) {
            int i = b.rays[k].s.y;
            int j = b.rays[k].s.x;
            Rgb v = b.rays[k].val;
            (*target)(i, j, 0) += filter_weight * v.red();
            (*target)(i, j, 1) += filter_weight * v.green();
            (*target)(i, j, 2) += filter_weight * v.blue();
            (*weight)(i, j, 0) += filter_weight;
            if ((*weight)(i, j, 0) == 0) continue;
            Float inv_weight = 1.0f / (*weight)(i, j, 0);
            (*img)(i, j, 0) = (*target)(i, j, 0) * inv_weight;
            (*img)(i, j, 1) = (*target)(i, j, 1) * inv_weight;
            (*img)(i, j, 2) = (*target)(i, j, 2) * inv_weight;
        }
    }
    virtual void on_frame_begin() {}
    virtual void on_frame_end() {}

protected:
    virtual void radiance(PrimaryRay& pr) { pr.val = Rgb(); }   // monte_carlo_integrator.cc:40-43

    ImageFloat* target;
    ImageFloat* weight;
};

}  // namespace Renzoku
