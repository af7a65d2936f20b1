// adapter_test.cpp — drives AggregateGpu through Aggregate* (the reference's
// virtual interface) on a scene and rays read from a file, checks the batched
// forms against the one-ray calls, and writes the answers for the Python test
// to compare with the oracle.
//
// input : int32 ntri, nsurf, nmat, nlight, nray, nseg
//         f32 pos[ntri*9], tan[ntri*3]; i32 tri_surface[ntri], surf_material, surf_light, surf_flags [nsurf]
//         f32 mat[nmat*12], radiance[nlight*3], rays[nray*6] (o, d unnormalised), seg_p[nseg*3], seg_q[nseg*3]
// output: per ray f32 t, beta, gamma, i32 tri, u8 any; per segment u8 blocked; per ray f32 normal[3],
//         shading_normal[3], uvn[9] (u, v, n), tangent[3], i32 surface index of record.surface and of
//         record.material (-1 for null), the record's pointers checked against the surface bindings
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../renzoku_amd/adapters/aggregate_gpu.h"

using namespace Renzoku;

template <typename T>
static bool rd(FILE* f, std::vector<T>& v, size_t n) {
    v.resize(n);
    return n == 0 || fread(v.data(), sizeof(T), n, f) == n;
}

int main(int argc, char** argv) {
    if (argc != 3) { fprintf(stderr, "usage: adapter_test in out\n"); return 2; }
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    int32_t hdr[6];
    if (fread(hdr, 4, 6, f) != 6) return 2;
    int ntri = hdr[0], nsurf = hdr[1], nmat = hdr[2], nlight = hdr[3], nray = hdr[4], nseg = hdr[5];
    std::vector<float> pos, tan, mat, rad, rays, sp, sq;
    std::vector<int32_t> ts, sm, sl, sf;
    bool ok = rd(f, pos, 9 * (size_t)ntri) && rd(f, tan, 3 * (size_t)ntri) && rd(f, ts, ntri) && rd(f, sm, nsurf) &&
              rd(f, sl, nsurf) && rd(f, sf, nsurf) && rd(f, mat, 12 * (size_t)nmat) && rd(f, rad, 3 * (size_t)nlight) &&
              rd(f, rays, 6 * (size_t)nray) && rd(f, sp, 3 * (size_t)nseg) && rd(f, sq, 3 * (size_t)nseg);
    fclose(f);
    if (!ok) return 2;
    rzk_ctx* ctx = nullptr;
    if (rzk_ctx_create(0, &ctx) != RZK_OK) { fprintf(stderr, "no GPU context\n"); return 3; }
    rzk_scene_desc d{};
    d.num_triangles = ntri; d.positions = pos.data(); d.tangents = tan.data(); d.tri_surface = ts.data();
    d.num_surfaces = nsurf; d.surf_material = sm.data(); d.surf_light = sl.data(); d.surf_flags = sf.data();
    d.num_materials = nmat; d.materials = mat.data(); d.num_lights = nlight; d.light_radiance = rad.data();
    int status = 0;
    {
        // one binding per surface: distinct addresses standing for the reference's Surface /
        // Material / AreaLight objects (material and light null where the surface has none)
        std::vector<char> surf_obj(nsurf + 1), mat_obj(nmat + 1), light_obj(nlight + 1);
        std::vector<SurfaceBinding> bind(nsurf);
        for (int k = 0; k < nsurf; ++k) {
            bind[k].surface = reinterpret_cast<Surface*>(&surf_obj[k]);
            bind[k].material = sm[k] >= 0 ? reinterpret_cast<Material*>(&mat_obj[sm[k]]) : nullptr;
            bind[k].light = sl[k] >= 0 ? reinterpret_cast<AreaLight*>(&light_obj[sl[k]]) : nullptr;
        }
        AggregateGpu gpu(ctx, d, bind);
        const Aggregate* agg = &gpu;                  // callers see only the reference's interface
        const Float tmin = 1e-3f, tmax = 1e6f;
        std::vector<Ray> rs(nray);
        std::vector<float> t(nray), b(nray), g(nray), frame(18 * (size_t)nray);
        std::vector<int32_t> tri(nray), sidx(nray), midx(nray);
        std::vector<uint8_t> any(nray), blocked(nseg);
        for (int i = 0; i < nray; ++i) {
            const float* r = &rays[6 * (size_t)i];
            rs[i] = Ray(Vec3(r[0], r[1], r[2]), Vec3(r[3], r[4], r[5]));
            HitRecord rec;
            bool h = agg->hit(rs[i], tmin, tmax, 0.f, rec);
            t[i] = h ? rec.t : 0.f; b[i] = h ? rec.beta : 0.f; g[i] = h ? rec.gamma : 0.f;
            any[i] = agg->hit(rs[i], tmin, tmax, 0.f) ? 1 : 0;
            if (any[i] != (h ? 1 : 0)) status = 1;    // any hit <=> a closest hit in [tmin, tmax]
            Vec3 v[6] = {rec.normal, rec.shading_normal, rec.uvn.u(), rec.uvn.v(), rec.uvn.n(), rec.tangent};
            for (int k = 0; k < 6; ++k)
                for (int c = 0; c < 3; ++c) frame[18 * (size_t)i + 3 * k + c] = h ? v[k][c] : 0.f;
            sidx[i] = rec.surface ? (int32_t)(reinterpret_cast<char*>(rec.surface) - surf_obj.data()) : -1;
            midx[i] = rec.material ? (int32_t)(reinterpret_cast<char*>(rec.material) - mat_obj.data()) : -1;
            if (h && sidx[i] >= 0 && (rec.material != bind[sidx[i]].material || rec.light != bind[sidx[i]].light))
                status = 1;                           // the record carries its surface's bindings
        }
        // triangle indices come out of band from the batched form
        {
            std::vector<HitRecord> hb;
            std::vector<int> tb;
            gpu.hit_batch(rs, tmin, tmax, hb, &tb);
            for (int i = 0; i < nray; ++i) tri[i] = tb[i];
        }
        std::vector<Vec3> P(nseg), Q(nseg);
        for (int i = 0; i < nseg; ++i) {
            P[i] = Vec3(sp[3 * i], sp[3 * i + 1], sp[3 * i + 2]);
            Q[i] = Vec3(sq[3 * i], sq[3 * i + 1], sq[3 * i + 2]);
            blocked[i] = agg->hit(P[i], Q[i], tmin, 0.f) ? 1 : 0;
        }
        // the batched forms answer the same
        std::vector<HitRecord> hb;
        gpu.hit_batch(rs, tmin, tmax, hb);
        for (int i = 0; i < nray; ++i)
            if (hb[i].hit != (tri[i] >= 0) || (tri[i] >= 0 && memcmp(&hb[i].t, &t[i], 4) != 0)) status = 1;
        std::vector<uint8_t> vb;
        gpu.visible_batch(P, Q, tmin, vb);
        if (vb != blocked) status = 1;
        FILE* o = fopen(argv[2], "wb");
        if (!o) return 2;
        fwrite(t.data(), 4, nray, o); fwrite(b.data(), 4, nray, o); fwrite(g.data(), 4, nray, o);
        fwrite(tri.data(), 4, nray, o); fwrite(any.data(), 1, nray, o); fwrite(blocked.data(), 1, nseg, o);
        fwrite(frame.data(), 4, 18 * (size_t)nray, o); fwrite(sidx.data(), 4, nray, o); fwrite(midx.data(), 4, nray, o);
        fclose(o);
    }
    rzk_ctx_destroy(ctx);
    printf("adapter_test: %d rays, %d segments, batch == single: %s\n", nray, nseg, status == 0 ? "yes" : "NO");
    return status;
}
