// transport_test.cpp — drives GpuLightTransportMatrix::fill through
// TransportMatrix* (light_transport_matrix.h:16-55) with a GpuVplEvaluator, and
// the evaluator's per-pair calls through IVirtualPointLightEvaluator*
// (vpl_interface.h:126-139), and writes what the Python test compares with the
// oracle.
//
// input : i32 ntri, nsurf, nmat, nlight, nvpl, npix, npairs, clamp_mode; f32 clamp, incident_clamp
//         f32 pos[ntri*9], tan[ntri*3]; i32 tri_surface[ntri], surf_material, surf_light, surf_flags [nsurf]
//         f32 mat[nmat*12], radiance[nlight*3], camera[19]
//         f32 vpl[nvpl*21]; i32 vpl_mat[nvpl]; i32 pixels[npix*2] (x, y); i32 pairs[npairs*2] (pixel row, vpl)
// output: f32 M[npix][nvpl][3] (row major); per pair: f32 radiance[3], f32 visible, f32 incident[3]
#include <cstdio>
#include <vector>

#include "../../renzoku_amd/adapters/gpu_transport_matrix.h"

using namespace Renzoku;

template <typename T>
static bool rd(FILE* f, std::vector<T>& v, size_t n) {
    v.resize(n);
    return n == 0 || fread(v.data(), sizeof(T), n, f) == n;
}

int main(int argc, char** argv) {
    if (argc != 3) { fprintf(stderr, "usage: transport_test in out\n"); return 2; }
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    int32_t hdr[8];
    float thr[2];
    if (fread(hdr, 4, 8, f) != 8 || fread(thr, 4, 2, f) != 2) return 2;
    const int ntri = hdr[0], nsurf = hdr[1], nmat = hdr[2], nlight = hdr[3], nvpl = hdr[4], npix = hdr[5];
    const int npairs = hdr[6], clamp_mode = hdr[7];
    std::vector<float> pos, tan, mat, rad, cam, vrec;
    std::vector<int32_t> ts, sm, sl, sf, vmat, pix, pairs;
    bool ok = rd(f, pos, 9 * (size_t)ntri) && rd(f, tan, 3 * (size_t)ntri) && rd(f, ts, ntri) && rd(f, sm, nsurf) &&
              rd(f, sl, nsurf) && rd(f, sf, nsurf) && rd(f, mat, 12 * (size_t)nmat) && rd(f, rad, 3 * (size_t)nlight) &&
              rd(f, cam, RZK_CAM_FLOATS) && rd(f, vrec, RZK_VPL_FLOATS * (size_t)nvpl) && rd(f, vmat, nvpl) &&
              rd(f, pix, 2 * (size_t)npix) && rd(f, pairs, 2 * (size_t)npairs);
    fclose(f);
    if (!ok) { fprintf(stderr, "bad input\n"); return 2; }

    rzk_ctx* ctx = nullptr;
    if (rzk_ctx_create(0, &ctx) != RZK_OK) { fprintf(stderr, "no GPU context\n"); return 3; }
    rzk_scene_desc d{};
    d.num_triangles = ntri; d.positions = pos.data(); d.tangents = tan.data(); d.tri_surface = ts.data();
    d.num_surfaces = nsurf; d.surf_material = sm.data(); d.surf_light = sl.data(); d.surf_flags = sf.data();
    d.num_materials = nmat; d.materials = mat.data(); d.num_lights = nlight; d.light_radiance = rad.data();
    // stand-ins for the scene's Material objects: one distinct address per material
    std::vector<char> tags((size_t)nmat + 1);
    auto material = [&](int i) { return i >= 0 ? reinterpret_cast<Material*>(&tags[(size_t)i]) : nullptr; };
    std::vector<SurfaceBinding> bind((size_t)nsurf);
    for (int i = 0; i < nsurf; ++i) bind[i].material = material(sm[i]);
    FILE* o = fopen(argv[2], "wb");
    if (!o) return 2;
    int status = 0;
    try {
        AggregateGpu agg(ctx, d, bind);
        BrdfPointLights vpls;
        for (int k = 0; k < nvpl; ++k) {
            const float* v = &vrec[RZK_VPL_FLOATS * (size_t)k];
            LocalGeometry dg(Vec3(v[0], v[1], v[2]), Vec3(v[3], v[4], v[5]), Vec3(v[3], v[4], v[5]),
                             Onb(Vec3(v[6], v[7], v[8]), Vec3(v[9], v[10], v[11]), Vec3(v[12], v[13], v[14])), Vec3(),
                             Vec2());
            vpls.push_back(BrdfPointLight(dg, Rgb(v[18], v[19], v[20]), material(vmat[k]), Vec3(v[15], v[16], v[17])));
        }
        Pixels pixels;
        for (int i = 0; i < npix; ++i) pixels.push_back(Vec2((Float)pix[2 * i], (Float)pix[2 * i + 1]));
        Scene scene(Size2(0, 0));
        GpuVplEvaluator ev(&agg, clamp_mode, thr[0], thr[1]);
        GpuLightTransportMatrix gm(cam.data());
        TransportMatrix* tm = &gm;
        tm->fill(&scene, &ev, vpls, pixels);
        MatrixC<Rgb>& M = gm.get_matrix();
        for (int i = 0; i < M.get_rows(); ++i)
            for (int k = 0; k < M.get_cols(); ++k) fwrite(M(i, k).e, 4, 3, o);
        // per-pair calls on the receivers of the pixel-centre primary hits (the engine's records)
        IVirtualPointLightEvaluator* e = &ev;
        std::vector<int32_t> px(npix), py(npix), rmat(npix), rlight(npix);
        for (int i = 0; i < npix; ++i) { px[i] = pix[2 * i]; py[i] = pix[2 * i + 1]; }
        std::vector<float> rrec(RZK_RECV_FLOATS * (size_t)npix);
        if (rzk_camera_primary(ctx, agg.handle(), cam.data(), npix, px.data(), py.data(), nullptr, nullptr, rrec.data(),
                               rmat.data(), rlight.data()) != RZK_OK)
            throw std::runtime_error(rzk_ctx_last_error(ctx));
        for (int q = 0; q < npairs; ++q) {
            const int row = pairs[2 * q], k = pairs[2 * q + 1];
            float out[7] = {0, 0, 0, 0, 0, 0, 0};
            if (rmat[row] >= 0) {
                const float* c = &rrec[RZK_RECV_FLOATS * (size_t)row];
                Receiver recv;
                recv.p = Vec3(c[0], c[1], c[2]);
                recv.n = Vec3(c[3], c[4], c[5]);
                recv.shading_n = Vec3(c[6], c[7], c[8]);
                recv.wo = Vec3(c[9], c[10], c[11]);
                recv.shading_uvn = Onb(Vec3(c[12], c[13], c[14]), Vec3(c[15], c[16], c[17]), Vec3(c[18], c[19], c[20]));
                recv.m = material(rmat[row]);
                bool visible = false;
                Rgb L = e->radiance(&scene, recv, vpls[k], visible);
                Rgb I = e->incident_radiance(&scene, recv, vpls[k]);
                out[0] = L.red(); out[1] = L.green(); out[2] = L.blue(); out[3] = visible ? 1.f : 0.f;
                out[4] = I.red(); out[5] = I.green(); out[6] = I.blue();
            }
            fwrite(out, 4, 7, o);
        }
    } catch (const std::exception& ex) {
        fprintf(stderr, "transport_test: %s\n", ex.what());
        status = 1;
    }
    fclose(o);
    rzk_ctx_destroy(ctx);
    printf("transport_test: %d pixels x %d VPLs, %d pairs, status %d\n", npix, nvpl, npairs, status);
    return status;
}
