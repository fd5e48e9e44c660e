// Caller of the reference's shadow entry points (Reconstruction.hpp:47-48)
// through the facade: two v1.3 buffers (camera, light) and the cameras'
// matrices; writes reconstructDofMotionShadows' image, then runs
// reconstructShadows on the light's own samples as queries (w raised by
// `lift`, a LIGHTSPACE_DENSITY channel of 1.5) and writes their colours.
// Used by tests/test_gpu_facade.py.
#include <cstdio>
#include <cstdlib>

#include "facade/TreeGather.hpp"

int main(int argc, char** argv) {
    if (argc < 7) {
        fprintf(stderr, "usage: %s camera.bin light.bin matrices.f32 image.rgba queries.rgba out_spp\n", argv[0]);
        return 2;
    }
    try {
        FW::UVTSampleBuffer cam(argv[1]), light(argv[2]);
        // matrices.f32: camera projection, camera-to-world, light projection,
        // light world-to-camera (row-major), camera far plane, lift
        float m[66];
        FILE* f = fopen(argv[3], "rb");
        if (!f || fread(m, sizeof(float), 66, f) != 66) return 3;
        fclose(f);
        FW::CameraParams cp, lp;
        for (int k = 0; k < 16; k++) {
            cp.projection[k] = m[k];
            cp.cameraToWorld[k] = m[16 + k];
            lp.projection[k] = m[32 + k];
            lp.worldToCamera[k] = m[48 + k];
        }
        cp.windowWidth = lp.windowWidth = cam.getWidth();
        cp.farPlane = m[64];
        FW::TreeGather tg(cam, cp), shadowTG(light, lp);
        tg.setOutputSpp(atoi(argv[6]));
        FW::Image image(cam.getWidth(), cam.getHeight());
        tg.reconstructDofMotionShadows(image, shadowTG);
        f = fopen(argv[4], "wb");
        if (!f) return 3;
        fwrite(image.getPtr(), sizeof(float), (size_t)cam.getWidth() * cam.getHeight() * 4, f);
        fclose(f);

        std::vector<float> lifted = light.records();
        for (size_t k = 0; k < lifted.size(); k += 16) lifted[k + 3] += m[65];   // w of every query
        FW::UVTSampleBuffer q2(light.getWidth(), light.getHeight(), light.getNumSamples(), light.getCocCoeffs(),
                               lifted);
        const int cid = q2.reserveChannel("LIGHTSPACE_DENSITY");
        for (int y = 0; y < q2.getHeight(); y++)
            for (int x = 0; x < q2.getWidth(); x++)
                for (int i = 0; i < q2.getNumSamples(); i++) q2.setSampleExtra<float>(cid, x, y, i, 1.5f);
        shadowTG.reconstructShadows(&q2);
        f = fopen(argv[5], "wb");
        if (!f) return 3;
        for (int y = 0; y < q2.getHeight(); y++)
            for (int x = 0; x < q2.getWidth(); x++)
                for (int i = 0; i < q2.getNumSamples(); i++) {
                    const FW::Vec4f c = q2.getSampleColor(x, y, i);
                    const float v[4] = {c.x, c.y, c.z, c.w};
                    fwrite(v, sizeof(float), 4, f);
                }
        fclose(f);
        printf("shadows %dx%d, %.3f ms on device\n", cam.getWidth(), cam.getHeight(), tg.lastStats().total_ms);
    } catch (const FW::ReconstructionError& e) {
        fprintf(stderr, "error %d: %s\n", (int)e.status, e.what());
        return 1;
    }
    return 0;
}
