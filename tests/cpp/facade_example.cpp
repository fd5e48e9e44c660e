// Caller written against the reference interface (Reconstruction.hpp:42-48):
// load a v1.3 sample buffer, build a TreeGather, reconstruct into an Image,
// and write the image as raw RGBA floats.  Used by tests/test_gpu_facade.py.
#include <cstdio>
#include <cstdlib>

#include "facade/TreeGather.hpp"

int main(int argc, char** argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s samples.bin out.rgba out_spp\n", argv[0]);
        return 2;
    }
    try {
        FW::UVTSampleBuffer sbuf(argv[1]);
        FW::CameraParams params;
        FW::TreeGather tg(sbuf, params);
        tg.setOutputSpp(atoi(argv[3]));
        FW::Image image(sbuf.getWidth(), sbuf.getHeight());
        tg.reconstructDofMotion(image);
        FILE* f = fopen(argv[2], "wb");
        if (!f) return 3;
        fwrite(image.getPtr(), sizeof(float), (size_t)sbuf.getWidth() * sbuf.getHeight() * 4, f);
        fclose(f);
        printf("reconstructed %dx%d, %.3f ms on device\n", sbuf.getWidth(), sbuf.getHeight(),
               tg.lastStats().total_ms);
    } catch (const FW::ReconstructionError& e) {
        fprintf(stderr, "error %d: %s\n", (int)e.status, e.what());
        return 1;
    }
    return 0;
}
