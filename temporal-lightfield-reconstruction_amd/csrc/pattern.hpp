// Host generation of the output sampling pattern (pattern.cpp).
#pragma once
#include <vector>

namespace tlr {

// Fills out5 with 64*M templates (x, y, u, v, t); returns M (1 when
// override_uvt, else out_spp).
int make_output_pattern(int out_spp, bool override_uvt, const float* uvt, std::vector<float>& out5);

// (sobol(3, i + 1), sobol(4, i + 1)) for i < M: the light-source positions
// of reconstructDofMotionShadows' output samples (Reconstruction.cpp:784).
void light_sample_positions(int M, std::vector<float>& out2);

}  // namespace tlr
