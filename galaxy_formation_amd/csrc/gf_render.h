// Rendering kernels (gf_render.hip).  All pointers are device memory.
#pragma once
#include "gf_common.h"

int gf_render_deposit(int n, const double* pos, const uint8_t* group, int ngroups,
                      const double* rot, const double* center, double scale, int W, int H,
                      float* img, hipStream_t s);
int gf_render_tonemap(int W, int H, int ngroups, const float* img, const float* glow, int gw,
                      int gh, const float* colors, float inv_ref, float K, float exposure,
                      uint8_t* out, hipStream_t s);
