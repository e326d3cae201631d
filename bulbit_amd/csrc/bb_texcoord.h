// bb_texcoord.h — texel index arithmetic of ImageTexture, shared by the
// kernels (bb_shade.h) and host-side unit tests (compiled with g++).
//
//   include/bulbit/texture.h:26-46    FilterTexCoord (repeat, clamp)
//   include/bulbit/textures.h:60-88   ImageTexture::Evaluate (bilinear)
#pragma once

#include <math.h>
#include <stdint.h>

#include "../../include/bbgpu.h"

#if defined(__HIPCC__)
#define BB_TC __host__ __device__ inline
#else
#define BB_TC inline
#endif

namespace bb
{

// FilterTexCoord (include/bulbit/texture.h:26-46); repeat maps -k to
// n - (k % n) - 1, as the reference writes it
BB_TC int32_t filter_coord(int32_t i, int32_t n, int32_t filter)
{
    if (filter == BB_TEXCOORD_CLAMP) return i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
    return i >= 0 ? i % n : n - ((-i) % n) - 1;
}

// The integer texel coordinate floor(w) that ImageTexture::Evaluate converts
// to int32 before FilterTexCoord.  For |floor(w)| < 2^30 this is the plain
// conversion.  Beyond that range the conversion, or the +1 of the second
// bilinear tap, would overflow int32 (undefined in the reference, and an
// out-of-bounds texel read on the GPU), so far coordinates are reduced first
// to an index that FilterTexCoord maps to the same texel as the exact
// integer would:
//   repeat: a representative of floor(w) mod n (fmodf is exact) with the
//           sign of floor(w), and for negative floor(w) at most -n, so that
//           the second tap (+1) is negative too: FilterTexCoord's repeat is
//           periodic in n on each side of 0 but not across it
//   clamp:  -1 below the image, n above it
// NaN, and infinities under repeat, have no texel; they read texel 0.
BB_TC int32_t tex_floor_index(float w, int32_t n, int32_t filter)
{
    const float f = floorf(w);
    if (fabsf(f) < 1073741824.0f) return (int32_t)f;
    if (f != f) return 0;
    if (filter == BB_TEXCOORD_CLAMP) return f < 0.0f ? -1 : n;
    if (isinf(f)) return 0;
    const int32_t r = (int32_t)fmodf(f, (float)n);
    return f < 0.0f ? r - n : r;
}

} // namespace bb
