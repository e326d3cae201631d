// Host build of the kernels' texel index arithmetic (bulbit_amd/csrc/bb_texcoord.h):
// reads "w n filter" lines from stdin, prints "i0 i1 t0 t1" per line, where
// i0 = tex_floor_index(w), i1 = i0 + 1 and t0, t1 the texels filter_coord
// maps them to.
#include <cstdio>

#include "../bulbit_amd/csrc/bb_texcoord.h"

int main()
{
    unsigned int bits;
    int n, filter;
    while (std::scanf("%x %d %d", &bits, &n, &filter) == 3) {
        float w;
        static_assert(sizeof(w) == sizeof(bits), "float is 32 bits");
        __builtin_memcpy(&w, &bits, sizeof(w));
        const int i0 = bb::tex_floor_index(w, n, filter), i1 = i0 + 1;
        std::printf("%d %d %d %d\n", i0, i1, bb::filter_coord(i0, n, filter), bb::filter_coord(i1, n, filter));
    }
    return 0;
}
