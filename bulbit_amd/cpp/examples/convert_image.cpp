// Reads an image with ReadImage3 (PNM, Radiance .hdr, PFM) and writes it with
// WriteImage:  convert_image input output.{pfm,hdr,png} [non_color]
#include "bulbit/bulbit.h"

#include <cstdio>
#include <cstring>

int main(int argc, char** argv)
{
    using namespace bulbit;
    if (argc < 3) {
        std::fprintf(stderr, "usage: convert_image input output [non_color]\n");
        return 2;
    }
    const bool non_color = argc > 3 && std::strcmp(argv[3], "non_color") == 0;
    Image3 image = ReadImage3(argv[1], non_color);
    if (!image) return 1;
    return WriteImage(image, argv[2]) ? 0 : 1;
}
