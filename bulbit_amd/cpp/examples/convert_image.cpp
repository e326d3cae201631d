// Reads an image with ReadImage3 (PNM, PNG, JPEG, Radiance .hdr, PFM), or one
// channel of it with ReadImage1, and writes it with WriteImage:
//   convert_image input output.{pfm,hdr,png} [non_color] [channel=K]
// With channel=K the channel is written to all three components; an image
// that has no channel K (ReadImage1 returns an empty image) exits with 3.
#include "bulbit/bulbit.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>

int main(int argc, char** argv)
{
    using namespace bulbit;
    if (argc < 3) {
        std::fprintf(stderr, "usage: convert_image input output [non_color] [channel=K]\n");
        return 2;
    }
    bool non_color = false;
    int channel = -1;
    for (int i = 3; i < argc; ++i) {
        if (std::strcmp(argv[i], "non_color") == 0) non_color = true;
        else if (std::strncmp(argv[i], "channel=", 8) == 0) channel = std::atoi(argv[i] + 8);
    }
    if (channel < 0) {
        Image3 image = ReadImage3(argv[1], non_color);
        if (!image) return 1;
        return WriteImage(image, argv[2]) ? 0 : 1;
    }
    Image1 one = ReadImage1(argv[1], channel, non_color);
    if (!one) return 3;
    Image3 image(one.width, one.height);
    for (int32 i = 0; i < one.width * one.height; ++i) image[i] = Spectrum(one[i]);
    return WriteImage(image, argv[2]) ? 0 : 1;
}
