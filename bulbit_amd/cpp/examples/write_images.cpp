// write_images: writes a synthetic linear image through the C++ API's
// WriteImage in every supported format, for the image-output tests.
//   write_images <prefix>  ->  <prefix>.png .jpg .hdr .pfm
// Pixel (x, y) of the W x H image (row 0 is the bottom of the picture) holds
// r = x / (W - 1) * 1.2 (values above 1 clip in 8-bit output),
// g = y / (H - 1), b = 0.01 * ((x * 7 + y * 13) % 11) (near-black values,
// some below the sRGB curve's linear segment).
#include "bulbit/bulbit.h"

#include <cstdio>
#include <string>

int main(int argc, char** argv)
{
    if (argc != 2) {
        std::fprintf(stderr, "usage: write_images <prefix>\n");
        return 2;
    }
    const int W = 37, H = 21;
    bulbit::Image3 img(W, H);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
            img(x, y) = bulbit::Spectrum(float(x) / float(W - 1) * 1.2f, float(y) / float(H - 1),
                                         0.01f * float((x * 7 + y * 13) % 11));
    const std::string p = argv[1];
    bool ok = true;
    for (const char* ext : { ".png", ".jpg", ".hdr", ".pfm" }) ok = bulbit::WriteImage(img, p + ext) && ok;
    return ok ? 0 : 1;
}
