// Mutation test driver for the C++ API's file readers: feeds a file and
// seeded mutations of it (bytes replaced, flipped, deleted, duplicated or
// cut) to ReadImage3 / ReadImage1, LoadModel or LoadScene.  Built with
// AddressSanitizer and UBSan by tests/test_parsers_fuzz.py; any report
// aborts the run.  Reads that fail are expected; only faults are errors.
//   parsers_fuzz image|model|scene FILE ITERATIONS SEED
#include "bulbit/bulbit.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <random>
#include <sstream>
#include <string>

using namespace bulbit;

namespace
{

std::string slurp(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    std::stringstream s;
    s << f.rdbuf();
    return s.str();
}

std::string mutate(const std::string& base, std::mt19937& rng)
{
    std::string m = base;
    static const char kSyntax[] = "0123456789-.eE\"<>/=,{}[]: ";
    const int n = 1 + int(rng() % 6);
    for (int k = 0; k < n && !m.empty(); ++k) {
        const size_t i = rng() % m.size();
        switch (rng() % 6) {
        case 0: m[i] = char(rng()); break;
        case 1: m[i] = char(m[i] ^ char(1u << (rng() % 8))); break;
        case 2: m.resize(i + 1); break;
        case 3: m[i] = kSyntax[rng() % (sizeof kSyntax - 1)]; break;
        case 4: m.erase(i, 1 + rng() % 8); break;
        default: m.insert(i, m.substr(rng() % m.size(), 1 + rng() % 24)); break;
        }
    }
    return m;
}

} // namespace

int main(int argc, char** argv)
{
    if (argc != 5) {
        std::fprintf(stderr, "usage: parsers_fuzz image|model|scene FILE ITERATIONS SEED\n");
        return 2;
    }
    const std::string kind = argv[1], src = argv[2];
    const int iterations = std::atoi(argv[3]);
    std::mt19937 rng(uint32_t(std::atoi(argv[4])));
    const std::string base = slurp(src);
    const size_t slash = src.rfind('/'), dot = src.rfind('.');
    // the mutated file sits beside the original so that relative references resolve
    const std::string tmp = src.substr(0, slash + 1) + "mutated" + src.substr(dot);
    long loaded = 0;
    std::fflush(stdout);
    for (int it = 0; it < iterations; ++it) {
        {
            std::ofstream o(tmp, std::ios::binary);
            o << (it == 0 ? base : mutate(base, rng));
        }
        try {
            if (kind == "image") {
                Image3 image = ReadImage3(tmp);
                Image1 alpha = ReadImage1(tmp, 3, true);
                loaded += image ? 1 : 0;
            } else if (kind == "model") {
                Scene scene;
                ModelLoaderOptions options;
                options.use_fallback_material = true;
                options.fallback_material =
                    scene.CreateMaterial<DiffuseMaterial>(scene.CreateTexture<ConstantTexture, Spectrum>(Spectrum(0.5f)));
                LoadModel(scene, tmp, Transform(Vec3(0, 0, 0)), options);
                loaded += 1;
            } else if (kind == "scene") {
                RendererInfo info;
                loaded += LoadScene(&info, tmp) ? 1 : 0;
            } else {
                return 2;
            }
        } catch (const std::exception&) {
            // a refused file is an outcome, not a fault
        }
    }
    std::printf("%s %s: %d files, %ld read\n", kind.c_str(), src.c_str(), iterations, loaded);
    return 0;
}
