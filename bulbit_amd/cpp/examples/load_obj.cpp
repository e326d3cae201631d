// Renders an OBJ model with LoadModel (the CLI's model loader) under a
// uniform environment light of 0.2:
//   load_obj model.obj width height spp output.{pfm,hdr,png} [ty]
// The model is moved up by ty (default 0) through the loader's Transform;
// the camera sits at (0, 2, 6) looking at (0, 0.5, 0), vfov 40, box filter.
#include "bulbit/bulbit.h"

#include <cstdio>
#include <cstdlib>
#include <string>

int main(int argc, char** argv)
{
    using namespace bulbit;
    if (argc < 6) {
        std::fprintf(stderr, "usage: load_obj model.obj width height spp output [ty]\n");
        return 2;
    }
    const int32 width = std::atoi(argv[2]), height = std::atoi(argv[3]), spp = std::atoi(argv[4]);
    const Float ty = argc > 6 ? Float(std::atof(argv[6])) : Float(0);

    Scene scene;
    LoadModel(scene, argv[1], Transform(Vec3(0, ty, 0)));
    scene.CreateLight<UniformInfiniteLight>(Spectrum(0.2f));

    Point3 eye(0, 2, 6), target(0, 0.5f, 0);
    BoxFilter filter;
    PerspectiveCamera camera(Transform::LookAt(eye, target, y_axis), 40.0f, 0.0f, Dist(eye, target),
                             Point2i(width, height), nullptr, &filter);
    BVH accel(scene.GetPrimitives());
    IndependentSampler sampler(spp);
    PathIntegrator integrator(&accel, scene.GetLights(), &sampler, 8);
    Allocator alloc;
    Rendering* rendering = integrator.Render(alloc, &camera);
    rendering->WaitAndLogProgress();
    const bool written = WriteImage(rendering->GetFilm().GetRenderedImage(), argv[5]);
    alloc.delete_object(rendering);
    return written ? 0 : 1;
}
