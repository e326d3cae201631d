// Renders a Mitsuba 3 XML scene the way the reference's CLI does
// (cli/main.cpp:225-290): LoadScene, then Filter::Create, Camera::Create,
// Sampler::Create, the BVH, Integrator::Create and Render.
//   render_scene scene.xml output.{pfm,hdr,png,jpg} [spp]
// A positive spp overrides the scene's sample count.
#include "bulbit/bulbit.h"

#include <cstdio>
#include <cstdlib>

int main(int argc, char** argv)
{
    using namespace bulbit;
    if (argc < 3) {
        std::fprintf(stderr, "usage: render_scene scene.xml output [spp]\n");
        return 2;
    }
    RendererInfo ri;
    if (!LoadScene(&ri, argv[1])) return 1;
    if (argc > 3 && std::atoi(argv[3]) > 0) ri.camera_info.sampler_info.spp = std::atoi(argv[3]);

    Allocator alloc;
    Filter* filter = Filter::Create(alloc, ri.camera_info.film_info.filter_info);
    Camera* camera = Camera::Create(alloc, ri.camera_info, filter);
    Sampler* sampler = Sampler::Create(alloc, ri.camera_info.sampler_info);
    if (!filter || !camera || !sampler) {
        std::fprintf(stderr, "render_scene: camera, filter or sampler not supported\n");
        return 1;
    }
    BVH accel(ri.scene.GetPrimitives());
    Integrator* integrator = Integrator::Create(alloc, ri.integrator_info, &accel, ri.scene.GetLights(), sampler);
    if (!integrator) {
        std::fprintf(stderr, "render_scene: integrator not supported\n");
        return 1;
    }
    Rendering* rendering = integrator->Render(alloc, camera);
    rendering->WaitAndLogProgress();
    const bool written = WriteImage(rendering->GetFilm().GetRenderedImage(), argv[2]);
    alloc.delete_object(rendering);
    alloc.delete_object(integrator);
    alloc.delete_object(sampler);
    alloc.delete_object(camera);
    alloc.delete_object(filter);
    return written ? 0 : 1;
}
