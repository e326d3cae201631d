// The library example of Bulbit's README (README.md:136-190), compiled
// against bulbit/bulbit.h.  The body is the README's; only the resolution,
// sample count and output file come from the command line:
//   readme_example [width height spp output.{pfm,hdr}]
// (defaults: 1280 720 16 render.pfm).
#include "bulbit/bulbit.h"

#include <cstdlib>
#include <string>

int main(int argc, char** argv)
{
    using namespace bulbit;

    const int32 width = argc > 1 ? std::atoi(argv[1]) : 1280;
    const int32 height = argc > 2 ? std::atoi(argv[2]) : 720;
    const int32 samples = argc > 3 ? std::atoi(argv[3]) : 16;
    const std::string output = argc > 4 ? argv[4] : "render.pfm";

    uint32 num_threads = std::thread::hardware_concurrency();
    ThreadPool::global_thread_pool = std::make_unique<ThreadPool>(num_threads);

    Scene scene;

    // Create texture, material and sphere primitive
    Texture<Spectrum>* tex = scene.CreateTexture<ConstantTexture, Spectrum>(Spectrum(0.85f, 0.2f, 0.2f));
    Material* mat = scene.CreateMaterial<DiffuseMaterial>(tex);
    Point3 sphere_pos(0.0f, 1.0f, 0.0f);
    Shape* sphere = scene.CreateShape<Sphere>(sphere_pos, 1.0f);
    scene.CreatePrimitive(sphere, mat, MediumInterface{});

    // Create two lights
    Point3 light_pos(2.5f, 4.0f, 2.0f);
    Vec3 light_dir = Normalize(light_pos - sphere_pos);
    scene.CreateLight<SpotLight>(light_pos, light_dir, Spectrum(80.0f), 35.0f, 20.0f, nullptr);
    scene.CreateLight<UniformInfiniteLight>(Spectrum(0.1f));

    // Create camera
    Point3 camera_pos(3.0f, 2.0f, 5.0f);
    Point2i resolution(width, height);
    Float fov = 35;
    Float aperture = 0.01f;

    BoxFilter filter;
    PerspectiveCamera camera(
        Transform::LookAt(camera_pos, sphere_pos, y_axis), fov, aperture, Dist(camera_pos, sphere_pos), resolution,
        nullptr, &filter
    );

    // Build acceleration structure
    BVH accel(scene.GetPrimitives());

    // Create integrator
    int spp = samples;
    int max_bounces = 8;
    IndependentSampler sampler(spp);
    PathIntegrator integrator(&accel, scene.GetLights(), &sampler, max_bounces);

    // Render!
    Allocator alloc;
    Rendering* rendering = integrator.Render(alloc, &camera);
    rendering->WaitAndLogProgress();
    bool written = WriteImage(rendering->GetFilm().GetRenderedImage(), output);
    alloc.delete_object(rendering);

    return written ? 0 : 1;
}
