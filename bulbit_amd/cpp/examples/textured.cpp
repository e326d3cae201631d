// The README scene with textures, written against bulbit/bulbit.h: the
// sphere's reflectance is an ImageTexture and it carries a normal map; a
// ground quad with texture coordinates has a CheckerTexture reflectance.
//   textured_example [width height spp output.{pfm,hdr,png}]
// (defaults: 640 360 16 textured.pfm).
#include "bulbit/bulbit.h"

#include <cstdlib>
#include <string>

int main(int argc, char** argv)
{
    using namespace bulbit;

    const int32 width = argc > 1 ? std::atoi(argv[1]) : 640;
    const int32 height = argc > 2 ? std::atoi(argv[2]) : 360;
    const int32 samples = argc > 3 ? std::atoi(argv[3]) : 16;
    const std::string output = argc > 4 ? argv[4] : "textured.pfm";

    Scene scene;

    // 16 x 8 reflectance image and an 8 x 8 normal map of tilted stripes
    Image3 albedo(16, 8);
    for (int32 y = 0; y < 8; ++y)
        for (int32 x = 0; x < 16; ++x) albedo(x, y) = Spectrum(0.2f + 0.04f * x, 0.2f + 0.08f * y, 0.5f);
    Image3 bumps(8, 8);
    for (int32 y = 0; y < 8; ++y)
        for (int32 x = 0; x < 8; ++x) bumps(x, y) = Spectrum((x % 2) ? 0.7f : 0.3f, 0.5f, 1.0f);
    Texture<Spectrum>* reflectance = scene.CreateTexture<ImageTexture, Spectrum>(albedo);
    Texture<Spectrum>* normal = scene.CreateTexture<ImageTexture, Spectrum>(bumps);
    Material* sphere_mat = scene.CreateMaterial<DiffuseMaterial>(reflectance, nullptr, normal);
    Point3 sphere_pos(0.0f, 1.0f, 0.0f);
    Shape* sphere = scene.CreateShape<Sphere>(sphere_pos, 1.0f);
    scene.CreatePrimitive(sphere, sphere_mat, MediumInterface{});

    // checker ground: 8 x 8 squares over [-4, 4]^2 at y = 0
    Texture<Spectrum>* light_sq = scene.CreateTexture<ConstantTexture, Spectrum>(Spectrum(0.8f, 0.8f, 0.8f));
    Texture<Spectrum>* dark_sq = scene.CreateTexture<ConstantTexture, Spectrum>(Spectrum(0.1f, 0.1f, 0.15f));
    Texture<Spectrum>* checker = scene.CreateTexture<CheckerTexture, Spectrum>(light_sq, dark_sq, Point2(8.0f, 8.0f));
    Material* ground_mat = scene.CreateMaterial<DiffuseMaterial>(checker);
    std::vector<Point3> p = { Point3(-4, 0, -4), Point3(4, 0, -4), Point3(4, 0, 4), Point3(-4, 0, 4) };
    std::vector<Vec3> n(4, Vec3(0, 1, 0)), t(4, Vec3(1, 0, 0));
    std::vector<Point2> uv = { Point2(0, 0), Point2(1, 0), Point2(1, 1), Point2(0, 1) };
    std::vector<int32> idx = { 0, 2, 1, 0, 3, 2 };
    Mesh* ground = scene.CreateMesh(p, n, t, uv, idx, Mat4(identity));
    for (int32 i = 0; i < ground->GetTriangleCount(); ++i)
        scene.CreatePrimitive(scene.CreateShape<Triangle>(ground, i), ground_mat, MediumInterface{});

    Point3 light_pos(2.5f, 4.0f, 2.0f);
    Vec3 light_dir = Normalize(light_pos - sphere_pos);
    scene.CreateLight<SpotLight>(light_pos, light_dir, Spectrum(80.0f), 35.0f, 20.0f, nullptr);
    scene.CreateLight<UniformInfiniteLight>(Spectrum(0.1f));

    Point3 camera_pos(3.0f, 2.0f, 5.0f);
    BoxFilter filter;
    PerspectiveCamera camera(Transform::LookAt(camera_pos, sphere_pos, y_axis), 35.0f, 0.01f,
                             Dist(camera_pos, sphere_pos), Point2i(width, height), nullptr, &filter);

    BVH accel(scene.GetPrimitives());
    IndependentSampler sampler(samples);
    PathIntegrator integrator(&accel, scene.GetLights(), &sampler, 8);

    Allocator alloc;
    Rendering* rendering = integrator.Render(alloc, &camera);
    rendering->WaitAndLogProgress();
    bool written = WriteImage(rendering->GetFilm().GetRenderedImage(), output);
    alloc.delete_object(rendering);
    return written ? 0 : 1;
}
