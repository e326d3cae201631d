// Pins the order in which g++ evaluates two draws inside one argument list,
// the shape of camera->SampleRay(&ray, pixel, sampler->Next2D(), sampler->Next2D())
// (src/integrator/integrator.cpp:190) and
// bsdf.Sample_f(&s, wo, sampler.Next1D(), sampler.Next2D())
// (src/integrator/path_integrator.cpp:115).  Prints which argument got the
// first draw.  Test infrastructure for the oracle.
#include <cstdio>

struct Point2 { float x, y; };

struct Sampler {
    int n = 0;
    virtual ~Sampler() = default;
    virtual float Next1D() { return float(++n); }
    virtual Point2 Next2D() { float a = float(++n); return Point2{ a, 0.0f }; }
};

struct Camera {
    virtual ~Camera() = default;
    virtual void SampleRay(int* ray, const int& pixel, Point2 u0, Point2 u1) const
    {
        (void)ray; (void)pixel;
        std::printf("camera u0=%g u1=%g\n", u0.x, u1.x);
    }
};

struct BSDF {
    bool Sample_f(int* s, float wo, float u0, Point2 u12) const
    {
        (void)s; (void)wo;
        std::printf("bsdf u0=%g u12=%g\n", u0, u12.x);
        return true;
    }
};

int main()
{
    Sampler s1;
    Sampler* sampler = &s1;
    Camera c;
    const Camera* camera = &c;
    int ray = 0, pixel = 0;
    camera->SampleRay(&ray, pixel, sampler->Next2D(), sampler->Next2D());

    Sampler s2;
    Sampler& sampler2 = s2;
    BSDF bsdf;
    bsdf.Sample_f(&ray, 0.0f, sampler2.Next1D(), sampler2.Next2D());
    return 0;
}
