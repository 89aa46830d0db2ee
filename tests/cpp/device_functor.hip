// Device functors compiled into this translation unit (include/felt_amd/device.hpp) stepping
// Felt::Surface, checked bit for bit against the library's own device speeds and against the
// same functor run on the host (Surface.hpp:254-296, 310-386, 1754-1802).  Exit status 0 when
// every check holds; prints the first failure otherwise.
#include <felt_amd/device.hpp>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#define CHECK(c)                                                       \
    do {                                                               \
        if (!(c)) {                                                    \
            std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
            return 1;                                                  \
        }                                                              \
    } while (0)

using namespace Felt;
using S = Surface<3, 3>;
using G = device::Grid<3>;

static const VecDi<3> kSize{48, 40, 36}, kPart{8, 8, 8};

// Two spheres overlapping a little: a surface with curvature of both signs.
static void load(S& s) {
    const double c[6] = {-6.0, 2.5, 1.0, 9.0, -3.0, -2.0};
    const double r[2] = {9.5, 7.25};
    detail::check(felt_surface_load_spheres(s.handle(), 2, c, r));
}

static bool same(const S& a, const S& b) {
    const std::vector<float> x = a.isogrid().snapshot(), y = b.isogrid().snapshot();
    return x.size() == y.size() && std::memcmp(x.data(), y.data(), x.size() * sizeof(float)) == 0;
}

static bool same_layers(const S& a, const S& b) {
    for (int k = -3; k <= 3; k++)
        if (a.layer_size(k) != b.layer_size(k)) return false;
    return true;
}

int main() {
    {
        // fn(pos, grid) = 0.2 curv(pos): identical to the device descriptor b * curv, b = 0.2
        S a(kSize, kPart), b(kSize, kPart);
        load(a);
        load(b);
        for (int step = 0; step < 12; step++) {
            a.update(device_fn([] __device__(const VecDi<3>& p, const G& g) { return 0.2f * g.curv(p); }));
            b.update(speed_curvature(0.2f));
            CHECK(same(a, b));
            CHECK(same_layers(a, b));
        }
        CHECK(a.layer_size(0) > 100);
    }
    {
        // fn() = -0.5 and fn(pos): constant and position-dependent speeds
        S a(kSize, kPart), b(kSize, kPart), c(kSize, kPart), d(kSize, kPart);
        for (S* s : {&a, &b, &c, &d}) load(*s);
        for (int step = 0; step < 6; step++) {
            a.update(device_fn([] __device__() { return -0.5f; }));
            b.update(speed_constant(-0.5f));
            c.update(device_fn([] __device__(const VecDi<3>& p) {
                return p[0] > 0 ? -0.3f : 0.25f * (float)(p[1] % 3) / 3.0f;
            }));
            d.update([](const VecDi<3>& p) {
                return p[0] > 0 ? -0.3f : 0.25f * (float)(p[1] % 3) / 3.0f;
            });
            CHECK(same(a, b));
            CHECK(same(c, d));
        }
    }
    {
        // the grid API a functor calls on the device against the same functor on the host: the
        // upwind gradient, the divergence, the safe gradient at a real position (interp)
        S a(kSize, kPart), b(kSize, kPart);
        load(a);
        load(b);
        auto dev = [] __device__(const VecDi<3>& p, const G& g) {
            const VecDf<3> e = g.gradE(p);
            const float mag = sqrtf(e[0] * e[0] + (e[1] * e[1] + e[2] * e[2]));
            const VecDf<3> q{(float)p[0] + 0.25f, (float)p[1] - 0.5f, (float)p[2]};
            const VecDf<3> n = g.grad(q);
            float v = -0.3f * mag + 0.05f * g.divergence(p) + 0.01f * (n[0] - n[2]) +
                      0.02f * g.get(q);
            return v < -1.0f ? -1.0f : (v > 1.0f ? 1.0f : v);
        };
        auto host = [](const VecDi<3>& p, const S::IsoGrid& g) {
            const VecDf<3> e = g.gradE(p);
            const float mag = std::sqrt(e[0] * e[0] + (e[1] * e[1] + e[2] * e[2]));
            const VecDf<3> q{(float)p[0] + 0.25f, (float)p[1] - 0.5f, (float)p[2]};
            const VecDf<3> n = g.grad(q);
            float v = -0.3f * mag + 0.05f * g.divergence(p) + 0.01f * (n[0] - n[2]) +
                      0.02f * g.get(q);
            return v < -1.0f ? -1.0f : (v > 1.0f ? 1.0f : v);
        };
        for (int step = 0; step < 6; step++) {
            a.update(device_fn(dev));
            b.update(host);
            CHECK(same(a, b));
            CHECK(same_layers(a, b));
        }
    }
    {
        // bounded update(lo, hi, fn) with a device functor against the device descriptor
        S a(kSize, kPart), b(kSize, kPart);
        load(a);
        load(b);
        const VecDi<3> lo{-12, -8, -10}, hi{4, 9, 6};
        for (int step = 0; step < 8; step++) {
            a.update(lo, hi, device_fn([] __device__(const VecDi<3>& p, const G& g) {
                         return -0.3f + 0.1f * g.curv(p);
                     }));
            b.update(lo, hi, speed_const_curv(-0.3f, 0.1f));
            CHECK(same(a, b));
            CHECK(same_layers(a, b));
            CHECK(a.delta(S::layer_idx(0)) == b.delta(S::layer_idx(0)));
        }
    }
    {
        // a device-functor step blocks other calls until it is committed or cancelled, and a
        // cancelled step leaves the surface as it was
        S a(kSize, kPart), b(kSize, kPart);
        load(a);
        load(b);
        felt_zero_job job{};
        detail::check(felt_surface_update_begin(a.handle(), &job));
        CHECK(job.capacity > 0 && job.voxels && job.span && job.delta && job.stream);
        float v = 0.f;
        const int32_t p0[3] = {0, 0, 0};
        CHECK(felt_surface_get(a.handle(), 1, p0, &v) == FELT_E_STATE);
        CHECK(felt_surface_update_cancel(a.handle()) == FELT_OK);
        CHECK(felt_surface_update_commit(a.handle()) == FELT_E_STATE);
        CHECK(same(a, b));
        a.update(speed_constant(-0.4f));
        b.update(device_fn([] __device__() { return -0.4f; }));
        CHECK(same(a, b));
    }
    std::printf("all checks passed\n");
    return 0;
}
