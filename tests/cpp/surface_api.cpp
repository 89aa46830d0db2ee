// C++ use of Felt::Surface / Felt::Polys over the HIP library, written as the reference's own
// tests are (test_Surface.cpp:49-160, 1031-1053; test_Poly.cpp:278-358).  Exit status 0 when
// every check holds; prints the first failure otherwise.
#include <felt_amd/Surface.hpp>
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <set>
#include <sstream>

#define CHECK(c)                                                   \
    do {                                                           \
        if (!(c)) {                                                \
            std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
            return 1;                                              \
        }                                                          \
    } while (0)

using namespace Felt;

int main() {
    {
        // 9x9 2-layer surface, 3x3 partitions: seed then expand by a host lambda of each shape
        Surface<2, 2> surface(VecDi<2>{9, 9}, VecDi<2>{3, 3});
        surface.seed(VecDi<2>{0, 0});
        CHECK(surface.isogrid().get(VecDi<2>{0, 0}) == 0);
        CHECK(surface.layer_size(0) == 1 && surface.layer_size(1) == 4 && surface.layer_size(2) == 8);
        surface.update([](const VecDi<2>&, const Surface<2, 2>::IsoGrid&) { return -1.0f; });
        CHECK(surface.layer_size(-1) == 1 && surface.layer_size(0) == 4);
        surface.update([](const VecDi<2>&) { return 1.0f; });
        CHECK(surface.layer_size(-1) == 0 && surface.layer_size(0) == 1);
        surface.update([]() { return -1.0f; });
        CHECK(surface.layer_size(2) == 12);
        const std::vector<float> snap = surface.isogrid().snapshot();
        CHECK(snap.size() == 81 && snap[4 * 9 + 4] == -1.0f);
        surface.update(VecDi<2>{1, 0}, VecDi<2>{3, 3}, []() { return 0.0f; });
        CHECK(surface.delta(Surface<2, 2>::layer_idx(0)).size() >= 1);
        // the affected set holds the box's zero-layer points and reaches every layer
        const auto aff = surface.affected();
        CHECK(!aff.list(Surface<2, 2>::layer_idx(0)).empty());
        const std::size_t n0 = aff.leafs(Surface<2, 2>::layer_idx(0)).size() / 2;
        std::size_t nall = 0;
        for (int k = 0; k < 5; k++) nall += aff.leafs(k).size() / 2;
        CHECK(n0 >= 1 && nall > n0);
    }
    {
        // A host functor reading the grid's curvature (NumericMixin.hpp:73-158) steps exactly as
        // the device functor of the same speed; the other accessors on the city-block ball
        // |x|+|y|+|z| - 3 that three unit expansions of a seed leave.
        using S = Surface<3, 2>;
        S a(VecDi<3>{24, 24, 24}, VecDi<3>{8, 8, 8}), b(VecDi<3>{24, 24, 24}, VecDi<3>{8, 8, 8});
        a.seed(VecDi<3>{0, 0, 0});
        b.seed(VecDi<3>{0, 0, 0});
        for (int i = 0; i < 3; i++) {
            a.update(speed_constant(-1.0f));
            b.update(speed_constant(-1.0f));
        }
        const S::IsoGrid& g = a.isogrid();
        const VecDi<3> p2{2, 0, 0};
        CHECK(g.get(p2) == -1.0f && g.get(VecDi<3>{1, 0, 0}) == -2.0f);
        const VecDf<3> gc = g.gradC(p2), gf = g.gradF(p2), gb = g.gradB(p2), gs = g.grad(p2);
        CHECK(gc[0] == 1.0f && gc[1] == 0.0f && gc[2] == 0.0f);
        CHECK(gf[0] == 1.0f && gf[1] == 1.0f && gb[0] == 1.0f && gb[1] == -1.0f);
        CHECK(gs[0] == gc[0] && gs[1] == gc[1] && gs[2] == gc[2]);
        CHECK(g.divergence(p2) == 4.0f);
        const VecDf<3> ge = g.gradE(p2);  // upwind: max(c - back, 0) + min(fwd - c, 0)
        CHECK(ge[0] == 1.0f && ge[1] == 0.0f);
        CHECK(g.interp(VecDf<3>{2.0f, 0.0f, 0.0f}) == -1.0f);
        CHECK(std::fabs(g.interp(VecDf<3>{1.5f, 0.0f, 0.0f}) + 1.5f) < 1e-6f);
        CHECK(std::fabs(g.get(VecDf<3>{1.5f, 0.25f, 0.0f}) + 1.25f) < 1e-6f);
        CHECK(g.inside(VecDi<3>{-12, 11, 0}) && !g.inside(VecDi<3>{12, 0, 0}));
        std::vector<VecDi<3>> nb;
        S::IsoGrid::neighs(p2, [&](const VecDi<3>& q) { nb.push_back(q); });
        CHECK(nb.size() == 6 && nb[0] == (VecDi<3>{1, 0, 0}) && nb[1] == (VecDi<3>{3, 0, 0}) &&
              nb[4] == (VecDi<3>{2, 0, -1}));
        a.update([](const VecDi<3>& p, const S::IsoGrid& grid) { return 0.2f * grid.curv(p); });
        b.update(speed_curvature(0.2f));
        CHECK(a.isogrid().snapshot() == b.isogrid().snapshot());
        // the same through the bounded form
        const VecDi<3> lo{-4, -4, -2}, hi{2, 3, 4};
        a.update(lo, hi, [](const VecDi<3>& p, const S::IsoGrid& grid) { return 0.3f * grid.curv(p); });
        b.update(lo, hi, speed_curvature(0.3f));
        CHECK(a.isogrid().snapshot() == b.isogrid().snapshot());
        CHECK(a.isogrid().get(p2) == b.isogrid().get(p2));  // reads the device again
    }
    {
        // 3D layer counts 1/6/18/38/66 (test_Surface.cpp:1031-1053), device functor
        Surface<3, 3> surface(VecDi<3>{11, 11, 11}, VecDi<3>{3, 3, 3});
        surface.seed(VecDi<3>{0, 0, 0});
        surface.update(speed_constant(-1.0f));
        CHECK(surface.layer_size(-1) == 1 && surface.layer_size(0) == 6 &&
              surface.layer_size(1) == 18 && surface.layer_size(2) == 38 &&
              surface.layer_size(3) == 66);
        surface.update_start();
        surface.delta(VecDi<3>{0, 1, 0}, -1.0f);
        surface.update_end();
        bool threw = false;
        try {
            surface.update_start();
            surface.delta(VecDi<3>{5, 5, 5}, -0.5f);  // not a zero-layer point
            surface.update_end();
        } catch (const std::domain_error&) {
            threw = true;
        }
        CHECK(threw);
    }
    {
        // Polys counts in the central partition (test_Poly.cpp:329-358)
        Surface<3, 3> surface(VecDi<3>{15, 15, 15}, VecDi<3>{5, 5, 5});
        Polys<Surface<3, 3>> polys(surface);
        surface.seed(VecDi<3>{0, 0, 0});
        surface.update(speed_constant(-1.0f));
        polys.notify();
        polys.march();
        const PosIdx centre = (1 * 3 + 1) * 3 + 1;
        const auto child = polys.children().get(centre);
        CHECK(child.vtxs().size() == 30 && child.spxs().size() == 56);
        CHECK(polys.changes().size() == 7);
        for (const auto& v : child.vtxs()) {
            const float n = std::sqrt(v.norm[0] * v.norm[0] + v.norm[1] * v.norm[1] +
                                      v.norm[2] * v.norm[2]);
            CHECK(std::fabs(n - 1.0f) < 1e-5f);
        }
    }
    {
        // raycasting (test_Surface.cpp:1278-1330): radius-3 surface, rays along +x
        Surface<3, 3> surface(VecDi<3>{32, 32, 32}, VecDi<3>{5, 5, 5});
        surface.seed(VecDi<3>{0, 0, 0});
        for (int i = 0; i < 3; i++) surface.update([]() { return -1.0f; });
        const VecDf<3> hit = surface.ray(VecDf<3>{-35.0f, 0, 0}, VecDf<3>{1, 0, 0});
        CHECK(std::fabs(hit[0] + 3.0f) < 1e-4f && std::fabs(hit[1]) < 1e-4f &&
              std::fabs(hit[2]) < 1e-4f);
        CHECK(surface.ray(VecDf<3>{0, 0, 0}, VecDf<3>{1, 0, 0}) == surface.ray_miss);
        const auto hits = surface.ray({VecDf<3>{-6.0f, 0, 0}, VecDf<3>{0, 0, 0}},
                                      {VecDf<3>{1, 0, 0}, VecDf<3>{1, 0, 0}});
        CHECK(hits.size() == 2 && std::fabs(hits[0][0] + 3.0f) < 1e-4f && hits[1] == surface.ray_miss);
        // the same rays from device memory (felt_surface_ray_device)
        const float org[6] = {-6.0f, 0, 0, 0, 0, 0}, dir[6] = {1, 0, 0, 1, 0, 0};
        float *d_org = nullptr, *d_dir = nullptr, *d_out = nullptr, out[6] = {};
        CHECK(hipMalloc(&d_org, sizeof(org)) == hipSuccess);
        CHECK(hipMalloc(&d_dir, sizeof(dir)) == hipSuccess);
        CHECK(hipMalloc(&d_out, sizeof(out)) == hipSuccess);
        CHECK(hipMemcpy(d_org, org, sizeof(org), hipMemcpyHostToDevice) == hipSuccess);
        CHECK(hipMemcpy(d_dir, dir, sizeof(dir), hipMemcpyHostToDevice) == hipSuccess);
        CHECK(felt_surface_ray_device(surface.handle(), 2, d_org, d_dir, d_out) == FELT_OK);
        CHECK(hipMemcpy(out, d_out, sizeof(out), hipMemcpyDeviceToHost) == hipSuccess);
        for (int a = 0; a < 3; a++) {
            CHECK(out[a] == hits[0][a]);
            CHECK(out[3 + a] == surface.ray_miss[a]);
        }
        (void)hipFree(d_org);
        (void)hipFree(d_dir);
        (void)hipFree(d_out);
    }
    {
        // save / load round trip (test_Surface.cpp:464-509): expand by 0.6, save, load, then the
        // loaded surface contracts by 0.6 to layer sizes 0/0/1/4/8
        Surface<2, 2> surface(VecDi<2>{9, 9}, VecDi<2>{3, 3});
        surface.seed(VecDi<2>{0, 0});
        surface.update([]() { return -0.6f; });
        std::stringstream ss;
        surface.save(ss);
        Surface<2, 2> loaded{Surface<2, 2>::load(ss)};
        CHECK(loaded.isogrid().snapshot() == surface.isogrid().snapshot());
        loaded.update([](const VecDi<2>&, const Surface<2, 2>::IsoGrid&) { return 0.6f; });
        CHECK(loaded.layer_size(-2) == 0 && loaded.layer_size(-1) == 0 && loaded.layer_size(0) == 1 &&
              loaded.layer_size(1) == 4 && loaded.layer_size(2) == 8);
        bool threw = false;
        try {
            std::stringstream s3;
            surface.save(s3);
            Surface<3, 2>::load(s3);  // a 2D archive is not a Surface<3, 2>
        } catch (const std::domain_error&) {
            threw = true;
        }
        CHECK(threw);
    }
    std::printf("surface_api: all checks passed\n");
    {
        // Partition geometry, as test_Grid.cpp:1188-1201 and 1205-1215 state it for a 9x9x9 grid
        // in 3x3x3 partitions: child (-1,-1,-1) at offset (-4,-4,-4), child (1,1,1) at (2,2,2),
        // and the partitions holding (1,-4,-1), (2,-3,-2), (3,-2,-3), (4,-1,-4).
        using S = Surface<3, 2>;
        S s(VecDi<3>{9, 9, 9}, VecDi<3>{3, 3, 3});
        const S::IsoGrid& g = s.isogrid();
        CHECK(g.child_size() == (VecDi<3>{3, 3, 3}));
        CHECK(g.children().size() == (VecDi<3>{3, 3, 3}));
        CHECK(g.children().offset() == (VecDi<3>{-1, -1, -1}));
        CHECK(g.children().get(VecDi<3>{-1, -1, -1}).offset() == (VecDi<3>{-4, -4, -4}));
        CHECK(g.children().get(VecDi<3>{1, 1, 1}).offset() == (VecDi<3>{2, 2, 2}));
        CHECK(g.children().get(VecDi<3>{1, 1, 1}).size() == (VecDi<3>{3, 3, 3}));
        CHECK(g.pos_child(VecDi<3>{1, -4, -1}) == (VecDi<3>{0, -1, 0}));
        CHECK(g.pos_child(VecDi<3>{2, -3, -2}) == (VecDi<3>{1, -1, -1}));
        CHECK(g.pos_child(VecDi<3>{3, -2, -3}) == (VecDi<3>{1, -1, -1}));
        CHECK(g.pos_child(VecDi<3>{4, -1, -4}) == (VecDi<3>{1, 0, -1}));
        CHECK(g.pos_idx_child(VecDi<3>{4, -1, -4}) == g.children().index(VecDi<3>{1, 0, -1}));
        CHECK(g.children().index(g.children().index(PosIdx(17))) == 17);
        // the seed's partition is the only active one and holds the zero layer's one voxel
        s.seed(VecDi<3>{0, 0, 0});
        const PosIdx centre = g.pos_idx_child(VecDi<3>{0, 0, 0});
        CHECK(centre == 13);
        const auto child = g.children().get(centre);
        CHECK(child.is_active() && !g.children().get(PosIdx(0)).is_active());
        CHECK(child.list(S::layer_idx(0)).size() == 1);
        CHECK(child.index(child.list(S::layer_idx(0))[0]) == (VecDi<3>{0, 0, 0}));
        CHECK(g.children().lookup().list(S::layer_idx(0)) == PosIdxList{centre});
        CHECK(g.children().lookup().is_tracked(centre, S::layer_idx(1)));
        // leafs(k, fn) visits each voxel of the layer once, partition by partition
        for (int k = 0; k < 5; k++) {
            std::set<std::array<int, 3>> seen;
            PosIdx last = 0;
            std::size_t n = 0, runs = 0;
            g.leafs(TupleIdx(k), [&](const VecDi<3>& p) {
                seen.insert({p[0], p[1], p[2]});
                const PosIdx pc = g.pos_idx_child(p);
                if (n == 0 || pc != last) runs++;
                last = pc;
                n++;
            });
            CHECK(n == seen.size() && n == s.layer_size(k - 2));
            CHECK(runs == g.children().lookup().list(TupleIdx(k)).size());
        }
    }
    {
        // test_Surface.cpp:35-45: a 7x7 isogrid in 3x3 partitions has 9 children of 3x3
        Surface<2, 2> s(VecDi<2>{7, 7}, VecDi<2>{3, 3});
        CHECK(s.isogrid().num_children() == 9);
        CHECK(s.isogrid().children().get(VecDi<2>{0, 0}).size() == (VecDi<2>{3, 3}));
        CHECK(s.isogrid().get(VecDi<2>{0, 0}) == 3);
    }
    return 0;
}
