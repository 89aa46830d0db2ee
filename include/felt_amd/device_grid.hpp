// device_grid.hpp — Felt's isogrid API in device code, for HIP translation units (hipcc, gfx950).
//
// Felt::device::Grid<D> wraps a felt_device_grid (felt_surface_device_grid) and offers the
// members a speed functor calls on the reference's isogrid (Partitioned.hpp:160-189;
// NumericMixin.hpp:72-500): get at integer and real positions, interp, curv, divergence, grad,
// gradE, gradF, gradB, gradC, dx, inside, size, offset and neighs.  Every operator follows the
// reference's operation order in fp32; translation units that include it must be compiled with
// -ffp-contract=off for the results to round as the reference's do.  The reference's Eigen
// vector sums are evaluated as v0 + (v1 + v2) in 3D.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../felt_amd.h"
#include "types.hpp"

namespace Felt {
namespace device {

template <int D>
struct Grid {
    felt_device_grid g;

    __host__ __device__ VecDi<D> size() const {
        VecDi<D> v{};
        for (int a = 0; a < D; a++) v[a] = g.size[a];
        return v;
    }
    __host__ __device__ VecDi<D> offset() const {
        VecDi<D> v{};
        for (int a = 0; a < D; a++) v[a] = g.offset[a];
        return v;
    }
    // Grid spacing the derivatives divide by (NumericMixin.hpp:354-370).
    __device__ Distance dx() const { return g.dx; }

    // Half-open bounds, compared as float (Util.hpp:94-106).
    template <class V>
    __device__ bool inside(const V& pos) const {
        for (int a = 0; a < D; a++) {
            if ((float)pos[a] >= (float)(g.offset[a] + g.size[a])) return false;
            if ((float)pos[a] < (float)g.offset[a]) return false;
        }
        return true;
    }

    // get(pos) of the partitioned isogrid (PartitionedMixin.hpp:594-601): the grid background
    // L+1 outside the grid, the partition's background when it is inactive.
    __device__ Distance get(const VecDi<D>& pos) const {
        if (!inside(pos)) return (Distance)(g.L + 1);
        int p = g.pbase, lc[3] = {0, 0, 0};
        for (int a = 0; a < D; a++) {
            const int q = pos[a] - g.offset[a];
            const int c = g.part_shift[a] >= 0 ? q >> g.part_shift[a] : q / g.part[a];
            p += c * g.cstride[a];
            lc[a] = q - c * g.part[a];
        }
        const int s = g.page[p];
        if (s < 0) return g.pbg[p];
        return g.phi[(size_t)s * g.pv + leaf(lc)];
    }
    // A real position reads through the N-linear interpolation (NumericMixin.hpp:390-393).
    __device__ Distance get(const VecDf<D>& pos) const { return interp(pos); }

    // N-linear interpolation (NumericMixin.hpp:403-481): corner i takes +1 along axis a when bit
    // a of i is set; axis 0 is reduced first, each pair as t*hi + (1-t)*lo.
    __device__ Distance interp(const VecDf<D>& pos) const {
        VecDi<D> fl{};
        float t[D];
        for (int a = 0; a < D; a++) {
            const float f = __builtin_floorf(pos[a]);
            fl[a] = (int)f;
            t[a] = pos[a] - f;
        }
        float val[1 << D];
        for (int i = 0; i < (1 << D); i++) {
            VecDi<D> c = fl;
            for (int a = 0; a < D; a++) c[a] += (i >> a) & 1;
            val[i] = get(c);
        }
        int n = 1 << D;
        for (int a = 0; a < D; a++) {
            for (int i = 0; i < n / 2; i++)
                val[i] = t[a] * val[2 * i + 1] + (1.0f - t[a]) * val[2 * i];
            n /= 2;
        }
        return val[0];
    }

    // Mean curvature by the difference of normals (NumericMixin.hpp:72-156).
    template <class V>
    __device__ Distance curv(const V& pos) const {
        const float c = get(pos);
        float nf[D], nb[D];
        V dir = pos;
        for (int a = 0; a < D; a++) {
            dir[a] += 1;
            const float va = get(dir) - c;
            float sq = 0.0f;
            for (int b = 0; b < D; b++) {
                if (b == a) continue;
                V dn = dir;
                dn[b] -= 1;
                const float lo = get(dn);
                dn[b] += 2;
                const float hi = get(dn);
                const float t = (hi - lo) / 2.0f;
                sq += t * t;
            }
            nf[a] = va / __builtin_sqrtf(va * va + sq);
            dir[a] -= 1;
        }
        for (int a = 0; a < D; a++) {
            dir[a] -= 1;
            const float va = c - get(dir);
            float sq = 0.0f;
            for (int b = 0; b < D; b++) {
                if (b == a) continue;
                V dn = dir;
                dn[b] -= 1;
                const float lo = get(dn);
                dn[b] += 2;
                const float hi = get(dn);
                const float t = (hi - lo) / 2.0f;
                sq += t * t;
            }
            nb[a] = va / __builtin_sqrtf(va * va + sq);
            dir[a] += 1;
        }
        float d[D];
        for (int a = 0; a < D; a++) d[a] = nf[a] - nb[a];
        return sum(d) / 2.0f;
    }

    // Divergence of the gradient: sum(gradF - gradB) / dx^2 (NumericMixin.hpp:164-175).
    template <class V>
    __device__ Distance divergence(const V& pos) const {
        const VecDf<D> f = gradF(pos), b = gradB(pos);
        float d[D];
        for (int a = 0; a < D; a++) d[a] = f[a] - b[a];
        return sum(d) / (g.dx * g.dx);
    }

    // Safe gradient: central difference where both neighbours are inside the grid, one-sided
    // where one is, 0 where none is (NumericMixin.hpp:188-231).
    template <class V>
    __device__ VecDf<D> grad(const V& pos) const {
        VecDf<D> gv{};
        V pt = pos;
        const float centre = get(pos);
        for (int a = 0; a < D; a++) {
            float back = centre, fwd = centre;
            int order = 0;
            pt[a] -= 1;
            if (inside(pt)) {
                back = get(pt);
                order++;
            }
            pt[a] += 2;
            if (inside(pt)) {
                fwd = get(pt);
                order++;
            }
            pt[a] -= 1;
            gv[a] = order != 0 ? (fwd - back) / (float)order : 0.0f;
        }
        for (int a = 0; a < D; a++) gv[a] = gv[a] / g.dx;
        return gv;
    }

    // Upwind (entropy satisfying) gradient (NumericMixin.hpp:242-273).
    template <class V>
    __device__ VecDf<D> gradE(const V& pos) const {
        VecDf<D> gv{};
        V pt = pos;
        const float centre = get(pos);
        for (int a = 0; a < D; a++) {
            pt[a] -= 1;
            float back = get(pt);
            pt[a] += 2;
            float fwd = get(pt);
            pt[a] -= 1;
            // std::max(x, 0.0f) and std::min(x, 0.0f), signed zeros and NaN included
            const float xb = centre - back, xf = fwd - centre;
            back = xb < 0.0f ? 0.0f : xb;
            fwd = 0.0f < xf ? 0.0f : xf;
            gv[a] = fwd + back;
        }
        for (int a = 0; a < D; a++) gv[a] = gv[a] / g.dx;
        return gv;
    }

    // Forward, backward and central differences (NumericMixin.hpp:279-358).
    template <class V>
    __device__ VecDf<D> gradF(const V& pos) const {
        VecDf<D> gv{};
        V pt = pos;
        const float centre = get(pos);
        for (int a = 0; a < D; a++) {
            pt[a] += 1;
            gv[a] = get(pt) - centre;
            pt[a] -= 1;
        }
        for (int a = 0; a < D; a++) gv[a] = gv[a] / g.dx;
        return gv;
    }
    template <class V>
    __device__ VecDf<D> gradB(const V& pos) const {
        VecDf<D> gv{};
        V pt = pos;
        const float centre = get(pos);
        for (int a = 0; a < D; a++) {
            pt[a] -= 1;
            gv[a] = centre - get(pt);
            pt[a] += 1;
        }
        for (int a = 0; a < D; a++) gv[a] = gv[a] / g.dx;
        return gv;
    }
    template <class V>
    __device__ VecDf<D> gradC(const V& pos) const {
        VecDf<D> gv{};
        V pt = pos;
        for (int a = 0; a < D; a++) {
            pt[a] -= 1;
            const float back = get(pt);
            pt[a] += 2;
            const float fwd = get(pt);
            pt[a] -= 1;
            gv[a] = (fwd - back) / 2.0f;
        }
        for (int a = 0; a < D; a++) gv[a] = gv[a] / g.dx;
        return gv;
    }

    // Face neighbours, -1 then +1 along each axis in turn (NumericMixin.hpp:489-500).
    template <class Fn>
    __device__ static void neighs(VecDi<D> pos, Fn&& fn) {
        for (int a = 0; a < D; a++) {
            pos[a] -= 1;
            fn(pos);
            pos[a] += 2;
            fn(pos);
            pos[a] -= 1;
        }
    }

    // World position of pool voxel v (a felt_zero_job entry).
    __device__ VecDi<D> position(uint32_t v) const {
        const uint32_t slot = v / g.pv;
        uint32_t l = v - slot * g.pv;
        uint32_t w = l & ((1u << g.tile_vshift) - 1u), t = l >> g.tile_vshift;
        const int* org = g.slot_org + 4 * (size_t)slot;
        VecDi<D> pos{};
        for (int a = D - 1; a >= 0; a--) {
            const uint32_t in = w & ((1u << g.tile_shift[a]) - 1u);
            w >>= g.tile_shift[a];
            const uint32_t q = t / (uint32_t)g.tiles[a];
            const uint32_t lc = ((t - q * (uint32_t)g.tiles[a]) << g.tile_shift[a]) | in;
            t = q;
            pos[a] = org[a] + (int)lc + g.offset[a];
        }
        return pos;
    }

private:
    __device__ int leaf(const int lc[3]) const {
        int l = 0;
        for (int a = 0; a < D; a++)
            l += (lc[a] >> g.tile_shift[a]) * g.tile_stride[a] +
                 (lc[a] & ((1 << g.tile_shift[a]) - 1)) * g.in_stride[a];
        return l;
    }
    __device__ static float sum(const float* v) {
        if (D == 3) return v[0] + (v[1] + v[D - 1]);
        return v[0] + v[1];
    }
};

}  // namespace device
}  // namespace Felt
