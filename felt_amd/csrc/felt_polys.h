// felt_polys.h — host-side state of the device polygonisation of one surface.
#pragma once

#include <cstdint>
#include <vector>

#include "felt_surface.h"

// Polys children bookkeeping: a lookup-backed list of partitions with swap-with-last untrack
// (Lookup::SingleListSingleIdx, LookupMixin.hpp:225-387).
struct Changes {
    std::vector<int64_t> look;
    std::vector<int64_t> list;
    void track(int64_t p);
    void untrack(int64_t p);
    void reset();
};

// One Poly::Single child: its mesh lives in the vertex / simplex arenas.
struct PolyKid {
    bool active = false;
    long long voff = 0, nv = 0, toff = 0, nt = 0;
};

struct felt_polys {
    felt_surface* s = nullptr;
    std::vector<PolyKid> kids;
    Changes a, b;
    Changes alive;  // children with a mesh (active)
    Changes* pending = nullptr;
    Changes* done = nullptr;
    int vw = 6;
    DBuf<float> varena;
    DBuf<uint32_t> tarena;
    long long vtop = 0, ttop = 0;
    long long last_spx = 0;
    DBuf<uint8_t> cmask;
    DBuf<uint16_t> own;
    DBuf<uint32_t> ocand;  // per band cube: 2 bits per crossed edge, its owner among the sharers
    DBuf<int> vbase, tbase, tot, dparts, pstart, vst, item_j, item_l, nv, nt, vleaf, work, vwork;
    DBuf<long long> dseg;
    DBuf<uint8_t> bits;
    DBuf<int> plist;  // partition lists exchanged with the host
};

int felt_polys_init(felt_polys* q);
void felt_polys_release(felt_polys* q);
int felt_polys_notify_impl(felt_polys* q);
int felt_polys_invalidate_impl(felt_polys* q);
int felt_polys_march_impl(felt_polys* q);
int felt_polys_child_mesh_impl(felt_polys* q, uint64_t p, float* vtx, size_t vcap, uint32_t* spx,
                               size_t scap);
// Device bytes the polygoniser holds: mesh arenas and march scratch.
uint64_t felt_polys_device_bytes(const felt_polys* q);
void felt_mc_table_host(uint64_t* out);
void felt_ms_table_host(uint32_t* out);
