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

// One Poly::Single child: its mesh lives in a segment of the vertex / simplex arenas, of
// capacity vcap / tcap entries, reused in place while a new mesh of the child fits.
struct PolyKid {
    bool active = false;
    long long voff = 0, nv = 0, toff = 0, nt = 0;
    long long vcap = 0, tcap = 0;
};

struct felt_polys {
    felt_surface* s = nullptr;
    std::vector<PolyKid> kids;
    Changes a, b;
    Changes alive;  // children with a mesh (active)
    Changes* pending = nullptr;
    Changes* done = nullptr;
    int vw = 6;
    DBuf<float> varena, varena2;      // the arena and the spare a compaction copies into
    DBuf<uint32_t> tarena, tarena2;
    long long vtop = 0, ttop = 0;     // bump pointers (entries)
    DBuf<long long> dst;              // per marched partition: vertex and simplex destinations
    HBuf<long long> hdst;
    HBuf<int> hbuf;                   // staging of partition lists and counts
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
// Allocates a full march batch's scratch and a spare arena ahead of time.
int felt_polys_reserve(felt_polys* q);
int felt_polys_child_mesh_impl(felt_polys* q, uint64_t p, float* vtx, size_t vcap, uint32_t* spx,
                               size_t scap);
// Device bytes the polygoniser holds: mesh arenas and march scratch.
uint64_t felt_polys_device_bytes(const felt_polys* q);
void felt_mc_table_host(uint64_t* out);
void felt_ms_table_host(uint32_t* out);
