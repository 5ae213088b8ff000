// xcdplan.hpp -- per-XCD work lists of the batched K1 (host code, also run
// by the CPU test mirror).
//
// Workgroups reach the 8 XCDs round robin (workgroup b on XCD b mod 8), each
// XCD with its own L2.  The batched K1 therefore walks one list of work items
// per XCD instead of a plain grid-stride loop, so that an XCD gathers from
// the x vectors of the few members it is working on.
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

namespace blqh {

constexpr int kXcds = 8;

// One batch member as the planner sees it: its work items [item0, item1),
// its nonzeros, its column panels and its value kind.
struct XcdMember {
    int64_t item0, item1;
    int64_t nnz;
    int npan;
    int real;
};

// Lists for kinds 0 (real members), 1 (complex members) and 2 (all members):
// items[seg[9 k + x] .. seg[9 k + x + 1]) is the list of XCD x for kind k
// (items of kinds 0 and 1 fill the first nitems entries, kind 2 the next).
// Per kind:
//  - a member of one panel (x of at most 65,536 rows, 1 MiB) is cut into
//    chunks of consecutive items: whole members when there are many, so that
//    an XCD works on few members at a time and their x stays in its L2; at
//    least 64 chunks in all, so that a few members still spread over every
//    XCD.  Chunks are dealt largest first, each to the XCD with the fewest
//    nonzeros so far (ties: lowest XCD);
//  - a member of several panels is dealt item by item, item i to XCD i mod 8,
//    so that all XCDs walk it together (measured faster for 204k-state
//    members than chunks);
//  - a list holds its items in ascending order.
inline void xcd_lists(const std::vector<XcdMember>& mem, std::vector<int32_t>& items, std::vector<int32_t>& seg) {
    const int P = (int)mem.size();
    int64_t nitems = 0;
    for (const XcdMember& m : mem) nitems += m.item1 - m.item0;
    items.clear();
    items.reserve(2 * nitems);
    seg.assign(3 * (kXcds + 1), 0);
    struct Chunk {
        int64_t i0, i1;
        double w;
    };
    for (int kind = 0; kind < 3; ++kind) {
        auto mine = [&](int p) { return kind == 2 || (mem[p].real ? 0 : 1) == kind; };
        auto per_item = [&](int p) { return (double)mem[p].nnz / (double)(mem[p].item1 - mem[p].item0); };
        std::vector<double> load(kXcds, 0.0);
        std::vector<std::vector<Chunk>> own(kXcds);
        int64_t small_items = 0;
        for (int p = 0; p < P; ++p) {
            if (!mine(p)) continue;
            if (mem[p].npan == 1) {
                small_items += mem[p].item1 - mem[p].item0;
                continue;
            }
            for (int64_t it = mem[p].item0; it < mem[p].item1; ++it) {
                own[it % kXcds].push_back({it, it + 1, per_item(p)});
                load[it % kXcds] += per_item(p);
            }
        }
        const int64_t cmax = std::max<int64_t>(1, (small_items + 63) / 64);
        std::vector<Chunk> ch;
        for (int p = 0; p < P; ++p) {
            if (!mine(p) || mem[p].npan != 1) continue;
            for (int64_t i0 = mem[p].item0; i0 < mem[p].item1; i0 += cmax) {
                const int64_t i1 = std::min(i0 + cmax, mem[p].item1);
                ch.push_back({i0, i1, per_item(p) * (double)(i1 - i0)});
            }
        }
        std::stable_sort(ch.begin(), ch.end(), [](const Chunk& a, const Chunk& c) { return a.w > c.w; });
        for (const Chunk& c : ch) {
            const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            load[x] += c.w;
            own[x].push_back(c);
        }
        for (int x = 0; x < kXcds; ++x) {
            std::sort(own[x].begin(), own[x].end(), [](const Chunk& a, const Chunk& c) { return a.i0 < c.i0; });
            seg[kind * (kXcds + 1) + x] = (int32_t)items.size();
            for (const Chunk& c : own[x])
                for (int64_t it = c.i0; it < c.i1; ++it) items.push_back((int32_t)it);
        }
        seg[kind * (kXcds + 1) + kXcds] = (int32_t)items.size();
    }
}

}  // namespace blqh
