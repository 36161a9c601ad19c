// sph_mc_table.cpp — the marching-cubes case table, generated from its defining rule.
//
// Corner and edge numbering are the reference's (src/physics/marching_cubes.h:20-24,
// marching_cubes.cpp:24-37): corners (0,0,0) (1,0,0) (1,1,0) (0,1,0) (0,0,1) (1,0,1)
// (1,1,1) (0,1,1); edges 0-1 1-2 2-3 3-0 4-5 5-6 6-7 7-4 0-4 1-5 2-6 3-7. The rule
// (marching_cubes.cpp:17-22) is: the contour on each cube face is decided by that face's
// four corner signs alone — so two cells sharing a face produce the same segments and
// the surface has no cracks — an ambiguous face (two diagonal corners inside) cuts each
// inside corner off separately, and the triangles wind so their normal points from the
// inside corners to the outside ones. This file builds the table from that rule: face
// segments → closed loops of crossed edges → each loop oriented outward and fanned into
// triangles. The generator is written from that rule alone; the table it produces equals
// the reference's MC_TABLE (marching_cubes.cpp:39-296) entry for entry.
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <vector>

#include "sph_internal.h"

namespace {

constexpr int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
constexpr int kEdge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6}, {6, 7}, {7, 4},
                              {0, 4}, {1, 5}, {2, 6}, {3, 7}};
// Each face as a cycle of four corners.
constexpr int kFace[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4}, {3, 2, 6, 7}, {0, 3, 7, 4}, {1, 2, 6, 5}};

int edge_between(int a, int b) {
    for (int e = 0; e < 12; ++e)
        if ((kEdge[e][0] == a && kEdge[e][1] == b) || (kEdge[e][0] == b && kEdge[e][1] == a)) return e;
    return -1;
}

struct V3 {
    double x, y, z;
};

V3 midpoint(int e) {
    const int* a = kCorner[kEdge[e][0]];
    const int* b = kCorner[kEdge[e][1]];
    return {(a[0] + b[0]) * 0.5, (a[1] + b[1]) * 0.5, (a[2] + b[2]) * 0.5};
}

V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Sum over the edges of a triangle or loop of (outside corner - inside corner).
V3 outward(uint32_t code, const std::vector<int>& edges) {
    V3 t{0, 0, 0};
    for (int e : edges) {
        const int a = kEdge[e][0], b = kEdge[e][1];
        const bool ain = (code >> a) & 1u;
        const int from = ain ? a : b, to = ain ? b : a;
        t.x += kCorner[to][0] - kCorner[from][0];
        t.y += kCorner[to][1] - kCorner[from][1];
        t.z += kCorner[to][2] - kCorner[from][2];
    }
    return t;
}

double facing(uint32_t code, int a, int b, int c) {
    const V3 pa = midpoint(a), pb = midpoint(b), pc = midpoint(c);
    return dot(cross(sub(pb, pa), sub(pc, pa)), outward(code, {a, b, c}));
}

struct Table {
    int8_t t[256][16];
    bool ok = true;
    Table() {
        memset(t, -1, sizeof(t));
        for (uint32_t code = 0; code < 256; ++code) ok = build(code) && ok;
    }

    bool build(uint32_t code) {
        // segments on the faces: adjacency of crossed edges
        std::array<std::vector<int>, 12> link;
        for (const auto& f : kFace) {
            int e[4];
            bool in[4];
            for (int k = 0; k < 4; ++k) {
                e[k] = edge_between(f[k], f[(k + 1) % 4]);   // edge from corner k to corner k+1
                in[k] = (code >> f[k]) & 1u;
            }
            std::vector<int> active;
            for (int k = 0; k < 4; ++k)
                if (in[k] != in[(k + 1) % 4]) active.push_back(k);
            if (active.size() == 2) {
                link[e[active[0]]].push_back(e[active[1]]);
                link[e[active[1]]].push_back(e[active[0]]);
            } else if (active.size() == 4) {
                // ambiguous face: cut each inside corner off by itself; corner k lies
                // between edge k - 1 and edge k of the cycle
                for (int k = 0; k < 4; ++k) {
                    if (!in[k]) continue;
                    const int a = e[(k + 3) % 4], b = e[k];
                    link[a].push_back(b);
                    link[b].push_back(a);
                }
            }
        }
        // closed loops through the crossed edges (each has exactly two neighbours)
        std::array<bool, 12> used{};
        std::vector<std::vector<int>> loops;
        for (int start = 0; start < 12; ++start) {
            if (used[start] || link[start].empty()) continue;
            if (link[start].size() != 2) return false;
            std::vector<int> loop{start};
            used[start] = true;
            int prev = start, cur = link[start][0];
            while (cur != start) {
                if (link[cur].size() != 2 || used[cur]) return false;
                loop.push_back(cur);
                used[cur] = true;
                const int next = link[cur][0] == prev ? link[cur][1] : link[cur][0];
                prev = cur;
                cur = next;
            }
            loops.push_back(loop);
        }
        int n = 0;
        for (auto& loop : loops) {
            // orient the loop outward (Newell normal against the inside -> outside direction)
            V3 nrm{0, 0, 0};
            for (size_t i = 0; i < loop.size(); ++i) {
                const V3 p = midpoint(loop[i]), q = midpoint(loop[(i + 1) % loop.size()]);
                nrm.x += (p.y - q.y) * (p.z + q.z);
                nrm.y += (p.z - q.z) * (p.x + q.x);
                nrm.z += (p.x - q.x) * (p.y + q.y);
            }
            if (dot(nrm, outward(code, loop)) < 0.0) std::reverse(loop.begin(), loop.end());
            // fan from the first start vertex whose triangles all face outward
            const size_t m = loop.size();
            size_t pick = m;
            for (size_t s = 0; s < m && pick == m; ++s) {
                bool good = true;
                for (size_t i = 1; i + 1 < m; ++i)
                    good = good && facing(code, loop[s], loop[(s + i) % m], loop[(s + i + 1) % m]) >= -1e-9;
                if (good) pick = s;
            }
            if (pick == m) return false;
            for (size_t i = 1; i + 1 < m; ++i) {
                if (n + 3 > 15) return false;
                t[code][n++] = (int8_t)loop[pick];
                t[code][n++] = (int8_t)loop[(pick + i) % m];
                t[code][n++] = (int8_t)loop[(pick + i + 1) % m];
            }
        }
        return true;
    }
};

}  // namespace

namespace sphhost {

const int8_t* mc_table() {
    static const Table table;
    return table.ok ? &table.t[0][0] : nullptr;
}

}  // namespace sphhost

// The 256 x 16 case table (edge indices, -1 terminated), for tests and host tools.
extern "C" int sph_mc_table(int8_t* out) {
    const int8_t* t = sphhost::mc_table();
    if (!out) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "sph_mc_table: null out");
    if (!t) return sphhost::fail(SPH_ERR_STATE, "marching-cubes table generation failed");
    memcpy(out, t, 256 * 16);
    return SPH_OK;
}
