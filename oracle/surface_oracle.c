/*
 * surface_oracle.c — CPU oracle of the fluid surface. TEST INFRASTRUCTURE ONLY: loaded
 * by the test suite, never by the product library.
 *
 * Restates src/physics/marching_cubes.cpp:312-487 (sampleField, fieldAt, buildFluidField,
 * marchFluidField) and src/asset/vertex_pack.h:19-52 (packSnorm2x16, packUnitVector) in
 * C, FP32, in the reference's evaluation order (glm's scalar order as in sph_oracle.c).
 * Serial: the vertices come out in cell order.
 *
 * Pins (tests/test_surface_cpu.py): the field against an O(N x samples) brute force
 * over all particles (the reference's own check, tests/marching_cubes_test.cpp:260-318)
 * and the zero faces; packUnitVector against its inverse and closed cases; marching
 * against a sphere's closed form (vertices within a cell of the radius, normals and
 * windings outward) and the table's crack-free property.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "sph_c_api.h"

static uint32_t hash3(int32_t x, int32_t y, int32_t z, uint32_t cellCount) {
    return (((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u)) & (cellCount - 1u);
}

/* FluidGrid::build (fluid_sph.cpp:160-188) over pos4: bucket of each particle, members
 * in ascending index, at most 32 kept per bucket. */
typedef struct {
    uint32_t* start;   /* cellCount + 1 */
    uint32_t* members; /* n */
} Grid;

static void grid_build(Grid* g, const float* pos4, uint32_t n, float h, uint32_t cellCount) {
    uint32_t* key = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    g->start = (uint32_t*)calloc(cellCount + 1u, sizeof(uint32_t));
    g->members = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t i = 0; i < n; ++i) {
        const float* p = pos4 + 4 * (size_t)i;
        key[i] = hash3((int32_t)floorf(p[0] / h), (int32_t)floorf(p[1] / h), (int32_t)floorf(p[2] / h), cellCount);
        g->start[key[i] + 1u] += 1u;
    }
    for (uint32_t b = 0; b < cellCount; ++b) g->start[b + 1u] += g->start[b];
    uint32_t* fill = (uint32_t*)calloc(cellCount, sizeof(uint32_t));
    for (uint32_t i = 0; i < n; ++i) g->members[g->start[key[i]] + fill[key[i]]++] = i;
    free(fill);
    free(key);
}

static void grid_free(Grid* g) {
    free(g->start);
    free(g->members);
}

/* sampleField (:312-343). */
static float sample_field(const float* pos4, const Grid* g, uint32_t cellCount, const float pt[3], float radius) {
    const float radiusSq = radius * radius;
    float total = 0.0f;
    const int32_t bx = (int32_t)floorf(pt[0] / radius), by = (int32_t)floorf(pt[1] / radius),
                  bz = (int32_t)floorf(pt[2] / radius);
    uint32_t visited[27];
    uint32_t nvisited = 0;
    for (int z = -1; z <= 1; ++z) {
        for (int y = -1; y <= 1; ++y) {
            for (int x = -1; x <= 1; ++x) {
                const uint32_t b = hash3(bx + x, by + y, bz + z, cellCount);
                int seen = 0;
                for (uint32_t v = 0; v < nvisited; ++v) seen |= visited[v] == b;
                if (seen) continue;
                visited[nvisited++] = b;
                uint32_t count = g->start[b + 1u] - g->start[b];
                if (count > SPH_CELL_CAPACITY) count = SPH_CELL_CAPACITY;
                for (uint32_t k = 0; k < count; ++k) {
                    const float* q = pos4 + 4 * (size_t)g->members[g->start[b] + k];
                    const float dx = q[0] - pt[0], dy = q[1] - pt[1], dz = q[2] - pt[2];
                    const float distanceSq = (dx * dx + dy * dy) + dz * dz;
                    if (distanceSq >= radiusSq) continue;
                    const float d = 1.0f - distanceSq / radiusSq;
                    total += d * d * d;
                }
            }
        }
    }
    return total;
}

/* buildFluidField (:358-391). */
void orc_surface_field(const float* pos4, uint32_t n, const SphParams* p, uint32_t resolution, float* field) {
    const uint32_t samples = resolution + 1u;
    const size_t total = (size_t)samples * samples * samples;
    memset(field, 0, total * sizeof(float));
    if (n == 0) return;
    Grid g;
    grid_build(&g, pos4, n, p->smoothingRadius, p->cellCount);
    float cell[3];
    for (int a = 0; a < 3; ++a) cell[a] = (p->containerMax[a] - p->containerMin[a]) / (float)resolution;
    for (size_t i = 0; i < total; ++i) {
        const uint32_t x = (uint32_t)(i % samples), y = (uint32_t)((i / samples) % samples),
                       z = (uint32_t)(i / ((size_t)samples * samples));
        if (x == 0 || y == 0 || z == 0 || x == resolution || y == resolution || z == resolution) continue;
        const float pt[3] = {p->containerMin[0] + (float)x * cell[0], p->containerMin[1] + (float)y * cell[1],
                             p->containerMin[2] + (float)z * cell[2]};
        field[i] = sample_field(pos4, &g, p->cellCount, pt, p->smoothingRadius);
    }
    grid_free(&g);
}

/* asset::detail::packSnorm2x16 and asset::packUnitVector (vertex_pack.h:24-52). */
static uint32_t quantize(float x) {
    const float c = x < -1.0f ? -1.0f : (1.0f < x ? 1.0f : x);
    return (uint32_t)(int32_t)lroundf(c * 32767.0f) & 0xFFFFu;
}

uint32_t orc_pack_unit(const float nin[3]) {
    float n[3] = {nin[0], nin[1], nin[2]};
    const float sum = (fabsf(n[0]) + fabsf(n[1])) + fabsf(n[2]);
    if (sum <= 0.0f) return quantize(0.0f) | (quantize(0.0f) << 16);
    for (int a = 0; a < 3; ++a) n[a] = n[a] / sum;
    float u = n[0], v = n[1];
    if (n[2] < 0.0f) {
        const float sx = n[0] >= 0.0f ? 1.0f : -1.0f, sy = n[1] >= 0.0f ? 1.0f : -1.0f;
        u = (1.0f - fabsf(n[1])) * sx;
        v = (1.0f - fabsf(n[0])) * sy;
    }
    return quantize(u) | (quantize(v) << 16);
}

static const int kCorners[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
static const int kEdges[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6}, {6, 7}, {7, 4},
                                  {0, 4}, {1, 5}, {2, 6}, {3, 7}};

/* fieldAt (:346-355). */
static float field_at(const float* f, uint32_t s, int x, int y, int z) {
    if (x < 0 || y < 0 || z < 0 || x >= (int)s || y >= (int)s || z >= (int)s) return 0.0f;
    return f[(size_t)z * s * s + (size_t)y * s + (size_t)x];
}

/* marchFluidField (:393-487), serial; `table` is the 256 x 16 case table. */
uint32_t orc_march_field(const float* f, uint32_t resolution, const float origin[3], const float cellSize[3],
                         float iso, const int8_t* table, float* outPos, uint32_t* outNormal, uint32_t capacity) {
    const uint32_t s = resolution + 1u;
    if (capacity < 3) return 0;
    uint32_t written = 0;
    for (uint32_t i = 0; i < resolution * resolution * resolution; ++i) {
        const int base[3] = {(int)(i % resolution), (int)((i / resolution) % resolution),
                             (int)(i / (resolution * resolution))};
        float values[8];
        uint32_t code = 0;
        for (int c = 0; c < 8; ++c) {
            values[c] = field_at(f, s, base[0] + kCorners[c][0], base[1] + kCorners[c][1], base[2] + kCorners[c][2]);
            if (values[c] >= iso) code |= 1u << c;
        }
        const int8_t* entry = table + 16u * code;
        if (entry[0] < 0) continue;
        uint32_t vertices = 0;
        while (vertices < 16u && entry[vertices] >= 0) ++vertices;
        if (written + vertices > capacity) break;   /* serial: the first overflow ends the list */
        for (uint32_t v = 0; v < vertices; ++v) {
            const int e = entry[v], a = kEdges[e][0], b = kEdges[e][1];
            const float va = values[a], vb = values[b];
            float t = fabsf(vb - va) > 1.0e-8f ? (iso - va) / (vb - va) : 0.5f;
            t = t < 0.0f ? 0.0f : (1.0f < t ? 1.0f : t);
            int ca[3], cb[3];
            float ga[3], gb[3], n[3];
            for (int k = 0; k < 3; ++k) {
                ca[k] = base[k] + kCorners[a][k];
                cb[k] = base[k] + kCorners[b][k];
                const float pa = origin[k] + (float)ca[k] * cellSize[k];
                const float pb = origin[k] + (float)cb[k] * cellSize[k];
                outPos[3 * (size_t)(written + v) + k] = pa + (pb - pa) * t;
            }
            for (int k = 0; k < 3; ++k) {
                const int u[3] = {k == 0, k == 1, k == 2};
                ga[k] = (field_at(f, s, ca[0] + u[0], ca[1] + u[1], ca[2] + u[2]) -
                         field_at(f, s, ca[0] - u[0], ca[1] - u[1], ca[2] - u[2])) / cellSize[k];
                gb[k] = (field_at(f, s, cb[0] + u[0], cb[1] + u[1], cb[2] + u[2]) -
                         field_at(f, s, cb[0] - u[0], cb[1] - u[1], cb[2] - u[2])) / cellSize[k];
            }
            for (int k = 0; k < 3; ++k) n[k] = -(ga[k] + (gb[k] - ga[k]) * t);
            if (!(sqrtf((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]) > 1.0e-12f)) {
                n[0] = 0.0f;
                n[1] = 1.0f;
                n[2] = 0.0f;
            }
            const float inv = 1.0f / sqrtf((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
            const float unit[3] = {n[0] * inv, n[1] * inv, n[2] * inv};
            outNormal[written + v] = orc_pack_unit(unit);
        }
        written += vertices;
    }
    return written;
}
