// Fragment geometry for the rigid handoff: the exposed cell faces of a
// fragment, the convex collision proxy sampled from them, and the greedy merge
// of a fragment's cells into boxes. These are the geometric half of the
// reference's fragment build (fracture/FragmentGeometry.cpp:54-291,
// fracture/PieceGeometry.cpp:9-46); the mass half runs on the device
// (blx_backend_fragment_mass_properties), over the canonical labels the device
// also computes. Host code: a fragment's surface is linear in its cells and is
// read once, when the fragment leaves the lattice.

#include "banjo_lattice.h"
#include "lattice_errors.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

namespace blx {
namespace {

// A cell's grid coordinates packed into one sortable key (21 bits per axis,
// offset so negative coordinates order correctly).
constexpr int64_t kGridBias = int64_t(1) << 20;
uint64_t cellKey(int32_t x, int32_t y, int32_t z) {
    return (uint64_t(int64_t(z) + kGridBias) << 42) | (uint64_t(int64_t(y) + kGridBias) << 21) |
           uint64_t(int64_t(x) + kGridBias);
}
void checkCell(int32_t x, int32_t y, int32_t z) {
    const int64_t lim = kGridBias - 1;
    if (std::llabs(x) >= lim || std::llabs(y) >= lim || std::llabs(z) >= lim)
        throw Invalid("grid coordinates must lie within (-2^20 + 1, 2^20 - 1)");
}

// The six faces of a unit cell, in the order the surface lists them: the
// direction to the neighbour whose absence exposes the face, its outward
// normal, and its four corners (outward winding) in cell units about the
// centre.
struct Face {
    int32_t d[3];
    double n[3];
    double c[4][3];
};
constexpr Face kFaces[6] = {
    {{1, 0, 0}, {1, 0, 0}, {{0.5, -0.5, -0.5}, {0.5, 0.5, -0.5}, {0.5, 0.5, 0.5}, {0.5, -0.5, 0.5}}},
    {{-1, 0, 0}, {-1, 0, 0}, {{-0.5, -0.5, 0.5}, {-0.5, 0.5, 0.5}, {-0.5, 0.5, -0.5}, {-0.5, -0.5, -0.5}}},
    {{0, 1, 0}, {0, 1, 0}, {{-0.5, 0.5, -0.5}, {-0.5, 0.5, 0.5}, {0.5, 0.5, 0.5}, {0.5, 0.5, -0.5}}},
    {{0, -1, 0}, {0, -1, 0}, {{-0.5, -0.5, 0.5}, {-0.5, -0.5, -0.5}, {0.5, -0.5, -0.5}, {0.5, -0.5, 0.5}}},
    {{0, 0, 1}, {0, 0, 1}, {{-0.5, -0.5, 0.5}, {0.5, -0.5, 0.5}, {0.5, 0.5, 0.5}, {-0.5, 0.5, 0.5}}},
    {{0, 0, -1}, {0, 0, -1}, {{0.5, -0.5, -0.5}, {-0.5, -0.5, -0.5}, {-0.5, 0.5, -0.5}, {0.5, 0.5, -0.5}}},
};

struct P3 {
    double x, y, z;
};
double dotP(const P3 &a, const P3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

} // namespace
} // namespace blx

using namespace blx;

extern "C" {

int blx_exposed_voxel_surface(const int32_t *grid, const double *position, uint32_t count,
                              const double center_of_mass[3], double voxel_size_m, double *vertex_position,
                              double *vertex_normal, uint32_t *indices, uint64_t face_capacity,
                              uint64_t *face_count) {
    return guard([&] {
        if (count == 0 || !grid || !position || !center_of_mass || !face_count)
            throw Invalid("surface mesh requires an active material component");
        std::vector<uint64_t> cells(count);
        for (uint32_t k = 0; k < count; ++k) {
            checkCell(grid[3 * size_t(k)], grid[3 * size_t(k) + 1], grid[3 * size_t(k) + 2]);
            cells[k] = cellKey(grid[3 * size_t(k)], grid[3 * size_t(k) + 1], grid[3 * size_t(k) + 2]);
        }
        std::sort(cells.begin(), cells.end());
        auto occupied = [&](int32_t x, int32_t y, int32_t z) {
            return std::binary_search(cells.begin(), cells.end(), cellKey(x, y, z));
        };
        uint64_t faces = 0;
        for (uint32_t k = 0; k < count; ++k) {
            const int32_t *g = grid + 3 * size_t(k);
            const double *p = position + 3 * size_t(k);
            // The cell centre relative to the centre of mass, then each corner
            // at centre + h * corner, as double sums.
            const double cx = p[0] - center_of_mass[0], cy = p[1] - center_of_mass[1], cz = p[2] - center_of_mass[2];
            for (const Face &f : kFaces) {
                if (occupied(g[0] + f.d[0], g[1] + f.d[1], g[2] + f.d[2])) continue;
                if (faces < face_capacity) {
                    for (int q = 0; q < 4; ++q) {
                        double *v = vertex_position + 12 * faces + 3 * q;
                        v[0] = cx + voxel_size_m * f.c[q][0];
                        v[1] = cy + voxel_size_m * f.c[q][1];
                        v[2] = cz + voxel_size_m * f.c[q][2];
                        double *nrm = vertex_normal + 12 * faces + 3 * q;
                        nrm[0] = f.n[0];
                        nrm[1] = f.n[1];
                        nrm[2] = f.n[2];
                    }
                    const uint32_t base = uint32_t(4 * faces);
                    const uint32_t tri[6] = {base, base + 1, base + 2, base, base + 2, base + 3};
                    std::copy(tri, tri + 6, indices + 6 * faces);
                }
                ++faces;
            }
        }
        *face_count = faces;
    });
}

int blx_collision_points(const double *vertex_position, uint64_t vertex_count, const double first_voxel_center[3],
                         double voxel_size_m, uint64_t maximum_points, double *points, uint64_t capacity,
                         uint64_t *count) {
    return guard([&] {
        if (vertex_count == 0 || !vertex_position || !first_voxel_center || !count)
            throw Invalid("collision proxy requires surface geometry");
        const uint64_t cap_points = std::max<uint64_t>(maximum_points, 8);
        std::vector<P3> out;
        out.reserve(cap_points + 8);
        auto vtx = [&](uint64_t k) {
            return P3{vertex_position[3 * k], vertex_position[3 * k + 1], vertex_position[3 * k + 2]};
        };
        // A point within 1e-8 m of one already taken is the same point.
        auto add = [&](const P3 &c) {
            for (const P3 &e : out) {
                const P3 d{e.x - c.x, e.y - c.y, e.z - c.z};
                if (dotP(d, d) <= 1.0e-16) return;
            }
            out.push_back(c);
        };
        // The support point of the surface in each of the 26 lattice
        // directions (the first vertex that attains the largest projection),
        // so the proxy keeps the piece's extent in every direction.
        for (int dz = -1; dz <= 1 && out.size() < cap_points; ++dz)
            for (int dy = -1; dy <= 1 && out.size() < cap_points; ++dy)
                for (int dx = -1; dx <= 1 && out.size() < cap_points; ++dx) {
                    if (dx == 0 && dy == 0 && dz == 0) continue;
                    const P3 raw{double(dx), double(dy), double(dz)};
                    const double len = std::sqrt(dotP(raw, raw));
                    const P3 dir = len > 1.0e-12 ? P3{raw.x / len, raw.y / len, raw.z / len} : P3{1.0, 0.0, 0.0};
                    uint64_t best = 0;
                    double best_p = dotP(vtx(0), dir);
                    for (uint64_t k = 0; k < vertex_count; ++k) {
                        const double p = dotP(vtx(k), dir);
                        if (p > best_p) {
                            best = k;
                            best_p = p;
                        }
                    }
                    add(vtx(best));
                }
        // Then an even sample of the surface's vertices up to the budget.
        const uint64_t left = cap_points - out.size();
        if (left > 0) {
            const uint64_t stride = std::max<uint64_t>(1, (vertex_count + left - 1) / left);
            for (uint64_t k = 0; k < vertex_count && out.size() < cap_points; k += stride) add(vtx(k));
        }
        // A proxy of fewer than four distinct points gets the first cell's corners.
        if (out.size() < 4) {
            const double half = 0.5 * voxel_size_m;
            const P3 c{first_voxel_center[0], first_voxel_center[1], first_voxel_center[2]};
            for (int z = -1; z <= 1; z += 2)
                for (int y = -1; y <= 1; y += 2)
                    for (int x = -1; x <= 1; x += 2)
                        add(P3{c.x + double(x) * half, c.y + double(y) * half, c.z + double(z) * half});
        }
        *count = out.size();
        if (points)
            for (uint64_t k = 0; k < std::min<uint64_t>(capacity, out.size()); ++k) {
                points[3 * k] = out[k].x;
                points[3 * k + 1] = out[k].y;
                points[3 * k + 2] = out[k].z;
            }
    });
}

int blx_merge_cells_into_boxes(const int32_t *cells, uint64_t count, int32_t *boxes, uint64_t capacity,
                               uint64_t *box_count) {
    return guard([&] {
        if (!box_count || (count && !cells)) throw Invalid("merge needs the cells");
        // The cells as sorted keys, ascending (z, y, x); a taken cell is marked.
        std::vector<uint64_t> keys(count);
        for (uint64_t k = 0; k < count; ++k) {
            checkCell(cells[3 * k], cells[3 * k + 1], cells[3 * k + 2]);
            keys[k] = cellKey(cells[3 * k], cells[3 * k + 1], cells[3 * k + 2]);
        }
        std::sort(keys.begin(), keys.end());
        keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
        std::vector<uint8_t> taken(keys.size(), 0);
        auto slot = [&](int32_t x, int32_t y, int32_t z) -> int64_t {
            const auto it = std::lower_bound(keys.begin(), keys.end(), cellKey(x, y, z));
            return it != keys.end() && *it == cellKey(x, y, z) ? int64_t(it - keys.begin()) : -1;
        };
        auto free_cell = [&](int32_t x, int32_t y, int32_t z) {
            const int64_t s = slot(x, y, z);
            return s >= 0 && !taken[size_t(s)];
        };
        auto coord = [](uint64_t key, int shift) { return int32_t(int64_t((key >> shift) & 0x1fffffULL) - kGridBias); };
        uint64_t n = 0;
        size_t next = 0;
        for (;;) {
            while (next < keys.size() && taken[next]) ++next;
            if (next == keys.size()) break;
            // The smallest remaining cell in (z, y, x) order starts a box: a run
            // along x, grown row by row along y and slab by slab along z while
            // every cell of the next row or slab is free.
            const int32_t x0 = coord(keys[next], 0), y0 = coord(keys[next], 21), z0 = coord(keys[next], 42);
            int32_t x1 = x0, y1 = y0, z1 = z0;
            while (free_cell(x1 + 1, y0, z0)) ++x1;
            for (;;) {
                bool full = true;
                for (int32_t x = x0; x <= x1 && full; ++x) full = free_cell(x, y1 + 1, z0);
                if (!full) break;
                ++y1;
            }
            for (;;) {
                bool full = true;
                for (int32_t y = y0; y <= y1 && full; ++y)
                    for (int32_t x = x0; x <= x1 && full; ++x) full = free_cell(x, y, z1 + 1);
                if (!full) break;
                ++z1;
            }
            for (int32_t z = z0; z <= z1; ++z)
                for (int32_t y = y0; y <= y1; ++y)
                    for (int32_t x = x0; x <= x1; ++x) taken[size_t(slot(x, y, z))] = 1;
            if (boxes && n < capacity) {
                int32_t *b = boxes + 6 * n;
                b[0] = x0; b[1] = y0; b[2] = z0; b[3] = x1; b[4] = y1; b[5] = z1;
            }
            ++n;
        }
        *box_count = n;
    });
}

} // extern "C"
