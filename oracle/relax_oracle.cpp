// ORACLE — test infrastructure only. This file is the CPU restatement of
// TANGLE's device relaxation loop used to check the HIP backend; it is never
// linked into, loaded by or called from the product library. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
//
// It restates, as plain sequential (OpenMP-parallel where the reference runs
// one GPU thread per independent unit) C++, the reference's Rust/CubeCL code:
//
//   packing             crates/tangle_relax/src/device/packed.rs:150-320, 527-580, 1058-1063
//   upload sizing       crates/tangle_relax/src/device/world.rs:308-417, 496-509, 868-917
//   cell list           crates/tangle_contact/src/device/cell_list.rs:7-378
//   neighbour lists     crates/tangle_contact/src/device/neighbor.rs:25-465
//   contact             crates/tangle_contact/src/device/contact.rs:232-902
//   relax kernels       crates/tangle_relax/src/device/kernels.rs:3-162, 925-1537
//   batch loop/status   crates/tangle_relax/src/device/world.rs:1020-1505, 2337-2369
//   compaction          crates/tangle_relax/src/device/kernels.rs:712-923, world.rs:1931-2100
//   contact capture     crates/tangle_contact/src/device/contact.rs:9-230, world.rs:2193-2334
//
// Arithmetic follows Rust f32 semantics: no FMA contraction (compile with
// -ffp-contract=off), IEEE division and square root, f32::round = half away
// from zero, saturating float->int casts, f32::max/min ignore NaN,
// f32::clamp passes NaN through. Only round (circular) sections and the
// non-adaptive packing are restated; there every segment is a single proxy
// piece (world.rs:370-391 gives list weight 1 to every segment), so the
// proxy-piece arithmetic reduces to the segment midpoint (cell_list.rs:50-58).
//
// Parity pins for this file are in tests/test_oracle_*.py (known answers and
// invariants taken from crates/tangle_relax/src/device/world_tests.rs and
// crates/tangle_core/src/curvature.rs, brute-force pair enumeration, and the
// independent numpy restatement of the cell hash in oracle/cell_hash.py).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../include/tangle_hip.h"

namespace {

constexpr uint32_t kNone = 0xffffffffu;
constexpr size_t kDefaultMaxListBytes = size_t(120) << 20;  // world.rs:36

thread_local std::string g_error;

int fail(int code, const std::string &message) {
    g_error = message;
    return code;
}

// ---- Rust f32 semantics -------------------------------------------------

inline float rclamp(float x, float lo, float hi) {  // f32::clamp
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}
inline float rmax(float a, float b) { return std::fmax(a, b); }  // f32::max
inline float rmin(float a, float b) { return std::fmin(a, b); }  // f32::min
inline int32_t sat_i32(float x) {                                // `as i32`
    if (std::isnan(x)) return 0;
    if (x <= -2147483648.0f) return INT32_MIN;
    if (x >= 2147483648.0f) return INT32_MAX;
    return (int32_t)x;
}
inline uint32_t sat_u32(float x) {  // `as u32`
    if (std::isnan(x) || x <= 0.0f) return 0;
    if (x >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)x;
}
inline uint32_t f2b(float x) {
    uint32_t b;
    std::memcpy(&b, &x, 4);
    return b;
}
inline float b2f(uint32_t b) {
    float x;
    std::memcpy(&x, &b, 4);
    return x;
}

struct World {
    // packed assembly (packed.rs:47-141)
    uint32_t F = 0, V = 0, S = 0;
    std::vector<float> pos, intr;          // 3V
    std::vector<uint32_t> seg_vtx;         // 2S
    std::vector<uint32_t> seg_fiber;       // S
    std::vector<float> seg_radius, seg_rest;
    std::vector<uint32_t> vtx_seg;         // 2V (previous, next)
    std::vector<float> vtx_kmax;           // V
    std::vector<uint32_t> vtx_pinned;      // V
    std::vector<uint32_t> fiber_span;      // 2F (start, count)
    std::vector<float> arc;                // V
    std::vector<float> wall_extent;        // 3V
    float lower[3], upper[3];
    uint32_t periodic[3];
    // formation (packed.rs:94, 202-210, 289, 303): every fiber active at upload
    std::vector<uint32_t> fiber_step, fiber_layer, fiber_active;  // F
    // installed targets (world.rs:1723-1927)
    std::vector<float> layer_targets;
    uint32_t layer_axis = 0, first_layer = 0, last_layer = 0;
    float layer_stiffness = 0, layer_max = 0;
    std::vector<uint32_t> vt_index;
    std::vector<float> vt_target, vt_command;
    uint32_t vt_axis = 0;
    float vt_stiffness = 0, vt_max = 0;

    // sizing (world.rs:322-417)
    float skin = 0, cell_size = 0, segment_cell_size = 0, max_radius = 0;
    uint32_t cx = 1, cy = 1, cz = 1, C = 1;
    size_t slots = 1;
    uint32_t capacity = 48, block = 48;

    // state
    uint32_t control[4] = {0, 0, 0, 0};     // running, lifetime, batch, terminal
    float metrics[3];
    uint32_t nstate[2] = {1, 0};            // rebuild pending, rebuild count
    uint32_t reduced[3] = {0, 0, 0};
    std::vector<uint32_t> cell_counts, cell_offsets, cell_segments;
    std::vector<uint32_t> ncount, nlist;    // nlist[block * s + k]
    std::vector<float> refpos;              // 3V
    std::vector<float> corr;                // 6S
    std::vector<float> seg_max;             // S
    std::vector<float> curv;                // V
    std::vector<float> icorr;               // 3V
    std::vector<float> vstep;               // V
    std::vector<float> wall;                // 3V
    // per fiber, at its last curvature cleanup: projections applied, and
    // the triplets the GPU sweep visits (its sweeps stop after a forward
    // pass that projects nothing); observation only
    std::vector<uint32_t> gs_projections, gs_visits;
};

// ---- packing: packed.rs:150-320 ------------------------------------------

double dist64(const double *a, const double *b) {  // packed.rs:1058-1063
    double x = b[0] - a[0], y = b[1] - a[1], z = b[2] - a[2];
    return std::sqrt(x * x + y * y + z * z);
}

int pack(World &w, const tgh_assembly *a) {
    if (!a || a->fiber_count == 0 || a->vertex_count == 0 || !a->fiber_vertex_counts ||
        !a->placed || !a->intrinsic || !a->fiber_radius)
        return fail(TGH_ERR_INVALID, "assembly is empty or missing arrays");
    for (int k = 0; k < 3; ++k)
        if (!(a->cell_lengths[k] > 0.0))
            return fail(TGH_ERR_INVALID, "cell lengths must be positive");
    w.F = a->fiber_count;
    w.V = a->vertex_count;
    uint64_t total = 0;
    for (uint32_t f = 0; f < w.F; ++f) {
        if (a->fiber_vertex_counts[f] < 2)
            return fail(TGH_ERR_INVALID, "every fiber needs at least two vertices");
        total += a->fiber_vertex_counts[f];
    }
    if (total != w.V) return fail(TGH_ERR_INVALID, "fiber vertex counts do not sum to vertex_count");
    w.S = w.V - w.F;
    w.pos.resize(3 * (size_t)w.V);
    w.intr.resize(3 * (size_t)w.V);
    for (size_t i = 0; i < 3 * (size_t)w.V; ++i) {
        w.pos[i] = (float)a->placed[i];
        w.intr[i] = (float)a->intrinsic[i];
    }
    w.seg_vtx.assign(2 * (size_t)w.S, 0);
    w.seg_fiber.assign(w.S, 0);
    w.seg_radius.assign(w.S, 0.f);
    w.seg_rest.assign(w.S, 0.f);
    w.vtx_seg.assign(2 * (size_t)w.V, kNone);
    w.vtx_kmax.assign(w.V, 0.f);
    w.vtx_pinned.assign(w.V, 0);
    w.fiber_span.assign(2 * (size_t)w.F, 0);
    uint32_t vstart = 0, s = 0;
    for (uint32_t f = 0; f < w.F; ++f) {
        uint32_t n = a->fiber_vertex_counts[f];
        float radius = (float)a->fiber_radius[f];
        if (!(radius > 0.f)) return fail(TGH_ERR_INVALID, "fiber radius must be positive");
        w.fiber_span[2 * f] = vstart;
        w.fiber_span[2 * f + 1] = n;
        if (a->pin_fiber_ends) {
            w.vtx_pinned[vstart] = 1;
            w.vtx_pinned[vstart + n - 1] = 1;
        }
        for (uint32_t l = 0; l + 1 < n; ++l, ++s) {
            w.seg_vtx[2 * s] = vstart + l;
            w.seg_vtx[2 * s + 1] = vstart + l + 1;
            w.seg_fiber[s] = f;
            w.seg_radius[s] = radius;
            w.seg_rest[s] = (float)dist64(&a->intrinsic[3 * (size_t)(vstart + l)],
                                          &a->intrinsic[3 * (size_t)(vstart + l + 1)]);
            w.vtx_seg[2 * (size_t)(vstart + l) + 1] = s;
            w.vtx_seg[2 * (size_t)(vstart + l + 1)] = s;
        }
        double mbr = a->fiber_min_bend_radius ? a->fiber_min_bend_radius[f] : 0.0;
        if (mbr > 0.0) {
            float kmax = (float)(1.0 / mbr);  // FiberBendLimit::maximum_curvature
            for (uint32_t l = 1; l + 1 < n; ++l) w.vtx_kmax[vstart + l] = kmax;
        }
        vstart += n;
    }
    if (a->vertex_pinned)
        for (uint32_t v = 0; v < w.V; ++v)
            if (a->vertex_pinned[v]) w.vtx_pinned[v] = 1;
    // vertex_arc_lengths: packed.rs:561-580 (f64 accumulation of f32 data)
    w.arc.assign(w.V, 0.f);
    for (uint32_t f = 0; f < w.F; ++f) {
        uint32_t st = w.fiber_span[2 * f], n = w.fiber_span[2 * f + 1];
        double acc = 0.0;
        for (uint32_t v = st + 1; v < st + n; ++v) {
            double sum = 0.0;
            for (int k = 0; k < 3; ++k) {
                double d = (double)w.intr[3 * (size_t)v + k] - (double)w.intr[3 * (size_t)(v - 1) + k];
                sum += d * d;
            }
            acc += std::sqrt(sum);
            w.arc[v] = (float)acc;
        }
    }
    // vertex_wall_extents: packed.rs:527-550 (round sections: the radius)
    w.wall_extent.assign(3 * (size_t)w.V, 0.f);
    for (uint32_t f = 0; f < w.F; ++f) {
        uint32_t st = w.fiber_span[2 * f], n = w.fiber_span[2 * f + 1];
        float r = w.seg_radius[st - f];  // first segment of fiber f
        for (uint32_t v = st; v < st + n; ++v)
            for (int k = 0; k < 3; ++k) w.wall_extent[3 * (size_t)v + k] = r;
    }
    for (int k = 0; k < 3; ++k) {
        w.lower[k] = (float)a->cell_origin[k];
        w.upper[k] = (float)(a->cell_origin[k] + a->cell_lengths[k]);
        w.periodic[k] = a->cell_periodic[k] ? 1u : 0u;
    }
    w.fiber_step.assign(w.F, 0u);
    w.fiber_layer.assign(w.F, 0xffffffffu);
    w.fiber_active.assign(w.F, 1u);
    for (uint32_t f = 0; f < w.F; ++f) {
        if (a->fiber_formation_step) w.fiber_step[f] = a->fiber_formation_step[f];
        if (a->fiber_formation_layer) w.fiber_layer[f] = a->fiber_formation_layer[f];
    }
    return TGH_OK;
}

inline bool seg_on(const World &w, uint32_t s) { return w.fiber_active[w.seg_fiber[s]] != 0; }
inline uint32_t vertex_fiber(const World &w, uint32_t v) {
    uint32_t s = w.vtx_seg[2 * v] != kNone ? w.vtx_seg[2 * v] : w.vtx_seg[2 * v + 1];
    return w.seg_fiber[s];
}
inline bool vtx_on(const World &w, uint32_t v) { return w.fiber_active[vertex_fiber(w, v)] != 0; }

// ---- upload sizing: world.rs:322-417, 496-509, 868-917 ---------------------

float seg_current_length(const World &w, uint32_t s) {
    uint32_t a = w.seg_vtx[2 * s], b = w.seg_vtx[2 * s + 1];
    float dx = w.pos[3 * b] - w.pos[3 * a];
    float dy = w.pos[3 * b + 1] - w.pos[3 * a + 1];
    float dz = w.pos[3 * b + 2] - w.pos[3 * a + 2];
    return std::sqrt(dx * dx + dy * dy + dz * dz);
}

int size_world(World &w, const tgh_config *c) {
    if (!(c->cell_size_scale >= 1.0f)) return fail(TGH_ERR_INVALID, "cell_size_scale must be >= 1");
    if (!(std::isfinite(c->neighbor_skin_scale) && c->neighbor_skin_scale >= 0.0f))
        return fail(TGH_ERR_INVALID, "neighbor_skin_scale must be finite and >= 0");
    if (c->neighbor_capacity == 0) return fail(TGH_ERR_INVALID, "neighbor_capacity must be > 0");
    if (!(c->max_step > 0.0f)) return fail(TGH_ERR_INVALID, "max_step must be positive");
    float max_rest = 0.f, max_init = 0.f, max_radius = 0.f;
    for (uint32_t s = 0; s < w.S; ++s) {
        max_rest = rmax(max_rest, w.seg_rest[s]);
        max_init = rmax(max_init, seg_current_length(w, s));
        max_radius = rmax(max_radius, w.seg_radius[s]);
    }
    float max_len = rmax(max_rest, max_init);
    w.skin = c->neighbor_skin_scale * max_radius;
    float margin = 2.0f * max_radius + rmax(2.0f * c->max_step, w.skin);
    // every segment is a leaf here, so the longest leaf is the longest segment
    float max_leaf = 0.f;
    for (uint32_t s = 0; s < w.S; ++s) max_leaf = rmax(max_leaf, rmax(w.seg_rest[s], seg_current_length(w, s)));
    float block_len = rmin(max_len, rmax(max_leaf, 2.0f * margin));
    for (uint32_t s = 0; s < w.S; ++s) {
        uint32_t wgt = sat_u32(std::ceil(rmax(w.seg_rest[s], seg_current_length(w, s)) / block_len));
        if (wgt > 1) return fail(TGH_ERR_INVALID, "segments split into proxy pieces are not supported");
    }
    float piece = block_len < max_len ? 1.05f * block_len : block_len;
    w.cell_size = (piece + margin) * c->cell_size_scale;
    w.segment_cell_size = (max_len + margin) * c->cell_size_scale;  // world.rs:405
    w.max_radius = max_radius;
    float ext[3];
    for (int k = 0; k < 3; ++k) ext[k] = w.upper[k] - w.lower[k];
    w.cx = sat_u32(rmax(std::floor(ext[0] / w.cell_size), 1.0f));
    w.cy = sat_u32(rmax(std::floor(ext[1] / w.cell_size), 1.0f));
    w.cz = sat_u32(rmax(std::floor(ext[2] / w.cell_size), 1.0f));
    uint64_t C = (uint64_t)w.cx * w.cy * w.cz;
    if (C > (1u << 28)) return fail(TGH_ERR_INVALID, "too many cells");
    w.C = (uint32_t)C;
    size_t max_bytes = c->maximum_neighbor_list_bytes ? c->maximum_neighbor_list_bytes : kDefaultMaxListBytes;
    w.capacity = c->neighbor_capacity;
    w.slots = std::max<size_t>(1, std::min<size_t>((size_t)w.capacity * w.S, max_bytes / 4));
    size_t total = std::max<size_t>(1, w.S);  // every list weight is 1
    w.block = (uint32_t)std::min<size_t>(w.capacity, std::min<size_t>(w.slots / total, 0xffffffffu));
    return TGH_OK;
}

// ---- cell list: cell_list.rs:32-79, 114-378 --------------------------------

inline void cell_coords(const World &w, float px, float py, float pz, float qx, float qy, float qz,
                        uint32_t &x, uint32_t &y, uint32_t &z) {
    float mx = 0.5f * (px + qx), my = 0.5f * (py + qy), mz = 0.5f * (pz + qz);
    float wx = (w.upper[0] - w.lower[0]) / (float)w.cx;
    float wy = (w.upper[1] - w.lower[1]) / (float)w.cy;
    float wz = (w.upper[2] - w.lower[2]) / (float)w.cz;
    int32_t rx = sat_i32(std::floor((mx - w.lower[0]) / wx));
    int32_t ry = sat_i32(std::floor((my - w.lower[1]) / wy));
    int32_t rz = sat_i32(std::floor((mz - w.lower[2]) / wz));
    int32_t nx = (int32_t)w.cx, ny = (int32_t)w.cy, nz = (int32_t)w.cz;
    x = (uint32_t)std::min(std::max(rx, 0), nx - 1);
    y = (uint32_t)std::min(std::max(ry, 0), ny - 1);
    z = (uint32_t)std::min(std::max(rz, 0), nz - 1);
    if (w.periodic[0]) x = (uint32_t)(((rx % nx) + nx) % nx);
    if (w.periodic[1]) y = (uint32_t)(((ry % ny) + ny) % ny);
    if (w.periodic[2]) z = (uint32_t)(((rz % nz) + nz) % nz);
}

inline uint32_t segment_cell(const World &w, const float *P, uint32_t s) {
    uint32_t a = w.seg_vtx[2 * s], b = w.seg_vtx[2 * s + 1];
    uint32_t x, y, z;
    cell_coords(w, P[3 * a], P[3 * a + 1], P[3 * a + 2], P[3 * b], P[3 * b + 1], P[3 * b + 2], x, y, z);
    return (z * w.cy + y) * w.cx + x;
}

// count, exclusive scan, scatter, rank: the ranked result puts each cell's
// segments in ascending order, which a stable counting sort over ascending
// segment indices reproduces directly.
void rebuild_cell_list(World &w) {
    w.cell_counts.assign(w.C, 0);
    w.cell_offsets.assign(w.C, 0);
    w.cell_segments.assign(w.S, 0);
    std::vector<uint32_t> cell_of(w.S);
    for (uint32_t s = 0; s < w.S; ++s) {
        if (!seg_on(w, s)) continue;  // only active segments are binned
        cell_of[s] = segment_cell(w, w.pos.data(), s);
        w.cell_counts[cell_of[s]]++;
    }
    uint32_t run = 0;
    for (uint32_t c = 0; c < w.C; ++c) {
        w.cell_offsets[c] = run;
        run += w.cell_counts[c];
    }
    std::vector<uint32_t> cursor(w.C, 0);
    for (uint32_t s = 0; s < w.S; ++s) {
        if (!seg_on(w, s)) continue;
        uint32_t c = cell_of[s];
        w.cell_segments[w.cell_offsets[c] + cursor[c]++] = s;
    }
}

// ---- neighbour lists: neighbor.rs:25-86, 155-408 ---------------------------

inline float along_fiber_gap(const World &w, uint32_t a0, uint32_t a1, uint32_t b0, uint32_t b1) {
    float after = w.arc[b0] - w.arc[a1];
    float before = w.arc[a0] - w.arc[b1];
    return rmax(after, before);
}

// proxy_pair_axis_distance with one proxy per segment (the canonical check
// always holds because proxy_of(s, 1) == 0 for s in [0, 1]).
float axis_distance(float p1x, float p1y, float p1z, float d1x, float d1y, float d1z, float p2x,
                    float p2y, float p2z, float d2x, float d2y, float d2z) {
    const float eps_len = 1.0e-20f, eps_par = 1.0e-6f;
    float rx = p1x - p2x, ry = p1y - p2y, rz = p1z - p2z;
    float a = d1x * d1x + d1y * d1y + d1z * d1z;
    float e = d2x * d2x + d2y * d2y + d2z * d2z;
    float f = d2x * rx + d2y * ry + d2z * rz;
    float s = 0.f, t = 0.f;
    if (a <= eps_len && e > eps_len) {
        t = rclamp(f / e, 0.f, 1.f);
    } else if (a > eps_len) {
        float c = d1x * rx + d1y * ry + d1z * rz;
        if (e <= eps_len) {
            s = rclamp(-c / a, 0.f, 1.f);
        } else {
            float b = d1x * d2x + d1y * d2y + d1z * d2z;
            float den = a * e - b * b;
            if (std::fabs(den) > eps_par * a * e) s = rclamp((b * f - c * e) / den, 0.f, 1.f);
            float proj = (b * s + f) / e;
            if (proj < 0.f) {
                s = rclamp(-c / a, 0.f, 1.f);
            } else if (proj > 1.f) {
                t = 1.f;
                s = rclamp((b - c) / a, 0.f, 1.f);
            } else {
                t = proj;
            }
        }
    }
    float dx = (p2x + d2x * t) - (p1x + d1x * s);
    float dy = (p2y + d2y * t) - (p1y + d1y * s);
    float dz = (p2z + d2z * t) - (p1z + d1z * s);
    return std::sqrt(dx * dx + dy * dy + dz * dz);
}

// The 27-cell stencil of a home cell with periodic de-duplication
// (neighbor.rs:242-276, contact.rs:393-428). Returns false for a skipped cell.
inline bool stencil_cell(const World &w, uint32_t home, uint32_t n, uint32_t &cell) {
    uint32_t hx = home % w.cx, hy = (home / w.cx) % w.cy, hz = home / (w.cx * w.cy);
    int32_t ox = (int32_t)(n % 3) - 1, oy = (int32_t)((n / 3) % 3) - 1, oz = (int32_t)(n / 9) - 1;
    int32_t rx = (int32_t)hx + ox, ry = (int32_t)hy + oy, rz = (int32_t)hz + oz;
    int32_t nx = (int32_t)w.cx, ny = (int32_t)w.cy, nz = (int32_t)w.cz;
    bool vx = w.periodic[0] || (rx >= 0 && rx < nx);
    bool vy = w.periodic[1] || (ry >= 0 && ry < ny);
    bool vz = w.periodic[2] || (rz >= 0 && rz < nz);
    bool ux = !w.periodic[0] || !((w.cx == 1 && ox != 0) || (w.cx == 2 && ((hx == 0 && ox > 0) || (hx == 1 && ox < 0))));
    bool uy = !w.periodic[1] || !((w.cy == 1 && oy != 0) || (w.cy == 2 && ((hy == 0 && oy > 0) || (hy == 1 && oy < 0))));
    bool uz = !w.periodic[2] || !((w.cz == 1 && oz != 0) || (w.cz == 2 && ((hz == 0 && oz > 0) || (hz == 1 && oz < 0))));
    if (!(vx && vy && vz && ux && uy && uz)) return false;
    uint32_t x = (uint32_t)(((rx % nx) + nx) % nx);
    uint32_t y = (uint32_t)(((ry % ny) + ny) % ny);
    uint32_t z = (uint32_t)(((rz % nz) + nz) % nz);
    cell = (z * w.cy + y) * w.cx + x;
    return true;
}

void build_neighbor_lists(World &w) {
    const float len_x = w.upper[0] - w.lower[0], len_y = w.upper[1] - w.lower[1], len_z = w.upper[2] - w.lower[2];
    w.ncount.assign(w.S, 0);
    if (w.nlist.size() != w.slots) w.nlist.assign(w.slots, 0);
    const float *P = w.pos.data();
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t si = 0; si < (int64_t)w.S; ++si) {
        uint32_t seg = (uint32_t)si;
        if (!seg_on(w, seg)) continue;  // inactive: no list
        uint32_t v1 = w.seg_vtx[2 * seg], v2 = w.seg_vtx[2 * seg + 1];
        float p1x = P[3 * v1], p1y = P[3 * v1 + 1], p1z = P[3 * v1 + 2];
        float q1x = P[3 * v2], q1y = P[3 * v2 + 1], q1z = P[3 * v2 + 2];
        float d1x = q1x - p1x, d1y = q1y - p1y, d1z = q1z - p1z;
        uint32_t owner = w.seg_fiber[seg];
        float radius = w.seg_radius[seg];
        float half_length = 0.5f * std::sqrt(d1x * d1x + d1y * d1y + d1z * d1z);
        uint32_t home = segment_cell(w, P, seg);
        size_t start = (size_t)w.block * seg;
        uint32_t count = 0;
        for (uint32_t n = 0; n < 27; ++n) {
            uint32_t cell;
            if (!stencil_cell(w, home, n, cell)) continue;
            uint32_t cs = w.cell_offsets[cell], cn = w.cell_counts[cell];
            for (uint32_t k = 0; k < cn; ++k) {
                uint32_t other = w.cell_segments[cs + k];
                uint32_t v3 = w.seg_vtx[2 * other], v4 = w.seg_vtx[2 * other + 1];
                float gap = along_fiber_gap(w, v1, v2, v3, v4);
                bool near = w.seg_fiber[other] == owner && gap < radius + w.seg_radius[other];
                if (other == seg || near) continue;
                float p2x = P[3 * v3], p2y = P[3 * v3 + 1], p2z = P[3 * v3 + 2];
                float q2x = P[3 * v4], q2y = P[3 * v4 + 1], q2z = P[3 * v4 + 2];
                if (w.periodic[0]) {
                    float sh = std::round(0.5f * (p2x + q2x - p1x - q1x) / len_x) * len_x;
                    p2x -= sh;
                    q2x -= sh;
                }
                if (w.periodic[1]) {
                    float sh = std::round(0.5f * (p2y + q2y - p1y - q1y) / len_y) * len_y;
                    p2y -= sh;
                    q2y -= sh;
                }
                if (w.periodic[2]) {
                    float sh = std::round(0.5f * (p2z + q2z - p1z - q1z) / len_z) * len_z;
                    p2z -= sh;
                    q2z -= sh;
                }
                float d2x = q2x - p2x, d2y = q2y - p2y, d2z = q2z - p2z;
                float inter = (radius + w.seg_radius[other] + w.skin) * (1.0f + 1.0e-4f);
                float mdx = 0.5f * (p2x + q2x - p1x - q1x);
                float mdy = 0.5f * (p2y + q2y - p1y - q1y);
                float mdz = 0.5f * (p2z + q2z - p1z - q1z);
                float reach = half_length + 0.5f * std::sqrt(d2x * d2x + d2y * d2y + d2z * d2z) + inter;
                if (mdx * mdx + mdy * mdy + mdz * mdz <= reach * reach &&
                    axis_distance(p1x, p1y, p1z, d1x, d1y, d1z, p2x, p2y, p2z, d2x, d2y, d2z) <= inter) {
                    if (count < w.block) w.nlist[start + count] = other;
                    ++count;
                }
            }
        }
        w.ncount[seg] = count;
        if (count <= w.block) std::sort(w.nlist.begin() + start, w.nlist.begin() + start + count);
    }
}

void rebuild_if_requested(World &w) {  // world/neighbor_list.rs:19-62
    if (w.nstate[0] == 0) return;
    rebuild_cell_list(w);
    build_neighbor_lists(w);
    w.refpos = w.pos;
    w.nstate[0] = 0;
    w.nstate[1] += 1;
}

void flag_displacement(World &w) {  // neighbor.rs:443-465
    float m = 0.5f * w.skin;
    for (uint32_t v = 0; v < w.V; ++v) {
        float dx = w.pos[3 * v] - w.refpos[3 * v];
        float dy = w.pos[3 * v + 1] - w.refpos[3 * v + 1];
        float dz = w.pos[3 * v + 2] - w.refpos[3 * v + 2];
        if (dx * dx + dy * dy + dz * dz > m * m) {
            w.nstate[0] = 1;
            return;
        }
    }
}

// ---- contact: contact.rs:232-902 (round sections, one lane each) -----------

void find_segment_corrections(World &w, const tgh_config *c) {
    const float len_x = w.upper[0] - w.lower[0], len_y = w.upper[1] - w.lower[1], len_z = w.upper[2] - w.lower[2];
    const float *P = w.pos.data();
    const float cf = c->correction_fraction;
    const uint32_t agg = c->contact_aggregation;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t si = 0; si < (int64_t)w.S; ++si) {
        uint32_t seg = (uint32_t)si;
        if (!seg_on(w, seg)) {  // inactive segments take no contact
            std::fill(w.corr.begin() + 6 * (size_t)seg, w.corr.begin() + 6 * (size_t)seg + 6, 0.f);
            w.seg_max[seg] = 0.f;
            continue;
        }
        uint32_t v1 = w.seg_vtx[2 * seg], v2 = w.seg_vtx[2 * seg + 1];
        float p1x = P[3 * v1], p1y = P[3 * v1 + 1], p1z = P[3 * v1 + 2];
        float q1x = P[3 * v2], q1y = P[3 * v2 + 1], q1z = P[3 * v2 + 2];
        float d1x = q1x - p1x, d1y = q1y - p1y, d1z = q1z - p1z;
        uint32_t owner = w.seg_fiber[seg];
        float radius = w.seg_radius[seg];
        float half_length = 0.5f * std::sqrt(d1x * d1x + d1y * d1y + d1z * d1z);
        const float eps_len = 1.0e-20f, eps_par = 1.0e-6f;
        float c0x = 0, c0y = 0, c0z = 0, c1x = 0, c1y = 0, c1z = 0;
        float max_pen = 0.f, cweight = 0.f;
        uint32_t listed_count = w.ncount[seg];
        bool listed = listed_count <= w.block;
        uint32_t ranges = listed ? 1 : 27;
        uint32_t home = 0;
        if (!listed) {
            const float *R = w.refpos.data();
            uint32_t x, y, z;
            cell_coords(w, R[3 * v1], R[3 * v1 + 1], R[3 * v1 + 2], R[3 * v2], R[3 * v2 + 1], R[3 * v2 + 2], x, y, z);
            home = (z * w.cy + y) * w.cx + x;
        }
        for (uint32_t range = 0; range < ranges; ++range) {
            uint32_t count = 0;
            size_t start = 0;
            if (listed) {
                count = listed_count;
                start = (size_t)w.block * seg;
            } else {
                uint32_t cell;
                if (stencil_cell(w, home, range, cell)) {
                    count = w.cell_counts[cell];
                    start = w.cell_offsets[cell];
                }
            }
            for (uint32_t k = 0; k < count; ++k) {
                uint32_t other = listed ? w.nlist[start + k] : w.cell_segments[start + k];
                uint32_t v3 = w.seg_vtx[2 * other], v4 = w.seg_vtx[2 * other + 1];
                float gap = along_fiber_gap(w, v1, v2, v3, v4);
                bool near = w.seg_fiber[other] == owner && gap < radius + w.seg_radius[other];
                if (other == seg || near) continue;
                float p2x = P[3 * v3], p2y = P[3 * v3 + 1], p2z = P[3 * v3 + 2];
                float q2x = P[3 * v4], q2y = P[3 * v4 + 1], q2z = P[3 * v4 + 2];
                float idx = 0.5f * (p2x + q2x - p1x - q1x);
                float idy = 0.5f * (p2y + q2y - p1y - q1y);
                float idz = 0.5f * (p2z + q2z - p1z - q1z);
                if (w.periodic[0]) {
                    float sh = std::round(idx / len_x) * len_x;
                    p2x -= sh;
                    q2x -= sh;
                }
                if (w.periodic[1]) {
                    float sh = std::round(idy / len_y) * len_y;
                    p2y -= sh;
                    q2y -= sh;
                }
                if (w.periodic[2]) {
                    float sh = std::round(idz / len_z) * len_z;
                    p2z -= sh;
                    q2z -= sh;
                }
                float d2x = q2x - p2x, d2y = q2y - p2y, d2z = q2z - p2z;
                float mdx = 0.5f * (p2x + q2x - p1x - q1x);
                float mdy = 0.5f * (p2y + q2y - p1y - q1y);
                float mdz = 0.5f * (p2z + q2z - p1z - q1z);
                float other_radius = w.seg_radius[other];
                float reach = (half_length + 0.5f * std::sqrt(d2x * d2x + d2y * d2y + d2z * d2z) + radius +
                               other_radius) * (1.0f + 1.0e-4f);
                if (!(mdx * mdx + mdy * mdy + mdz * mdz <= reach * reach)) continue;
                // exact closest points (contact.rs:653-728)
                float rx = p1x - p2x, ry = p1y - p2y, rz = p1z - p2z;
                float a = d1x * d1x + d1y * d1y + d1z * d1z;
                float e = d2x * d2x + d2y * d2y + d2z * d2z;
                float f = d2x * rx + d2y * ry + d2z * rz;
                float s = 0.f, t = 0.f;
                if (a <= eps_len && e > eps_len) {
                    t = rclamp(f / e, 0.f, 1.f);
                } else if (a > eps_len) {
                    float cc = d1x * rx + d1y * ry + d1z * rz;
                    if (e <= eps_len) {
                        s = rclamp(-cc / a, 0.f, 1.f);
                    } else {
                        float b = d1x * d2x + d1y * d2y + d1z * d2z;
                        float den = a * e - b * b;
                        bool parallel = std::fabs(den) <= eps_par * a * e;
                        if (!parallel) s = rclamp((b * f - cc * e) / den, 0.f, 1.f);
                        float proj = (b * s + f) / e;
                        if (proj < 0.f) {
                            s = rclamp(-cc / a, 0.f, 1.f);
                        } else if (proj > 1.f) {
                            t = 1.f;
                            s = rclamp((b - cc) / a, 0.f, 1.f);
                        } else {
                            t = proj;
                        }
                        if (parallel) {
                            float es = 1.f, et = 0.f;
                            float eproj = (b + f) / e;
                            if (eproj < 0.f) {
                                es = rclamp(-cc / a, 0.f, 1.f);
                            } else if (eproj > 1.f) {
                                et = 1.f;
                                es = rclamp((b - cc) / a, 0.f, 1.f);
                            } else {
                                et = eproj;
                            }
                            float gx = rx + d1x * s - d2x * t, gy = ry + d1y * s - d2y * t, gz = rz + d1z * s - d2z * t;
                            float ex = rx + d1x * es - d2x * et, ey = ry + d1y * es - d2y * et, ez = rz + d1z * es - d2z * et;
                            if (ex * ex + ey * ey + ez * ez < gx * gx + gy * gy + gz * gz) {
                                s = es;
                                t = et;
                            }
                        }
                    }
                }
                float x1 = p1x + d1x * s, y1 = p1y + d1y * s, z1 = p1z + d1z * s;
                float x2 = p2x + d2x * t, y2 = p2y + d2y * t, z2 = p2z + d2z * t;
                float dx = x2 - x1, dy = y2 - y1, dz = z2 - z1;
                float distance = std::sqrt(dx * dx + dy * dy + dz * dz);
                float pen = radius + other_radius - distance;
                if (!(pen > 0.f)) continue;
                bool deepest = pen > max_pen;
                if (deepest) max_pen = pen;
                float nx, ny, nz;
                if (distance > 1.0e-7f) {
                    float inv = 1.0f / distance;
                    nx = dx * inv;
                    ny = dy * inv;
                    nz = dz * inv;
                } else {  // contact.rs:768-812
                    nx = d1y * d2z - d1z * d2y;
                    ny = d1z * d2x - d1x * d2z;
                    nz = d1x * d2y - d1y * d2x;
                    float nl = std::sqrt(nx * nx + ny * ny + nz * nz);
                    if (nl <= 1.0e-7f) {
                        if (std::fabs(d1x) <= std::fabs(d1y) && std::fabs(d1x) <= std::fabs(d1z)) {
                            nx = 0.f; ny = d1z; nz = -d1y;
                        } else if (std::fabs(d1y) <= std::fabs(d1z)) {
                            nx = -d1z; ny = 0.f; nz = d1x;
                        } else {
                            nx = d1y; ny = -d1x; nz = 0.f;
                        }
                        nl = std::sqrt(nx * nx + ny * ny + nz * nz);
                    }
                    if (nl > 1.0e-7f) {
                        nx /= nl;
                        ny /= nl;
                        nz /= nl;
                    } else {
                        nx = 1.0f;
                    }
                    float dom = (std::fabs(nx) >= std::fabs(ny) && std::fabs(nx) >= std::fabs(nz))
                                    ? nx : (std::fabs(ny) >= std::fabs(nz) ? ny : nz);
                    if (dom < 0.f) { nx = -nx; ny = -ny; nz = -nz; }
                    if (seg > other) { nx = -nx; ny = -ny; nz = -nz; }
                }
                float w0 = 1.0f - s, w1 = s, ow0 = 1.0f - t;
                // round sections: every director term is exactly zero (contact.rs:835-852)
                float den = w0 * w0 + w1 * w1 + ow0 * ow0 + t * t + 0.0f;
                float mag = -cf * pen / den;
                float aw = agg == 1 ? pen : 1.0f;
                if (agg == 2) {
                    if (deepest) {
                        c0x = nx * mag * w0; c0y = ny * mag * w0; c0z = nz * mag * w0;
                        c1x = nx * mag * w1; c1y = ny * mag * w1; c1z = nz * mag * w1;
                        cweight = 1.0f;
                    }
                } else {
                    c0x += nx * mag * w0 * aw; c0y += ny * mag * w0 * aw; c0z += nz * mag * w0 * aw;
                    c1x += nx * mag * w1 * aw; c1y += ny * mag * w1 * aw; c1z += nz * mag * w1 * aw;
                    cweight += aw;
                }
            }
        }
        if (agg != 2 && cweight > 1.0e-20f) {
            c0x /= cweight; c0y /= cweight; c0z /= cweight;
            c1x /= cweight; c1y /= cweight; c1z /= cweight;
        }
        float *o = &w.corr[6 * (size_t)seg];
        o[0] = c0x; o[1] = c0y; o[2] = c0z; o[3] = c1x; o[4] = c1y; o[5] = c1z;
        w.seg_max[seg] = max_pen;
    }
}

// ---- curvature ratio: kernels.rs:925-980 -----------------------------------

void measure_curvature_ratio(World &w) {
    const float *P = w.pos.data();
#pragma omp parallel for schedule(static)
    for (int64_t vi = 0; vi < (int64_t)w.V; ++vi) {
        uint32_t v = (uint32_t)vi;
        float kmax = w.vtx_kmax[v];
        float ratio = 0.f;
        uint32_t ps = w.vtx_seg[2 * v], ns = w.vtx_seg[2 * v + 1];
        if (kmax > 0.f && ps != kNone && ns != kNone && vtx_on(w, v)) {
            uint32_t pa = w.seg_vtx[2 * ps], pb = w.seg_vtx[2 * ps + 1];
            uint32_t prev = pa == v ? pb : pa;
            uint32_t na = w.seg_vtx[2 * ns], nb = w.seg_vtx[2 * ns + 1];
            uint32_t next = na == v ? nb : na;
            float ax = P[3 * v] - P[3 * prev], ay = P[3 * v + 1] - P[3 * prev + 1], az = P[3 * v + 2] - P[3 * prev + 2];
            float bx = P[3 * next] - P[3 * v], by = P[3 * next + 1] - P[3 * v + 1], bz = P[3 * next + 2] - P[3 * v + 2];
            float l1 = std::sqrt(ax * ax + ay * ay + az * az);
            float l2 = std::sqrt(bx * bx + by * by + bz * bz);
            if (l1 > 1.0e-7f && l2 > 1.0e-7f) {
                float cosine = (ax * bx + ay * by + az * bz) / (l1 * l2);
                cosine = rclamp(cosine, -1.f, 1.f);
                float dual = 0.5f * (l1 + l2);
                float kappa = 2.0f * std::sqrt(0.5f * (1.0f - cosine)) / dual;
                ratio = kappa / kmax;
            }
        }
        w.curv[v] = ratio;
    }
}

// ---- reductions + assess: kernels.rs:86-162 --------------------------------

void reduce_and_assess(World &w, const tgh_config *c, uint32_t iterations) {
    uint32_t r0 = 0, r1 = 0, r2 = 0;
    for (uint32_t s = 0; s < w.S; ++s) r0 = std::max(r0, f2b(rmax(w.seg_max[s], 0.f)));
    for (uint32_t v = 0; v < w.V; ++v) {
        r1 = std::max(r1, f2b(rmax(w.vstep[v], 0.f)));
        r2 = std::max(r2, f2b(rmax(w.curv[v], 0.f)));
    }
    if (w.control[0] == 0) return;
    float pen = b2f(r0), curv = b2f(r2);
    w.metrics[0] = pen;
    w.metrics[1] = b2f(r1);
    w.metrics[2] = curv;
    bool overflow = false;  // compact cell lists never overflow (world.rs:105-106)
    if (overflow || (c->force_full_iterations == 0 && pen <= c->penetration_tolerance &&
                     curv <= 1.0f + c->curvature_ratio_tolerance) ||
        w.control[2] >= iterations) {
        w.control[0] = 0;
    } else {
        w.control[1] += 1;
        w.control[2] += 1;
    }
}

// ---- internal constraints: kernels.rs:1027-1163, 1417-1453 ------------------

void find_internal(World &w, const tgh_config *c) {
    const float *P = w.pos.data();
    const float ks = c->stretch_stiffness, kb = c->bend_stiffness;
#pragma omp parallel for schedule(static)
    for (int64_t vi = 0; vi < (int64_t)w.V; ++vi) {
        uint32_t v = (uint32_t)vi;
        float *o = &w.icorr[3 * (size_t)v];
        if (w.vtx_pinned[v] || !vtx_on(w, v)) {
            o[0] = o[1] = o[2] = 0.f;
            continue;
        }
        float x = P[3 * v], y = P[3 * v + 1], z = P[3 * v + 2];
        float cx = 0.f, cy = 0.f, cz = 0.f;
        uint32_t ps = w.vtx_seg[2 * v], ns = w.vtx_seg[2 * v + 1];
        for (int side = 0; side < 2; ++side) {  // stretch: previous then next segment
            uint32_t sg = side == 0 ? ps : ns;
            if (sg == kNone) continue;
            uint32_t a = w.seg_vtx[2 * sg], b = w.seg_vtx[2 * sg + 1];
            uint32_t other = a == v ? b : a;
            float dx = P[3 * other] - x, dy = P[3 * other + 1] - y, dz = P[3 * other + 2] - z;
            float len = std::sqrt(dx * dx + dy * dy + dz * dz);
            if (len > 1.0e-7f) {
                float mag = 0.5f * ks * (len - w.seg_rest[sg]) / len;
                cx += dx * mag;
                cy += dy * mag;
                cz += dz * mag;
            }
        }
        for (int side = 0; side < 2; ++side) {  // rest chord to the second neighbours
            uint32_t sg = side == 0 ? ps : ns;
            if (sg == kNone) continue;
            uint32_t a = w.seg_vtx[2 * sg], b = w.seg_vtx[2 * sg + 1];
            uint32_t middle = a == v ? b : a;
            uint32_t outer = w.vtx_seg[2 * middle + (side == 0 ? 0 : 1)];
            if (outer == kNone) continue;
            uint32_t oa = w.seg_vtx[2 * outer], ob = w.seg_vtx[2 * outer + 1];
            uint32_t other = oa == middle ? ob : oa;
            float dx = P[3 * other] - x, dy = P[3 * other + 1] - y, dz = P[3 * other + 2] - z;
            float len = std::sqrt(dx * dx + dy * dy + dz * dz);
            if (len > 1.0e-7f) {
                const float *I = w.intr.data();
                float rx = I[3 * other] - I[3 * v], ry = I[3 * other + 1] - I[3 * v + 1], rz = I[3 * other + 2] - I[3 * v + 2];
                float chord = std::sqrt(rx * rx + ry * ry + rz * rz);
                float mag = 0.5f * kb * (len - chord) / len;
                cx += dx * mag;
                cy += dy * mag;
                cz += dz * mag;
            }
        }
        o[0] = cx;
        o[1] = cy;
        o[2] = cz;
    }
}

void apply_internal(World &w) {
#pragma omp parallel for schedule(static)
    for (int64_t vi = 0; vi < (int64_t)w.V; ++vi) {
        uint32_t v = (uint32_t)vi;
        if (w.vtx_pinned[v] || !vtx_on(w, v)) continue;
        for (int k = 0; k < 3; ++k) {
            size_t i = 3 * (size_t)v + k;
            float r = w.wall_extent[i];
            float moved = w.pos[i] + w.icorr[i];
            w.pos[i] = w.periodic[k] ? moved : rclamp(moved, w.lower[k] + r, w.upper[k] - r);
            if (!w.periodic[k]) w.wall[i] += moved - w.pos[i];
        }
    }
}

void apply_contact(World &w, const tgh_config *c) {  // kernels.rs:1455-1537
    const float ms = c->max_step;
#pragma omp parallel for schedule(static)
    for (int64_t vi = 0; vi < (int64_t)w.V; ++vi) {
        uint32_t v = (uint32_t)vi;
        if (w.vtx_pinned[v] || !vtx_on(w, v)) {
            w.vstep[v] = 0.f;
            continue;
        }
        uint32_t ps = w.vtx_seg[2 * v], ns = w.vtx_seg[2 * v + 1];
        float cx = 0.f, cy = 0.f, cz = 0.f;
        if (ps != kNone) {
            cx += w.corr[6 * (size_t)ps + 3];
            cy += w.corr[6 * (size_t)ps + 4];
            cz += w.corr[6 * (size_t)ps + 5];
        }
        if (ns != kNone) {
            cx += w.corr[6 * (size_t)ns];
            cy += w.corr[6 * (size_t)ns + 1];
            cz += w.corr[6 * (size_t)ns + 2];
        }
        float len = std::sqrt(cx * cx + cy * cy + cz * cz);
        if (len > ms) {
            float sc = ms / len;
            cx *= sc;
            cy *= sc;
            cz *= sc;
        }
        float d[3] = {cx, cy, cz};
        for (int k = 0; k < 3; ++k) {
            size_t i = 3 * (size_t)v + k;
            float r = w.wall_extent[i];
            float moved = w.pos[i] + d[k];
            w.pos[i] = w.periodic[k] ? moved : rclamp(moved, w.lower[k] + r, w.upper[k] - r);
            if (!w.periodic[k]) w.wall[i] += moved - w.pos[i];
        }
        w.vstep[v] = std::sqrt(cx * cx + cy * cy + cz * cz);
    }
}

// ---- curvature Gauss–Seidel: kernels.rs:1165-1415 --------------------------

// Diagnostics of the curvature cleanup (projections applied, sweeps run).
uint64_t g_gs_projections = 0, g_gs_calls = 0, g_gs_fiber_sweeps = 0, g_gs_max_fiber_projections = 0;

bool project_triplet(World &w, uint32_t prev, uint32_t center, uint32_t next, float stiffness, float margin) {
    float *P = w.pos.data();
    float kmax = w.vtx_kmax[center] * (1.0f - margin);
    if (!(kmax > 0.f)) return false;
    float ux = P[3 * center] - P[3 * prev], uy = P[3 * center + 1] - P[3 * prev + 1], uz = P[3 * center + 2] - P[3 * prev + 2];
    float vx = P[3 * next] - P[3 * center], vy = P[3 * next + 1] - P[3 * center + 1], vz = P[3 * next + 2] - P[3 * center + 2];
    float l1 = std::sqrt(ux * ux + uy * uy + uz * uz);
    float l2 = std::sqrt(vx * vx + vy * vy + vz * vz);
    if (!(l1 > 1.0e-7f && l2 > 1.0e-7f)) return false;
    float unx = ux / l1, uny = uy / l1, unz = uz / l1;
    float vnx = vx / l2, vny = vy / l2, vnz = vz / l2;
    float cosine = rclamp(unx * vnx + uny * vny + unz * vnz, -1.f, 1.f);
    float sine_half = 0.25f * kmax * (l1 + l2);
    if (!(sine_half < 1.0f)) return false;
    float min_cos = 1.0f - 2.0f * sine_half * sine_half;
    float violation = min_cos - cosine;
    float sine = std::sqrt(rmax(1.0f - cosine * cosine, 0.f));
    if (violation > 0.f && sine > 1.0e-3f) violation = (std::acos(cosine) - std::acos(min_cos)) * sine;
    if (!(violation > 0.f)) return false;
    float gux = (vnx - cosine * unx) / l1, guy = (vny - cosine * uny) / l1, guz = (vnz - cosine * unz) / l1;
    float gvx = (unx - cosine * vnx) / l2, gvy = (uny - cosine * vny) / l2, gvz = (unz - cosine * vnz) / l2;
    float gcx = gux - gvx, gcy = guy - gvy, gcz = guz - gvz;
    float wp = w.vtx_pinned[prev] ? 0.f : 1.f;
    float wc = w.vtx_pinned[center] ? 0.f : 1.f;
    float wn = w.vtx_pinned[next] ? 0.f : 1.f;
    float den = wp * (gux * gux + guy * guy + guz * guz) + wc * (gcx * gcx + gcy * gcy + gcz * gcz) +
                wn * (gvx * gvx + gvy * gvy + gvz * gvz);
    if (!(den > 1.0e-12f)) return false;
    float scale = stiffness * violation / den;
    float gp[3] = {gux, guy, guz}, gc[3] = {gcx, gcy, gcz}, gn[3] = {gvx, gvy, gvz};
    for (int k = 0; k < 3; ++k) {
        size_t ip = 3 * (size_t)prev + k, ic = 3 * (size_t)center + k, in = 3 * (size_t)next + k;
        float mp = P[ip] - wp * scale * gp[k];
        float mc = P[ic] + wc * scale * gc[k];
        float mn = P[in] + wn * scale * gn[k];
        float vp = mp, vc = mc, vn = mn;
        if (!w.periodic[k]) {
            vp = rclamp(mp, w.lower[k] + w.wall_extent[ip], w.upper[k] - w.wall_extent[ip]);
            vc = rclamp(mc, w.lower[k] + w.wall_extent[ic], w.upper[k] - w.wall_extent[ic]);
            vn = rclamp(mn, w.lower[k] + w.wall_extent[in], w.upper[k] - w.wall_extent[in]);
            if (wp > 0.f) w.wall[ip] += mp - vp;
            if (wc > 0.f) w.wall[ic] += mc - vc;
            if (wn > 0.f) w.wall[in] += mn - vn;
        }
        if (wp > 0.f) P[ip] = vp;
        if (wc > 0.f) P[ic] = vc;
        if (wn > 0.f) P[in] = vn;
    }
    return true;
}

void curvature_cleanup(World &w, const tgh_config *c) {
    ++g_gs_calls;
    const float st = c->curvature_limit_stiffness, mg = c->curvature_limit_safety_margin;
    w.gs_projections.assign(w.F, 0u);
    w.gs_visits.assign(w.F, 0u);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t fi = 0; fi < (int64_t)w.F; ++fi) {
        uint32_t start = w.fiber_span[2 * fi], count = w.fiber_span[2 * fi + 1];
        if (count < 3 || !w.fiber_active[fi]) continue;
        uint64_t projections = 0, sweeps_run = 0, visits = 0;
        bool visiting = true;
        for (uint32_t sweep = 0; sweep < c->curvature_cleanup_sweeps; ++sweep) {
            uint64_t forward = 0;
            for (uint32_t v = start + 2; v < start + count; ++v) forward += project_triplet(w, v - 2, v - 1, v, st, mg);
            if (visiting) visits += count - 2;
            if (visiting && forward != 0) visits += count - 2;  // the backward pass
            for (uint32_t l = 2; l < count; ++l) {
                uint32_t v = start + count - 1 - l;
                projections += project_triplet(w, v, v + 1, v + 2, st, mg);
            }
            projections += forward;
            if (forward != 0) ++sweeps_run;  // sweeps the GPU kernel's early exit still runs
            if (forward == 0) visiting = false;
        }
        w.gs_projections[fi] = (uint32_t)projections;
        w.gs_visits[fi] = (uint32_t)visits;
#pragma omp atomic
        g_gs_projections += projections;
#pragma omp atomic
        g_gs_fiber_sweeps += sweeps_run;
#pragma omp critical
        g_gs_max_fiber_projections = std::max(g_gs_max_fiber_projections, projections);
    }
}

// ---- formation: kernels.rs:164-221, 538-710 --------------------------------

void activate(World &w, uint32_t maximum_step) {  // activate_formation_step
    for (uint32_t f = 0; f < w.F; ++f) w.fiber_active[f] = w.fiber_step[f] > maximum_step ? 0u : 1u;
    w.nstate[0] = 1;  // request_neighbor_list_rebuild (world.rs:2176-2178)
}

float fiber_center(const World &w, uint32_t f, uint32_t axis) {
    uint32_t start = w.fiber_span[2 * f], count = w.fiber_span[2 * f + 1];
    float center = 0.0f;
    for (uint32_t i = 0; i < count; ++i) center += w.pos[3 * (size_t)(start + i) + axis];
    return center / (float)count;
}

bool layer_held(const World &w, uint32_t f) {
    uint32_t layer = w.fiber_layer[f];
    return !(layer == 0xffffffffu || layer >= w.layer_targets.size() || layer < w.first_layer ||
             layer > w.last_layer) &&
           w.fiber_span[2 * f + 1] > 0 && w.fiber_active[f];
}

void apply_layer_targets(World &w) {  // apply_formation_layer_targets (kernels.rs:538-580)
    if (w.layer_targets.empty()) return;
    for (uint32_t f = 0; f < w.F; ++f) {
        if (!layer_held(w, f)) continue;
        float center = fiber_center(w, f, w.layer_axis);
        float t = rclamp((w.layer_targets[w.fiber_layer[f]] - center) * w.layer_stiffness, -w.layer_max, w.layer_max);
        uint32_t start = w.fiber_span[2 * f], count = w.fiber_span[2 * f + 1];
        for (uint32_t i = 0; i < count; ++i) w.pos[3 * (size_t)(start + i) + w.layer_axis] += t;
    }
}

void apply_vertex_targets(World &w) {  // apply_vertex_targets (kernels.rs:605-635)
    for (size_t i = 0; i < w.vt_index.size(); ++i) {
        uint32_t v = w.vt_index[i];
        if (!vtx_on(w, v)) continue;
        float c = w.vt_command[i];
        float t = rclamp((w.vt_target[i] - c) * w.vt_stiffness, -w.vt_max, w.vt_max);
        w.vt_command[i] = c + t;
        w.pos[3 * (size_t)v + w.vt_axis] = c + t;
    }
}

// ---- run_batch: world.rs:1020-1505; read_status: world.rs:2337-2369 --------

int run_batch(World &w, const tgh_config *c, uint32_t iterations, tgh_batch_status *st) {
    if (iterations == 0) return fail(TGH_ERR_INVALID, "iterations must be positive");
    w.control[2] = 0;  // begin_relaxation_batch
    if (w.control[3] == 0) w.control[0] = 1;
    flag_displacement(w);
    for (uint32_t step = 0; step <= iterations; ++step) {
        rebuild_if_requested(w);
        bool running = w.control[0] != 0;
        if (running) {
            find_segment_corrections(w, c);
            measure_curvature_ratio(w);
        }
        reduce_and_assess(w, c, iterations);
        if (w.control[0]) {
            for (uint32_t v = 0; v < 3 * w.V; ++v) w.wall[v] = 0.f;
            for (uint32_t ci = 0; ci < c->constraint_iterations; ++ci) {
                find_internal(w, c);
                apply_internal(w);
            }
            apply_contact(w, c);
            curvature_cleanup(w, c);
            if (step < iterations) {  // manufacturing targets (world.rs:1420-1482)
                apply_layer_targets(w);
                apply_vertex_targets(w);
            }
        }
        flag_displacement(w);
    }
    st->total_iterations = w.control[1];
    st->batch_iterations = w.control[2];
    st->max_penetration = w.metrics[0];
    st->max_displacement = w.metrics[1];
    st->max_curvature_ratio = w.metrics[2];
    bool overflow = w.control[3] != 0;
    st->cell_list_overflow = overflow;
    st->converged = !c->force_full_iterations && w.metrics[0] <= c->penetration_tolerance &&
                    w.metrics[2] <= 1.0f + c->curvature_ratio_tolerance && !overflow;
    return TGH_OK;
}

// ---- compaction: kernels.rs:712-809, world.rs:1931-2053 ---------------------

int compact_cell(World &w, const float *nl, const float *nu, uint32_t kinematics) {
    for (int k = 0; k < 3; ++k) {
        if (!std::isfinite(nl[k]) || !std::isfinite(nu[k])) return fail(TGH_ERR_INVALID, "bounds must be finite");
        if (!(nu[k] - nl[k] > 2.0f * w.max_radius)) return fail(TGH_ERR_INVALID, "cell narrower than two radii");
        if (!(nl[k] >= w.lower[k]) || !(nu[k] <= w.upper[k])) return fail(TGH_ERR_INVALID, "cell outside the current");
    }
    uint32_t cells[3];
    for (int k = 0; k < 3; ++k) cells[k] = sat_u32(rmax(std::floor((nu[k] - nl[k]) / w.cell_size), 1.0f));
    if ((uint64_t)cells[0] * cells[1] * cells[2] > w.C) return fail(TGH_ERR_SIZE, "more cells than allocated");
    if (kinematics == 0) {  // compact_rigid_fiber_centers
        for (uint32_t f = 0; f < w.F; ++f) {
            const uint32_t start = w.fiber_span[2 * f], count = w.fiber_span[2 * f + 1];
            if (count == 0 || !vtx_on(w, start)) continue;
            float c[3] = {0.f, 0.f, 0.f};
            for (uint32_t i = 0; i < count; ++i)
                for (int k = 0; k < 3; ++k) c[k] += w.pos[3 * (size_t)(start + i) + k];
            for (int k = 0; k < 3; ++k) c[k] /= (float)count;
            float d[3];
            for (int k = 0; k < 3; ++k) {
                float fraction = (c[k] - w.lower[k]) / (w.upper[k] - w.lower[k]);
                d[k] = nl[k] + fraction * (nu[k] - nl[k]) - c[k];
            }
            for (uint32_t i = 0; i < count; ++i)
                for (int k = 0; k < 3; ++k) {
                    w.pos[3 * (size_t)(start + i) + k] += d[k];
                    w.wall[3 * (size_t)(start + i) + k] = 0.f;
                }
        }
    } else if (kinematics == 1) {  // compact_moving_walls
        for (uint32_t v = 0; v < w.V; ++v) {
            if (!vtx_on(w, v)) continue;
            for (int k = 0; k < 3; ++k) {
                const size_t i = 3 * (size_t)v + k;
                if (!w.periodic[k]) {
                    float previous = w.pos[i];
                    float projected = rclamp(previous, nl[k] + w.wall_extent[i], nu[k] - w.wall_extent[i]);
                    w.pos[i] = projected;
                    w.wall[i] = previous - projected;
                } else {
                    w.wall[i] = 0.f;
                }
            }
        }
    } else if (kinematics == 2) {  // compact_affine_vertices
        for (uint32_t v = 0; v < w.V; ++v) {
            if (!vtx_on(w, v)) continue;
            for (int k = 0; k < 3; ++k) {
                const size_t i = 3 * (size_t)v + k;
                float fraction = (w.pos[i] - w.lower[k]) / (w.upper[k] - w.lower[k]);
                w.pos[i] = nl[k] + fraction * (nu[k] - nl[k]);
                w.wall[i] = 0.f;
            }
        }
    } else {
        return fail(TGH_ERR_INVALID, "unknown kinematics");
    }
    for (int k = 0; k < 3; ++k) {
        w.lower[k] = nl[k];
        w.upper[k] = nu[k];
    }
    w.cx = cells[0];
    w.cy = cells[1];
    w.cz = cells[2];
    w.nstate[0] = 1;  // request_neighbor_list_rebuild
    return TGH_OK;
}

// measure_compaction_metrics (kernels.rs:811-923), one thread in index order
void compaction_metrics(const World &w, float cf, float kc, float ks, float kb, float *out) {
    float lx = 0.f, ly = 0.f, lz = 0.f, contact = 0.f, stretch = 0.f;
    for (uint32_t s = 0; s < w.S; ++s) {
        if (!seg_on(w, s)) continue;
        const float *c = &w.corr[6 * (size_t)s];
        lx += std::fabs(c[0]) + std::fabs(c[3]);
        ly += std::fabs(c[1]) + std::fabs(c[4]);
        lz += std::fabs(c[2]) + std::fabs(c[5]);
        float penetration = w.seg_max[s];
        contact += 0.5f * kc * penetration * penetration;
        const uint32_t a = w.seg_vtx[2 * s], b = w.seg_vtx[2 * s + 1];
        float dx = w.pos[3 * (size_t)b] - w.pos[3 * (size_t)a];
        float dy = w.pos[3 * (size_t)b + 1] - w.pos[3 * (size_t)a + 1];
        float dz = w.pos[3 * (size_t)b + 2] - w.pos[3 * (size_t)a + 2];
        float length = std::sqrt(dx * dx + dy * dy + dz * dz);
        float rest = w.seg_rest[s];
        if (rest > 1.0e-7f) {
            float extension = length - rest;
            stretch += 0.5f * ks * extension * extension / rest;
        }
    }
    float bending = 0.f;
    for (uint32_t v = 0; v < w.V; ++v) {
        float excess = rmax(w.curv[v] - 1.0f, 0.f);
        bending += 0.5f * kb * excess * excess;
    }
    float L[3];
    for (int k = 0; k < 3; ++k) L[k] = w.upper[k] - w.lower[k];
    float lo[3] = {0.f, 0.f, 0.f}, hi[3] = {0.f, 0.f, 0.f};
    for (uint32_t v = 0; v < w.V; ++v)
        for (int k = 0; k < 3; ++k) {
            float r = w.wall[3 * (size_t)v + k];
            lo[k] += rmax(-r, 0.f);
            hi[k] += rmax(r, 0.f);
        }
    float ip = 1.0f / rmax(cf, 1.0e-7f);
    float load[3] = {lx, ly, lz};
    float area[3] = {L[1] * L[2], L[0] * L[2], L[0] * L[1]};
    for (int k = 0; k < 3; ++k) {
        float pl = w.periodic[k] ? load[k] : lo[k] + hi[k];
        out[k] = kc * pl * ip / area[k];
        out[7 + 2 * k] = kc * lo[k] * ip / area[k];
        out[8 + 2 * k] = kc * hi[k] * ip / area[k];
    }
    out[3] = contact;
    out[4] = stretch;
    out[5] = bending;
    out[6] = contact + stretch + bending;
}

// capture_segment_contacts (contact.rs:9-230) on the whole-segment grid of
// capture_contacts (world.rs:2193-2334), then the host sort by (first,
// second, total order of the coordinates)
struct Candidate {
    uint32_t first, second;
    float s, t, gap, angle;
};

inline int32_t total_key(float x) {
    int32_t i;
    std::memcpy(&i, &x, 4);
    return i ^ (int32_t)((uint32_t)(i >> 31) >> 1);
}

int capture_contacts(World &w, float max_gap, size_t capacity, std::vector<Candidate> &out, bool &overflow) {
    if (!std::isfinite(max_gap) || !(max_gap >= 0.0f)) return fail(TGH_ERR_INVALID, "gap must be finite and >= 0");
    if (capacity == 0) return fail(TGH_ERR_INVALID, "capacity must be positive");
    const float size = rmax(w.segment_cell_size, w.cell_size) + max_gap;
    float L[3];
    uint32_t n[3];
    for (int k = 0; k < 3; ++k) {
        L[k] = w.upper[k] - w.lower[k];
        n[k] = sat_u32(rmax(std::floor(L[k] / size), 1.0f));
    }
    float width[3];
    for (int k = 0; k < 3; ++k) width[k] = L[k] / (float)n[k];
    auto home_of = [&](const float *p, const float *q, int32_t *h) {
        for (int k = 0; k < 3; ++k) {
            float mid = 0.5f * (p[k] + q[k]);
            int32_t raw = sat_i32(std::floor((mid - w.lower[k]) / width[k]));
            int32_t nk = (int32_t)n[k];
            h[k] = w.periodic[k] ? ((raw % nk) + nk) % nk : std::min(std::max(raw, 0), nk - 1);
        }
    };
    const uint64_t C = (uint64_t)n[0] * n[1] * n[2];
    std::vector<std::vector<uint32_t>> cells(C);
    for (uint32_t s = 0; s < w.S; ++s) {  // ascending within each cell (rank_cell_segments)
        if (!seg_on(w, s)) continue;
        int32_t h[3];
        home_of(&w.pos[3 * (size_t)w.seg_vtx[2 * s]], &w.pos[3 * (size_t)w.seg_vtx[2 * s + 1]], h);
        cells[((uint64_t)h[2] * n[1] + h[1]) * n[0] + h[0]].push_back(s);
    }
    std::vector<Candidate> found;
    const float eps_len = 1.0e-20f, eps_par = 1.0e-6f;
    for (uint32_t seg = 0; seg < w.S; ++seg) {
        if (!seg_on(w, seg)) continue;
        const float *P1 = &w.pos[3 * (size_t)w.seg_vtx[2 * seg]], *Q1 = &w.pos[3 * (size_t)w.seg_vtx[2 * seg + 1]];
        float d1[3] = {Q1[0] - P1[0], Q1[1] - P1[1], Q1[2] - P1[2]};
        int32_t h[3];
        home_of(P1, Q1, h);
        for (uint32_t nb = 0; nb < 27; ++nb) {
            int32_t o[3] = {(int32_t)(nb % 3) - 1, (int32_t)((nb / 3) % 3) - 1, (int32_t)(nb / 9) - 1};
            bool ok = true;
            int32_t raw[3], cell[3];
            float shift[3];
            for (int k = 0; k < 3; ++k) {
                int32_t nk = (int32_t)n[k];
                raw[k] = h[k] + o[k];
                bool valid = w.periodic[k] || (raw[k] >= 0 && raw[k] < nk);
                bool unique = !w.periodic[k] ||
                              !((n[k] == 1 && o[k] != 0) || (n[k] == 2 && ((h[k] == 0 && o[k] > 0) || (h[k] == 1 && o[k] < 0))));
                ok = ok && valid && unique;
                cell[k] = ((raw[k] % nk) + nk) % nk;
                shift[k] = raw[k] < 0 ? -L[k] : (raw[k] >= nk ? L[k] : 0.f);
            }
            if (!ok) continue;
            for (uint32_t other : cells[((uint64_t)cell[2] * n[1] + cell[1]) * n[0] + cell[0]]) {
                if (!(other > seg && w.seg_fiber[other] != w.seg_fiber[seg])) continue;
                const float *P2 = &w.pos[3 * (size_t)w.seg_vtx[2 * other]];
                const float *Q2 = &w.pos[3 * (size_t)w.seg_vtx[2 * other + 1]];
                float p2[3], q2[3];
                for (int k = 0; k < 3; ++k) {
                    p2[k] = P2[k] + shift[k];
                    q2[k] = Q2[k] + shift[k];
                }
                for (int k = 0; k < 3; ++k) {
                    float image = 0.5f * (p2[k] + q2[k] - P1[k] - Q1[k]);
                    if (w.periodic[k]) {
                        float sh = std::round(image / L[k]) * L[k];
                        p2[k] -= sh;
                        q2[k] -= sh;
                    }
                }
                float d2[3] = {q2[0] - p2[0], q2[1] - p2[1], q2[2] - p2[2]};
                float r[3] = {P1[0] - p2[0], P1[1] - p2[1], P1[2] - p2[2]};
                float a = d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2];
                float e = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2];
                float f = d2[0] * r[0] + d2[1] * r[1] + d2[2] * r[2];
                float sp = 0.f, tp = 0.f;
                if (a <= eps_len && e > eps_len) {
                    tp = rclamp(f / e, 0.f, 1.f);
                } else if (a > eps_len) {
                    float c = d1[0] * r[0] + d1[1] * r[1] + d1[2] * r[2];
                    if (e <= eps_len) {
                        sp = rclamp(-c / a, 0.f, 1.f);
                    } else {
                        float b = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2];
                        float den = a * e - b * b;
                        if (std::fabs(den) > eps_par * a * e) sp = rclamp((b * f - c * e) / den, 0.f, 1.f);
                        float proj = (b * sp + f) / e;
                        if (proj < 0.f) {
                            sp = rclamp(-c / a, 0.f, 1.f);
                        } else if (proj > 1.f) {
                            tp = 1.f;
                            sp = rclamp((b - c) / a, 0.f, 1.f);
                        } else {
                            tp = proj;
                        }
                    }
                }
                float dx = (p2[0] + d2[0] * tp) - (P1[0] + d1[0] * sp);
                float dy = (p2[1] + d2[1] * tp) - (P1[1] + d1[1] * sp);
                float dz = (p2[2] + d2[2] * tp) - (P1[2] + d1[2] * sp);
                float center = std::sqrt(dx * dx + dy * dy + dz * dz);
                float gap = center - w.seg_radius[seg] - w.seg_radius[other];
                if (!(gap <= max_gap)) continue;
                float product = std::sqrt(a * e);
                float angle = 0.f;
                if (product > eps_len)
                    angle = std::acos(rclamp(std::fabs((d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2]) / product), 0.f, 1.f));
                found.push_back(Candidate{seg, other, sp, tp, gap, angle});
            }
        }
    }
    overflow = found.size() > capacity;
    // the device keeps the first `capacity` appends in a run-dependent order;
    // which ones is unspecified when it overflows, so keep the sorted prefix
    std::sort(found.begin(), found.end(), [](const Candidate &l, const Candidate &r) {
        if (l.first != r.first) return l.first < r.first;
        if (l.second != r.second) return l.second < r.second;
        if (total_key(l.s) != total_key(r.s)) return total_key(l.s) < total_key(r.s);
        return total_key(l.t) < total_key(r.t);
    });
    if (found.size() > capacity) found.resize(capacity);
    out.swap(found);
    w.nstate[0] = 1;  // the shared grid held the capture layout
    return TGH_OK;
}

}  // namespace

extern "C" {

const char *orc_last_error(void) { return g_error.c_str(); }

int orc_world_create(const tgh_assembly *a, const tgh_config *c, void **out) {
    if (!out || !c) return fail(TGH_ERR_INVALID, "null argument");
    World *w = new World();
    int rc = pack(*w, a);
    if (rc == TGH_OK) rc = size_world(*w, c);
    if (rc != TGH_OK) {
        delete w;
        return rc;
    }
    w->metrics[0] = std::numeric_limits<float>::infinity();
    w->metrics[1] = 0.f;
    w->metrics[2] = std::numeric_limits<float>::infinity();
    w->cell_counts.assign(w->C, 0);
    w->cell_offsets.assign(w->C, 0);
    w->cell_segments.assign(w->S, 0);
    w->ncount.assign(w->S, 0);
    w->nlist.assign(w->slots, 0);
    w->refpos = w->pos;
    w->corr.assign(6 * (size_t)w->S, 0.f);
    w->seg_max.assign(w->S, 0.f);
    w->curv.assign(w->V, 0.f);
    w->icorr.assign(3 * (size_t)w->V, 0.f);
    w->vstep.assign(w->V, 0.f);
    w->wall.assign(3 * (size_t)w->V, 0.f);
    *out = w;
    return TGH_OK;
}

int orc_world_run_batch(void *h, const tgh_config *c, uint32_t iterations, tgh_batch_status *st) {
    if (!h || !c || !st) return fail(TGH_ERR_INVALID, "null argument");
    return run_batch(*(World *)h, c, iterations, st);
}

int orc_world_download_positions(void *h, float *out, size_t n) {
    World &w = *(World *)h;
    if (n < w.pos.size()) return fail(TGH_ERR_SIZE, "position buffer too small");
    std::memcpy(out, w.pos.data(), w.pos.size() * 4);
    return TGH_OK;
}

int orc_world_upload_positions(void *h, const float *in, size_t n) {
    World &w = *(World *)h;
    if (n != w.pos.size()) return fail(TGH_ERR_SIZE, "position count mismatch");
    std::memcpy(w.pos.data(), in, n * 4);
    w.nstate[0] = 1;
    return TGH_OK;
}

int orc_world_info(void *h, tgh_world_info *info) {
    World &w = *(World *)h;
    std::memset(info, 0, sizeof(*info));
    info->fiber_count = w.F;
    info->segment_count = w.S;
    info->vertex_count = w.V;
    info->cells[0] = w.cx;
    info->cells[1] = w.cy;
    info->cells[2] = w.cz;
    info->cell_count = w.C;
    info->cell_size = w.cell_size;
    info->neighbor_skin = w.skin;
    info->neighbor_block = w.block;
    info->neighbor_list_rebuilds = w.nstate[1];
    return TGH_OK;
}

// Rebuilds the cell list and neighbour lists from the current positions now
// (the body of rebuild_neighbor_lists_if_requested with the flag forced).
int orc_world_rebuild_now(void *h) {
    World &w = *(World *)h;
    w.nstate[0] = 1;
    rebuild_if_requested(w);
    return TGH_OK;
}

int orc_world_cell_list(void *h, uint32_t *counts, uint32_t *offsets, uint32_t *segments) {
    World &w = *(World *)h;
    std::memcpy(counts, w.cell_counts.data(), w.C * 4);
    std::memcpy(offsets, w.cell_offsets.data(), w.C * 4);
    std::memcpy(segments, w.cell_segments.data(), (size_t)w.S * 4);
    return TGH_OK;
}

int orc_world_neighbor_lists(void *h, uint32_t *counts, uint32_t *lists) {
    World &w = *(World *)h;
    std::memcpy(counts, w.ncount.data(), (size_t)w.S * 4);
    std::memcpy(lists, w.nlist.data(), std::min(w.nlist.size(), (size_t)w.block * w.S) * 4);
    return TGH_OK;
}

// Intermediate state for kernel-by-kernel comparisons: corrections[6S],
// segment max penetration[S], curvature ratio[V].
int orc_world_contact_state(void *h, float *corrections, float *segment_max, float *curvature) {
    World &w = *(World *)h;
    if (corrections) std::memcpy(corrections, w.corr.data(), w.corr.size() * 4);
    if (segment_max) std::memcpy(segment_max, w.seg_max.data(), w.seg_max.size() * 4);
    if (curvature) std::memcpy(curvature, w.curv.data(), w.curv.size() * 4);
    return TGH_OK;
}

// Curvature-cleanup counters since the last call: calls, projections
// applied, fiber sweeps whose forward pass projected, most projections of
// one fiber in one call.
void orc_gs_counters(uint64_t *out) {
    out[0] = g_gs_calls;
    out[1] = g_gs_projections;
    out[2] = g_gs_fiber_sweeps;
    out[3] = g_gs_max_fiber_projections;
    g_gs_calls = g_gs_projections = g_gs_fiber_sweeps = g_gs_max_fiber_projections = 0;
}

// Per fiber, at its last curvature cleanup: projections applied and
// triplets visited (fibers never cleaned up read 0).
void orc_world_gs_fibers(void *h, uint32_t *projections, uint32_t *visits) {
    const World &w = *(const World *)h;
    for (uint32_t f = 0; f < w.F; ++f) {
        projections[f] = f < w.gs_projections.size() ? w.gs_projections[f] : 0u;
        visits[f] = f < w.gs_visits.size() ? w.gs_visits[f] : 0u;
    }
}

void orc_world_activate(void *h, uint32_t maximum_step, uint64_t *segments, uint64_t *vertices) {
    World &w = *(World *)h;
    activate(w, maximum_step);
    uint64_t s = 0, v = 0;
    for (uint32_t f = 0; f < w.F; ++f)
        if (w.fiber_active[f]) {
            s += w.fiber_span[2 * f + 1] - 1;
            v += w.fiber_span[2 * f + 1];
        }
    *segments = s;
    *vertices = v;
}

void orc_world_set_layer_targets(void *h, uint32_t axis, const float *targets, size_t count, uint32_t first,
                                 uint32_t last, float stiffness, float max_translation) {
    World &w = *(World *)h;
    w.layer_targets.assign(targets, targets + count);
    w.layer_axis = axis;
    w.first_layer = first;
    w.last_layer = last;
    w.layer_stiffness = stiffness;
    w.layer_max = max_translation;
    apply_layer_targets(w);  // the first correction applies at once (world.rs:1766-1795)
}

void orc_world_clear_layer_targets(void *h) { ((World *)h)->layer_targets.clear(); }

void orc_world_set_vertex_targets(void *h, uint32_t axis, const uint32_t *vertices, size_t count, float displacement,
                                  float stiffness, float max_translation) {
    World &w = *(World *)h;  // initialize_vertex_displacement_targets (kernels.rs:582-603)
    w.vt_index.assign(vertices, vertices + count);
    w.vt_target.resize(count);
    w.vt_command.resize(count);
    for (size_t i = 0; i < count; ++i) {
        float c = w.pos[3 * (size_t)vertices[i] + axis];
        w.vt_target[i] = c + displacement;
        w.vt_command[i] = c;
    }
    w.vt_axis = axis;
    w.vt_stiffness = stiffness;
    w.vt_max = max_translation;
}

void orc_world_clear_vertex_targets(void *h) { ((World *)h)->vt_index.clear(); }

void orc_world_formation_target_error(void *h, float *layer_error, float *vertex_error, uint32_t *has_layer,
                                      uint32_t *has_vertex) {
    World &w = *(World *)h;  // measure_layer_target_error, measure_vertex_target_error
    float le = 0.f, ve = 0.f;
    if (!w.layer_targets.empty())
        for (uint32_t f = 0; f < w.F; ++f)
            if (layer_held(w, f))
                le = std::max(le, std::fabs(w.layer_targets[w.fiber_layer[f]] - fiber_center(w, f, w.layer_axis)));
    for (size_t i = 0; i < w.vt_index.size(); ++i)
        if (vtx_on(w, w.vt_index[i]))
            ve = std::max(ve, std::fabs(w.vt_target[i] - w.pos[3 * (size_t)w.vt_index[i] + w.vt_axis]));
    *layer_error = le;
    *vertex_error = ve;
    *has_layer = w.layer_targets.empty() ? 0u : 1u;
    *has_vertex = w.vt_index.empty() ? 0u : 1u;
}

int orc_world_compact_cell(void *h, const float *lower, const float *upper, uint32_t kinematics) {
    return compact_cell(*(World *)h, lower, upper, kinematics);
}

void orc_world_cell_bounds(void *h, float *lower, float *upper) {
    World &w = *(World *)h;
    for (int k = 0; k < 3; ++k) {
        lower[k] = w.lower[k];
        upper[k] = w.upper[k];
    }
}

void orc_world_compaction_metrics(void *h, float correction_fraction, float contact_stiffness,
                                  float stretch_stiffness, float bending_stiffness, float *out) {
    compaction_metrics(*(World *)h, correction_fraction, contact_stiffness, stretch_stiffness, bending_stiffness,
                       out);
}

// out: 6 words per candidate (first, second as u32 bits; s, t, gap, angle)
int orc_world_capture_contacts(void *h, float max_gap, size_t capacity, float *out, size_t *count,
                               uint32_t *overflow) {
    std::vector<Candidate> found;
    bool over = false;
    int rc = capture_contacts(*(World *)h, max_gap, capacity, found, over);
    if (rc) return rc;
    for (size_t i = 0; i < found.size(); ++i) {
        std::memcpy(&out[6 * i], &found[i].first, 4);
        std::memcpy(&out[6 * i + 1], &found[i].second, 4);
        out[6 * i + 2] = found[i].s;
        out[6 * i + 3] = found[i].t;
        out[6 * i + 4] = found[i].gap;
        out[6 * i + 5] = found[i].angle;
    }
    *count = found.size();
    *overflow = over ? 1u : 0u;
    return TGH_OK;
}

// set_fiber_maximum_curvatures (world.rs:2182-2191): the limit of every
// vertex of the fiber (only internal vertices are ever measured or projected)
void orc_world_set_fiber_maximum_curvatures(void *h, const float *values) {
    World &w = *(World *)h;
    for (uint32_t f = 0; f < w.F; ++f)
        for (uint32_t i = 0; i < w.fiber_span[2 * f + 1]; ++i) w.vtx_kmax[w.fiber_span[2 * f] + i] = values[f];
}

void orc_world_destroy(void *h) { delete (World *)h; }

}  // extern "C"
