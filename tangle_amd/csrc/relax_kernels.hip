// Relaxation kernels for CDNA4 (gfx950): cell-list binning, Verlet neighbour
// lists, capsule contact projection, stretch/bend Jacobi passes, per-fiber
// Gauss–Seidel curvature cleanup and the wave64 convergence reductions.
//
// These kernels implement the loop of DeviceFiberWorld::run_batch
// (crates/tangle_relax/src/device/world.rs:1020-1505) with the arithmetic of
// the reference kernels, restructured for the MI355X:
//
//   * Vertices live in float4 records (x, y, z, rest arc length): a segment's
//     two endpoints are one 32-byte contiguous load, and the along-fiber gap
//     test (crates/tangle_contact/src/device/cell_list.rs:99-110) reads its
//     arc lengths from the same load.
//   * Contact and list-build work is laid out in cell-list slot order, so the
//     lanes of a wave hold spatially adjacent segments and share their
//     candidate gathers in L2. The lists of each 8 consecutive slots are
//     packed one after another in those slots' 8 x block entries
//     (nbr[nbr_off[slot] + k]), so that short lists share lines; a group of
//     lanes reads a list, one candidate each.
//   * Each segment's list is sorted as it is stored, from LDS, which
//     replaces the separate sort pass (neighbor.rs:366-408).
//   * The Jacobi stretch/bend passes and the contact application read one
//     position buffer and write the other (kernels.rs:1027-1163, 1417-1537):
//     a find/apply pair becomes one launch with identical per-vertex
//     arithmetic, and the last pass also applies the contact corrections.
//   * The curvature Gauss–Seidel stages each fiber in LDS and, at 8 lanes
//     per fiber, walks every triplet in order with the triplet carried in
//     registers, one lane per axis; the wider groups screen the next
//     triplets in parallel with a conservative bound and project the first
//     candidate exactly (kernels.rs:1165-1415). Each stops after a sweep
//     direction that projected nothing, which leaves the result unchanged.
//   * Max penetration / curvature ratio / vertex step reduce within each
//     workgroup and then with one u32 atomicMax per workgroup into one of
//     the specimen's slots (order-free, so exact); one small workgroup then
//     folds the slots and assesses every specimen's convergence.
//   * Several independent specimens can share one world: every element
//     reads its specimen's grid and state, so one launch serves them all.
//
// All float arithmetic must be compiled without contraction
// (-ffp-contract=off) and with IEEE division and square root, as the Rust
// reference; cell lists and neighbour pair sets are then bit-exact. The
// curvature sweep's arc cosine is the host libm's algorithm (acosf_f32.h),
// so its projections reproduce a host evaluation bit for bit as well.

#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>
#include <stdint.h>

#include "acosf_f32.h"
#include "periodic_image.h"
#include "relax_device.h"

namespace tgh {

__device__ __forceinline__ float rclamp(float x, float lo, float hi) {  // f32::clamp
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}

__device__ __forceinline__ int32_t sat_i32(float x) {  // Rust `as i32`
    if (x != x) return 0;
    if (x <= -2147483648.0f) return INT32_MIN;
    if (x >= 2147483648.0f) return INT32_MAX;
    return (int32_t)x;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int offset = 32; offset > 0; offset >>= 1) {
        uint32_t o = (uint32_t)__shfl_xor((int)v, offset, 64);
        v = v > o ? v : o;
    }
    return v;
}

// The minimum-image data of one specimen, held in registers for the
// per-candidate loops.
struct Image {
    float length[3], inv_length[3];
    uint32_t periodic[3];
};
__device__ __forceinline__ Image image_of(const GridParams &g) {
    Image m;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        m.length[k] = g.length[k];
        m.inv_length[k] = 1.0f / g.length[k];
        m.periodic[k] = g.periodic[k];
    }
    return m;
}

// Cell of a segment midpoint (proxy_cell, cell_list.rs:32-79, one proxy).
__device__ __forceinline__ void cell_coords(const GridParams &g, float px, float py, float pz, float qx,
                                            float qy, float qz, uint32_t &x, uint32_t &y, uint32_t &z) {
    float mx = 0.5f * (px + qx), my = 0.5f * (py + qy), mz = 0.5f * (pz + qz);
    int32_t rx = sat_i32(floorf((mx - g.lower[0]) / g.width[0]));
    int32_t ry = sat_i32(floorf((my - g.lower[1]) / g.width[1]));
    int32_t rz = sat_i32(floorf((mz - g.lower[2]) / g.width[2]));
    int32_t nx = (int32_t)g.cells[0], ny = (int32_t)g.cells[1], nz = (int32_t)g.cells[2];
    x = (uint32_t)min(max(rx, 0), nx - 1);
    y = (uint32_t)min(max(ry, 0), ny - 1);
    z = (uint32_t)min(max(rz, 0), nz - 1);
    if (g.periodic[0]) x = (uint32_t)(((rx % nx) + nx) % nx);
    if (g.periodic[1]) y = (uint32_t)(((ry % ny) + ny) % ny);
    if (g.periodic[2]) z = (uint32_t)(((rz % nz) + nz) % nz);
}

__device__ __forceinline__ uint32_t cell_index(const GridParams &g, float4 p, float4 q) {
    uint32_t x, y, z;
    cell_coords(g, p.x, p.y, p.z, q.x, q.y, q.z, x, y, z);
    return (z * g.cells[1] + y) * g.cells[0] + x;
}

// Cell `n` (0..26, x fastest) of the stencil around `home`, with periodic
// de-duplication for one- and two-cell axes (neighbor.rs:242-276).
__device__ __forceinline__ bool stencil_cell(const GridParams &g, uint32_t home, uint32_t n, uint32_t &cell) {
    const uint32_t cx = g.cells[0], cy = g.cells[1], cz = g.cells[2];
    uint32_t hx = home % cx, hy = (home / cx) % cy, hz = home / (cx * cy);
    int32_t ox = (int32_t)(n % 3) - 1, oy = (int32_t)((n / 3) % 3) - 1, oz = (int32_t)(n / 9) - 1;
    int32_t rx = (int32_t)hx + ox, ry = (int32_t)hy + oy, rz = (int32_t)hz + oz;
    int32_t nx = (int32_t)cx, ny = (int32_t)cy, nz = (int32_t)cz;
    bool ok = (g.periodic[0] || (rx >= 0 && rx < nx)) && (g.periodic[1] || (ry >= 0 && ry < ny)) &&
              (g.periodic[2] || (rz >= 0 && rz < nz));
    ok = ok && (!g.periodic[0] || !((cx == 1 && ox != 0) || (cx == 2 && ((hx == 0 && ox > 0) || (hx == 1 && ox < 0)))));
    ok = ok && (!g.periodic[1] || !((cy == 1 && oy != 0) || (cy == 2 && ((hy == 0 && oy > 0) || (hy == 1 && oy < 0)))));
    ok = ok && (!g.periodic[2] || !((cz == 1 && oz != 0) || (cz == 2 && ((hz == 0 && oz > 0) || (hz == 1 && oz < 0)))));
    if (!ok) return false;
    uint32_t x = (uint32_t)(((rx % nx) + nx) % nx);
    uint32_t y = (uint32_t)(((ry % ny) + ny) % ny);
    uint32_t z = (uint32_t)(((rz % nz) + nz) % nz);
    cell = (z * cy + y) * cx + x;
    return true;
}

// stencil_cell from the home cell's coordinates (hx, hy, hz), without
// divisions: the same cells, for the same n.
__device__ __forceinline__ bool stencil_cell_at(const GridParams &g, uint32_t hx, uint32_t hy, uint32_t hz,
                                                uint32_t n, uint32_t &cell) {
    const uint32_t cx = g.cells[0], cy = g.cells[1], cz = g.cells[2];
    const int32_t ox = (int32_t)(n % 3u) - 1, oy = (int32_t)((n / 3u) % 3u) - 1, oz = (int32_t)(n / 9u) - 1;
    const int32_t rx = (int32_t)hx + ox, ry = (int32_t)hy + oy, rz = (int32_t)hz + oz;
    const int32_t nx = (int32_t)cx, ny = (int32_t)cy, nz = (int32_t)cz;
    bool ok = (g.periodic[0] || (rx >= 0 && rx < nx)) && (g.periodic[1] || (ry >= 0 && ry < ny)) &&
              (g.periodic[2] || (rz >= 0 && rz < nz));
    ok = ok && (!g.periodic[0] || !((cx == 1 && ox != 0) || (cx == 2 && ((hx == 0 && ox > 0) || (hx == 1 && ox < 0)))));
    ok = ok && (!g.periodic[1] || !((cy == 1 && oy != 0) || (cy == 2 && ((hy == 0 && oy > 0) || (hy == 1 && oy < 0)))));
    ok = ok && (!g.periodic[2] || !((cz == 1 && oz != 0) || (cz == 2 && ((hz == 0 && oz > 0) || (hz == 1 && oz < 0)))));
    // |offset| <= 1, so one wrap suffices
    const uint32_t x = (uint32_t)(rx < 0 ? rx + nx : (rx >= nx ? rx - nx : rx));
    const uint32_t y = (uint32_t)(ry < 0 ? ry + ny : (ry >= ny ? ry - ny : ry));
    const uint32_t z = (uint32_t)(rz < 0 ? rz + nz : (rz >= nz ? rz - nz : rz));
    cell = (z * cy + y) * cx + x;
    return ok;
}

// Distance between two segment axes (proxy_pair_axis_distance,
// neighbor.rs:25-86, one proxy per segment so the canonical test holds).
__device__ __forceinline__ float axis_distance(float p1x, float p1y, float p1z, float d1x, float d1y, float d1z,
                                               float p2x, float p2y, float p2z, float d2x, float d2y, float d2z) {
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
            if (fabsf(den) > eps_par * a * e) s = rclamp((b * f - c * e) / den, 0.f, 1.f);
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
    return sqrtf(dx * dx + dy * dy + dz * dz);
}

// ---------------------------------------------------------------------------
// batch begin (kernels.rs:3-12)

__global__ void k_begin_batch(SpecState *ss, uint32_t nspec) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nspec) return;
    ss[s].control[2] = 0;
    if (ss[s].control[3] == 0) ss[s].control[0] = 1;
}

__device__ __forceinline__ bool running(const SpecState *ss, uint32_t spec) { return ss[spec].control[0] != 0u; }

// ---------------------------------------------------------------------------
// neighbour-list rebuild (world/neighbor_list.rs:19-62). Every kernel returns
// at once unless some specimen's rebuild is pending, and then works only on
// the cells, segments, slots and vertices of the specimens whose own request
// flag is set: each specimen rebuilds exactly when it would alone.

__device__ __forceinline__ bool rebuilding(const SpecState *ss, uint32_t spec) { return ss[spec].rebuild != 0u; }

// The rebuild kernels run a capped grid over chunks of 256 elements (the
// grid-stride form), so that a pass without a pending rebuild costs one short
// launch each; a chunk none of whose specimens rebuilds is skipped whole.
constexpr uint32_t kRebuildBlocks = 2048;

// Whether any of specimens lo..hi (those of a chunk's first and last
// element) is rebuilding.
__device__ __forceinline__ bool any_rebuilding(const SpecState *ss, uint32_t lo, uint32_t hi) {
    for (uint32_t s = lo; s <= hi; ++s)
        if (ss[s].rebuild != 0u) return true;
    return false;
}

__device__ __forceinline__ uint32_t vertex_spec(const float4 *vB, uint32_t v) {
    return __float_as_uint(vB[v].z) >> kSpecShift;
}

__global__ __launch_bounds__(256) void k_rebuild_clear(const DevState *st, const SpecState *ss,
                                                       const uint32_t *cell_spec, uint32_t *cell_counts,
                                                       uint32_t *cell_cursor, uint32_t C, const float4 *pos,
                                                       float4 *refpos, const float4 *vB, uint32_t V,
                                                       uint32_t *cell_reach) {
    if (st->nstate[0] == 0) return;
    for (uint32_t first = blockIdx.x * 256u; first < C; first += gridDim.x * 256u) {
        if (!any_rebuilding(ss, cell_spec[first], cell_spec[min(first + 255u, C - 1u)])) continue;
        const uint32_t i = first + threadIdx.x;
        if (i < C && rebuilding(ss, cell_spec[i])) {
            cell_counts[i] = 0;
            cell_cursor[i] = 0;
            cell_reach[i] = 0;
        }
    }
    // The list snapshot (neighbor.rs:410-422) is taken from the positions the
    // lists are built from; nothing moves them before the build finishes.
    for (uint32_t first = blockIdx.x * 256u; first < V; first += gridDim.x * 256u) {
        if (!any_rebuilding(ss, vertex_spec(vB, first), vertex_spec(vB, min(first + 255u, V - 1u)))) continue;
        const uint32_t i = first + threadIdx.x;
        if (i < V && rebuilding(ss, vertex_spec(vB, i))) refpos[i] = pos[i];
    }
}

__global__ __launch_bounds__(256) void k_bin_segments(const DevState *st, const SpecState *ss,
                                                      const GridParams *grids, const float4 *pos, const uint4 *sstat,
                                                      const uint32_t *factive, uint32_t *seg_cell,
                                                      uint32_t *cell_counts, uint32_t *cell_reach, uint32_t S) {
    if (st->nstate[0] == 0) return;
    for (uint32_t first = blockIdx.x * 256u; first < S; first += gridDim.x * 256u) {
        if (!any_rebuilding(ss, sstat[first].w, sstat[min(first + 255u, S - 1u)].w)) continue;
        const uint32_t s = first + threadIdx.x;
        if (s >= S) continue;
        const uint4 me = sstat[s];
        if (!rebuilding(ss, me.w)) continue;
        if (factive[me.y] == 0u) {  // inactive fibers are not in the cell list
            seg_cell[s] = 0xffffffffu;
            continue;
        }
        const GridParams &g = grids[me.w];
        const uint32_t v0 = me.x;
        const float4 p = pos[v0], q = pos[v0 + 1];
        const uint32_t cell = g.cell_base + cell_index(g, p, q);
        seg_cell[s] = cell;
        atomicAdd(&cell_counts[cell], 1u);
        // half length plus radius, for the list build's bound (a max of
        // nonnegative floats is a max of their bits)
        const float dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
        const float ext = 0.5f * sqrtf(dx * dx + dy * dy + dz * dz) + __uint_as_float(me.z);
        // a NaN bounds nothing: its cell is never left out
        atomicMax(&cell_reach[cell], ext == ext ? __float_as_uint(ext) : 0x7f800000u);
    }
}

// Exclusive scan of one specimen's cell counts per 1024-thread workgroup
// (one workgroup per specimen): each thread sums a contiguous chunk, the 1024
// chunk sums are scanned in LDS, then each thread writes its chunk's offsets,
// which start at the specimen's first segment. The slots past its active
// segments are marked empty.
__global__ __launch_bounds__(1024) void k_scan_cells(const DevState *st, const SpecState *ss,
                                                     const GridParams *grids, const uint32_t *counts,
                                                     uint32_t *offsets, uint32_t *cell_segs, uint32_t *scattered,
                                                     uint4 *slotC) {
    if (st->nstate[0] == 0) return;
    const uint32_t sp = blockIdx.x;
    if (!rebuilding(ss, sp)) return;
    const GridParams &g = grids[sp];
    const uint32_t C = g.cells[0] * g.cells[1] * g.cells[2];
    counts += g.cell_base;
    offsets += g.cell_base;
    __shared__ uint32_t sums[1024];
    uint32_t t = threadIdx.x;
    uint32_t chunk = (C + 1023) / 1024;
    uint32_t begin = min(t * chunk, C);
    uint32_t end = min(begin + chunk, C);
    uint32_t local = 0;
    for (uint32_t i = begin; i < end; ++i) local += counts[i];
    sums[t] = local;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {
        uint32_t add = t >= d ? sums[t - d] : 0u;
        __syncthreads();
        sums[t] += add;
        __syncthreads();
    }
    uint32_t run = g.seg_base + sums[t] - local;
    for (uint32_t i = begin; i < end; ++i) {
        offsets[i] = run;
        run += counts[i];
    }
    const uint32_t filled = sums[1023];  // the specimen's active segments
    for (uint32_t i = g.seg_base + filled + t; i < g.seg_base + g.seg_count; i += 1024) {
        cell_segs[i] = kNoSegment;
        scattered[i] = kNoSegment;
        slotC[i] = make_uint4(kNoSegment, kNoSegment, 0u, sp);
    }
}

__global__ __launch_bounds__(256) void k_scatter_segments(const DevState *st, const SpecState *ss,
                                                          const uint4 *sstat, const uint32_t *cell_spec,
                                                          const uint32_t *seg_cell, const uint32_t *cell_offsets,
                                                          uint32_t *cell_cursor, uint32_t *scattered, uint32_t S) {
    if (st->nstate[0] == 0) return;
    for (uint32_t first = blockIdx.x * 256u; first < S; first += gridDim.x * 256u) {
        if (!any_rebuilding(ss, sstat[first].w, sstat[min(first + 255u, S - 1u)].w)) continue;
        const uint32_t s = first + threadIdx.x;
        if (s >= S) continue;
        const uint32_t cell = seg_cell[s];
        if (cell == 0xffffffffu) continue;  // inactive
        if (!rebuilding(ss, cell_spec[cell])) continue;
        const uint32_t slot = cell_offsets[cell] + atomicAdd(&cell_cursor[cell], 1u);
        scattered[slot] = s;
    }
}

// One thread per cell-list slot: the segment scattered into the slot is
// ranked against the other entries of its cell, which puts every cell in
// ascending segment order (the order rank_cell_segments produces,
// cell_list.rs:344-378), and the slot records the list build and the contact
// pass read are written (gather_cell_slot_geometry, neighbor.rs:99-153): the
// two vertex records, (segment, fiber, radius, specimen) and the compact
// records. Slot i belongs to the specimen of segment i (a specimen's slots
// are its segment range).
__global__ __launch_bounds__(256) void k_rank_slots(const DevState *st, const SpecState *ss, const uint4 *sstat,
                                                    const uint32_t *seg_cell, const uint32_t *cell_counts,
                                                    const uint32_t *cell_offsets, const uint32_t *scattered,
                                                    uint32_t *cell_segs, const float4 *pos, float4 *slotA,
                                                    float4 *slotB, uint4 *slotC, float4 *cP,
                                                    float4 *cQ, uint32_t S) {
    if (st->nstate[0] == 0) return;
    for (uint32_t first = blockIdx.x * 256u; first < S; first += gridDim.x * 256u) {
        if (!any_rebuilding(ss, sstat[first].w, sstat[min(first + 255u, S - 1u)].w)) continue;
        const uint32_t i = first + threadIdx.x;
        if (i >= S || !rebuilding(ss, sstat[i].w)) continue;
        const uint32_t mine = scattered[i];
        if (mine == kNoSegment) continue;  // past the specimen's active segments
        const uint32_t cell = seg_cell[mine];
        const uint32_t n = cell_counts[cell], off = cell_offsets[cell];
        uint32_t rank = 0;
        for (uint32_t k = 0; k < n; ++k) rank += scattered[off + k] < mine ? 1u : 0u;
        const uint32_t slot = off + rank;
        cell_segs[slot] = mine;
        const uint4 rec = sstat[mine];
        const float4 p = pos[rec.x], q = pos[rec.x + 1];
        slotA[slot] = p;
        slotB[slot] = q;
        slotC[slot] = make_uint4(mine, rec.y, rec.z, rec.w);
        cP[slot] = make_float4(p.x, p.y, p.z, __uint_as_float(rec.z));
        cQ[slot] = make_float4(q.x, q.y, q.z, __uint_as_float(mine));
    }
}

// The contact pass's compact slot records from the current vertex records
// (the first vertex of segment s of fiber f is s + f: fibers are
// contiguous), every iteration before the contact pass, so that a candidate
// is one 32-byte gather of records stored in cell order.
__global__ void k_refresh_slots(const float4 *pos, const uint4 *slotC, float4 *cP, float4 *cQ, uint32_t S) {
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= S) return;
    const uint4 sc = slotC[slot];
    if (sc.x == kNoSegment) return;
    const uint32_t v = sc.x + sc.y;
    const float4 p = pos[v], q = pos[v + 1];
    cP[slot] = make_float4(p.x, p.y, p.z, __uint_as_float(sc.z));
    cQ[slot] = make_float4(q.x, q.y, q.z, __uint_as_float(sc.x));
}

// Orders LDS and global accesses of the lanes of one wave: what a lane wrote
// before is seen by every lane of the wave after.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The list build: kBuildLanes lanes per cell-list slot, kBuildGroups slots
// per wave, one wave per workgroup.
constexpr uint32_t kBuildLanes = 8;
constexpr uint32_t kBuildGroups = 64 / kBuildLanes;
constexpr uint32_t kBuildKeys = 48;      // list entries per slot kept in LDS for the sort
constexpr uint32_t kBuildQueue = 128;    // candidates past the reach test awaiting the axis test
constexpr uint32_t kBuildBlocks = 16384; // grid cap of the one-wave build workgroups

// One slot's data the axis test reads from LDS.
struct BuildSlot {
    float p[3], q[3];
    float radius, skin;
    float length[3], inv_length[3];
    uint32_t periodic;
};

// The reference's list test of one candidate (neighbor.rs:155-364) is
// evaluated in two parts: the nearest-image shift and the reach bound on
// the segment midpoints, then the axis distance. Both parts compute their
// quantities with the same operations as the single test.
__device__ __forceinline__ void list_image(const float *length, const float *inv_length, uint32_t periodic,
                                           float p1x, float p1y, float p1z, float q1x, float q1y, float q1z,
                                           float &p2x, float &p2y, float &p2z, float &q2x, float &q2y, float &q2z) {
    if (periodic & 1u) {
        float sh = image_round(0.5f * (p2x + q2x - p1x - q1x), length[0], inv_length[0]) * length[0];
        p2x -= sh;
        q2x -= sh;
    }
    if (periodic & 2u) {
        float sh = image_round(0.5f * (p2y + q2y - p1y - q1y), length[1], inv_length[1]) * length[1];
        p2y -= sh;
        q2y -= sh;
    }
    if (periodic & 4u) {
        float sh = image_round(0.5f * (p2z + q2z - p1z - q1z), length[2], inv_length[2]) * length[2];
        p2z -= sh;
        q2z -= sh;
    }
}

// A slot's view of the stencil boxes, for the list build's cell bound: per
// axis the lower edge of the box of offset 0 (the cell holding the slot's
// midpoint s, unwrapped on a periodic axis, as the binning computes it
// before wrapping), the cell width and a rounding margin (the cell's binning
// and the reach test's midpoint difference each round the coordinates). On
// a periodic axis every candidate's nearest image lies in the box next to
// s's cell when the axis has five cells or more (any other image is at least
// 3 cells away); with fewer the axis bounds nothing. On a walled axis the
// first and last cells also hold the midpoints beyond the walls.
struct StencilBoxes {
    float s[3], lo0[3], w[3], margin[3];
    int32_t c0[3], cells[3];
    bool use[3], walled[3];
};
__device__ __forceinline__ StencilBoxes stencil_boxes(const GridParams &g, uint32_t hx, uint32_t hy, uint32_t hz,
                                                      float sx, float sy, float sz) {
    StencilBoxes b;
    const uint32_t h[3] = {hx, hy, hz};
    const float sv[3] = {sx, sy, sz};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        b.s[k] = sv[k];
        b.w[k] = g.width[k];
        b.cells[k] = (int32_t)g.cells[k];
        b.walled[k] = !g.periodic[k];
        b.use[k] = !g.periodic[k] || g.cells[k] >= 5u;
        const float c = g.periodic[k] ? floorf((sv[k] - g.lower[k]) / g.width[k]) : (float)h[k];
        b.c0[k] = (int32_t)h[k];
        b.lo0[k] = g.lower[k] + c * g.width[k];
        b.margin[k] = 1.0e-3f * g.width[k] + 1.0e-5f * (fabsf(g.lower[k]) + fabsf(g.upper[k]) + fabsf(sv[k]));
    }
    return b;
}
// Squared distance from the slot's midpoint to the box of stencil cell n,
// less the margins.
__device__ __forceinline__ float stencil_gap2(const StencilBoxes &b, uint32_t n) {
    const int32_t o[3] = {(int32_t)(n % 3u) - 1, (int32_t)((n / 3u) % 3u) - 1, (int32_t)(n / 9u) - 1};
    const float inf = __builtin_inff();
    float d2 = 0.f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (!b.use[k]) continue;
        float lo = b.lo0[k] + (float)o[k] * b.w[k], hi = lo + b.w[k];
        if (b.walled[k]) {
            const int32_t c = b.c0[k] + o[k];
            if (c <= 0) lo = -inf;
            if (c >= b.cells[k] - 1) hi = inf;
        }
        const float d = fmaxf(fmaxf(lo - b.s[k], b.s[k] - hi) - b.margin[k], 0.f);
        d2 += d * d;
    }
    return d2;
}

// The list build (build_segment_neighbor_lists, neighbor.rs:155-364), one
// wave per kBuildGroups consecutive cell-list slots, kBuildLanes lanes each.
// A slot's candidates are its 27 stencil cells' segments, cell by cell in
// stencil order and in slot (ascending segment) order within a cell: exactly
// the order in which the reference's overflow scan visits them
// (contact.rs:348-428). Each round every group tests its next kBuildLanes
// candidates (adjacent compact records) against the reach bound; the few
// that pass are queued in order in LDS and given the axis-distance test 64 at
// a time, so that the costly part keeps every lane busy. The accepted ones
// (the pairs within the skin, the only ones that can touch before the next
// rebuild) are appended to each slot's row of `ovf` (`ovf_cap` slots) in
// sequence order. A listed segment's row (at most `block` entries) is then
// written to `nbr` as slots in ascending segment order, which replaces the
// separate sort (neighbor.rs:366-408) and keeps the reference's summation
// order; an overflowed segment keeps its `ovf` row.
template <bool PRUNE>
__global__ __launch_bounds__(64) void k_build_lists(const DevState *st, const SpecState *ss, const GridParams *grids,
                                                    const uint4 *sstat, const uint4 *fspan, const float4 *slotA,
                                                    const float4 *slotB, const uint4 *slotC, const float4 *cP,
                                                    const float4 *cQ, const uint32_t *seg_cell,
                                                    const uint32_t *cell_counts, const uint32_t *cell_offsets,
                                                    uint32_t *ncount, uint32_t *nbr, uint32_t *nbr_off, uint32_t *ovf,
                                                    const uint32_t *cell_reach, uint32_t block, uint32_t ovf_cap,
                                                    uint32_t S) {
    constexpr uint32_t G = kBuildLanes, W = kBuildGroups;
    constexpr uint32_t per_lane = (27 + G - 1) / G;  // stencil cells per lane
    __shared__ BuildSlot slots[W];
    __shared__ uint32_t cell_start[W][28];  // a slot's candidate start of each stencil cell
    __shared__ uint32_t cell_slot[W][27];   // first slot of each stencil cell
    __shared__ uint32_t counts[W];
    __shared__ uint32_t specs[W];  // the slots' specimens (a slot's is its segment's)
    __shared__ uint2 queue[kBuildQueue];    // (group, candidate slot)
    __shared__ uint32_t keys[W][kBuildKeys];
    __shared__ uint32_t listed[W][kBuildKeys];
    if (st->nstate[0] == 0) return;
    const uint32_t lane = threadIdx.x % G, group = threadIdx.x / G;
    const uint64_t below = (1ull << threadIdx.x) - 1ull;
    // Workgroups are dealt to the 8 XCDs in turn; each XCD walks its own
    // eighth of the chunks, so that neighbouring slots share an L2.
    const uint32_t chunks = (S + W - 1u) / W;
    uint32_t c0 = blockIdx.x, c1 = chunks, step = gridDim.x;
    if (gridDim.x >= 8u) {
        const uint32_t xcd = blockIdx.x % 8u;
        c0 = (uint32_t)((uint64_t)chunks * xcd / 8u) + blockIdx.x / 8u;
        c1 = (uint32_t)((uint64_t)chunks * (xcd + 1u) / 8u);
        step = (gridDim.x - xcd + 7u) / 8u;  // workgroups on this XCD
    }
    for (uint32_t chunk = c0; chunk < c1; chunk += step) {
        const uint32_t first = chunk * W;
        if (!any_rebuilding(ss, sstat[first].w, sstat[min(first + W - 1u, S - 1u)].w)) continue;
        const uint32_t slot = first + group;
        const uint4 M = slotC[slot < S ? slot : first];
        const bool live = slot < S && M.x != kNoSegment && rebuilding(ss, M.w);
        const uint32_t seg = M.x, owner = M.y;
        const GridParams &g = grids[live ? M.w : 0u];
        uint32_t periodic = 0;
        for (int k = 0; k < 3; ++k)
            if (g.periodic[k]) periodic |= 1u << k;
        float length[3] = {g.length[0], g.length[1], g.length[2]};
        float inv_length[3] = {1.0f / g.length[0], 1.0f / g.length[1], 1.0f / g.length[2]};
        const float skin = g.skin;
        uint32_t total = 0;
        float4 A = make_float4(0.f, 0.f, 0.f, 0.f), B = A;
        uint32_t first_seg = 0, nseg = 0;
        if (live) {
            A = slotA[slot];
            B = slotB[slot];
            const uint4 span = fspan[owner];
            first_seg = span.x - owner;  // segment s of fiber f starts at vertex s + f
            nseg = span.y - 1u;
            const uint32_t home = seg_cell[seg] - g.cell_base;
            const uint32_t hx = home % g.cells[0], hy = (home / g.cells[0]) % g.cells[1],
                           hz = home / (g.cells[0] * g.cells[1]);
            // A stencil cell none of whose candidates can pass the reach test
            // below is left out (its candidates would all be rejected, so the
            // accepted sequence is the same): the test's reach is at most
            // this segment's half length, radius and skin plus the largest
            // half length and radius of the cell's segments, widened by its
            // (1 + 1e-4) and more.
            float rself = 0.f;
            StencilBoxes boxes;
            if (PRUNE) {
                const float bdx = B.x - A.x, bdy = B.y - A.y, bdz = B.z - A.z;
                rself = 0.5f * sqrtf(bdx * bdx + bdy * bdy + bdz * bdz) + __uint_as_float(M.z) + skin;
                boxes = stencil_boxes(g, hx, hy, hz, 0.5f * (A.x + B.x), 0.5f * (A.y + B.y), 0.5f * (A.z + B.z));
            }
            // the lane's stencil cells, then their starts in the candidate sequence
            uint32_t local = 0;
            for (uint32_t j = 0; j < per_lane; ++j) {
                const uint32_t n = lane * per_lane + j;
                if (n >= 27) break;
                uint32_t cell, cn = 0, cs = 0;
                if (stencil_cell_at(g, hx, hy, hz, n, cell)) {
                    cell += g.cell_base;
                    cn = cell_counts[cell];
                    cs = cell_offsets[cell];
                    if (PRUNE) {
                        const float rmax = (rself + __uint_as_float(cell_reach[cell])) * 1.001f;
                        if (stencil_gap2(boxes, n) > rmax * rmax) cn = 0;
                    }
                }
                cell_start[group][n] = local;  // lane-relative for now
                cell_slot[group][n] = cs;
                local += cn;
            }
            uint32_t incl = local;
#pragma unroll
            for (uint32_t d = 1; d < G; d <<= 1) {
                const uint32_t up = (uint32_t)__shfl_up((int)incl, d, G);
                if (lane >= d) incl += up;
            }
            const uint32_t excl = incl - local;
            total = (uint32_t)__shfl((int)incl, G - 1, G);
            for (uint32_t j = 0; j < per_lane; ++j) {
                const uint32_t n = lane * per_lane + j;
                if (n >= 27) break;
                cell_start[group][n] += excl;
            }
            if (lane == 0) {
                cell_start[group][27] = total;
                BuildSlot me;
                me.p[0] = A.x, me.p[1] = A.y, me.p[2] = A.z;
                me.q[0] = B.x, me.q[1] = B.y, me.q[2] = B.z;
                me.radius = __uint_as_float(M.z);
                me.skin = skin;
                me.length[0] = length[0], me.length[1] = length[1], me.length[2] = length[2];
                me.inv_length[0] = inv_length[0], me.inv_length[1] = inv_length[1], me.inv_length[2] = inv_length[2];
                me.periodic = periodic;
                slots[group] = me;
            }
        }
        if (lane == 0) {
            counts[group] = 0u;
            specs[group] = slot < S ? sstat[slot].w : 0xffffffffu;
        }
        wave_sync();
        const float arc_p = A.w, arc_q = B.w;
        const float p1x = A.x, p1y = A.y, p1z = A.z, radius = __uint_as_float(M.z);
        const float q1x = B.x, q1y = B.y, q1z = B.z;
        const float d1x = q1x - p1x, d1y = q1y - p1y, d1z = q1z - p1z;
        const float half_length = 0.5f * sqrtf(d1x * d1x + d1y * d1y + d1z * d1z);
        uint32_t n = 0, base = 0;               // this lane's stencil cell; the group's next candidate
        uint32_t head = 0, tail = 0;            // queue read and write counts
        bool screening = __ballot(base < total) != 0ull;
        while (screening || head != tail) {
            if (screening && tail - head < 64u) {
                // every group tests its next G candidates against the reach bound
                const uint32_t t = base + lane;
                bool pass = false;
                uint32_t o = 0;
                if (t < total) {
                    while (cell_start[group][n + 1] <= t) ++n;
                    o = cell_slot[group][n] + (t - cell_start[group][n]);
                    const float4 P2 = cP[o], Q2 = cQ[o];
                    const uint32_t other = __float_as_uint(Q2.w);
                    const float other_radius = P2.w;
                    bool skip = other == seg;
                    if (!skip && other - first_seg < nseg) {  // same fiber: along-fiber neighbours are skipped
                        const float gap = fmaxf(slotA[o].w - arc_q, arc_p - slotB[o].w);
                        skip = gap < radius + other_radius;
                    }
                    if (!skip) {
                        float p2x = P2.x, p2y = P2.y, p2z = P2.z, q2x = Q2.x, q2y = Q2.y, q2z = Q2.z;
                        list_image(length, inv_length, periodic, p1x, p1y, p1z, q1x, q1y, q1z, p2x, p2y, p2z, q2x, q2y,
                                   q2z);
                        float d2x = q2x - p2x, d2y = q2y - p2y, d2z = q2z - p2z;
                        float inter = (radius + other_radius + skin) * (1.0f + 1.0e-4f);
                        float mdx = 0.5f * (p2x + q2x - p1x - q1x);
                        float mdy = 0.5f * (p2y + q2y - p1y - q1y);
                        float mdz = 0.5f * (p2z + q2z - p1z - q1z);
                        float reach = half_length + 0.5f * sqrtf(d2x * d2x + d2y * d2y + d2z * d2z) + inter;
                        pass = mdx * mdx + mdy * mdy + mdz * mdz <= reach * reach;
                    }
                }
                const uint64_t mask = __ballot(pass);
                if (pass) queue[(tail + (uint32_t)__popcll(mask & below)) % kBuildQueue] = make_uint2(group, o);
                tail += (uint32_t)__popcll(mask);
                base += G;
                screening = __ballot(base < total) != 0ull;
                wave_sync();
            }
            const uint32_t waiting = tail - head;
            if (waiting < 64u && screening) continue;
            // the axis test of up to 64 queued candidates, in order
            const uint32_t take = waiting < 64u ? waiting : 64u;
            bool accept = false;
            uint32_t k = W, o = 0, other = 0;
            if (threadIdx.x < take) {
                const uint2 e = queue[(head + threadIdx.x) % kBuildQueue];
                k = e.x;
                o = e.y;
                const BuildSlot &me = slots[k];
                const float4 P2 = cP[o], Q2 = cQ[o];
                other = __float_as_uint(Q2.w);
                const float ax = me.p[0], ay = me.p[1], az = me.p[2];
                const float bx = me.q[0], by = me.q[1], bz = me.q[2];
                float p2x = P2.x, p2y = P2.y, p2z = P2.z, q2x = Q2.x, q2y = Q2.y, q2z = Q2.z;
                list_image(me.length, me.inv_length, me.periodic, ax, ay, az, bx, by, bz, p2x, p2y, p2z, q2x, q2y, q2z);
                const float inter = (me.radius + P2.w + me.skin) * (1.0f + 1.0e-4f);
                accept = axis_distance(ax, ay, az, bx - ax, by - ay, bz - az, p2x, p2y, p2z, q2x - p2x, q2y - p2y,
                                       q2z - p2z) <= inter;
            }
            head += take;
            // an accepted entry's place: its slot's count plus the accepted
            // entries of the same slot before it in the queue
#pragma unroll
            for (uint32_t w = 0; w < W; ++w) {
                const uint64_t mine = __ballot(accept && k == w);
                if (accept && k == w) {
                    const uint32_t at = counts[w] + (uint32_t)__popcll(mine & below);
                    if (at < ovf_cap) ovf[(size_t)(first + w) * ovf_cap + at] = o;
                    if (at < kBuildKeys) {
                        keys[w][at] = other;
                        listed[w][at] = o;
                    }
                }
                wave_sync();
                if (threadIdx.x == 0) counts[w] += (uint32_t)__popcll(mine);
            }
            wave_sync();
        }
        // counts, then the listed rows in ascending segment order. The 8
        // slots' rows are packed in their 8 x block entries: a slot's row
        // follows the listed rows of the slots before it of the same
        // specimen, from the entries of the first of them (a specimen's
        // slots rebuild together, so its part of the entries is rewritten
        // whole).
        const uint32_t c = counts[group];
        uint32_t part = group, at = 0;
        while (part > 0 && specs[part - 1] == specs[group]) --part;
        for (uint32_t k = part; k < group; ++k) at += counts[k] <= block ? counts[k] : 0u;
        const uint32_t off = (first + part) * block + at;
        if (live && lane == 0) {
            ncount[slot] = c;
            nbr_off[slot] = off;
        }
        uint32_t *out = nbr + off;
        if (live && c <= block && c <= kBuildKeys) {
            // rank sort: segments are distinct, so every entry has its own rank
            for (uint32_t i = lane; i < c; i += G) {
                const uint32_t key = keys[group][i];
                uint32_t rank = 0;
                for (uint32_t j = 0; j < c; ++j) rank += keys[group][j] < key ? 1u : 0u;
                out[rank] = listed[group][i];
            }
        }
        if (live && c <= block && c > kBuildKeys) {
            // long lists (blocks wider than the LDS keys): insertion sort of the
            // row this wave wrote
            __threadfence();
            if (lane == 0) {
                const uint32_t *row = ovf + (size_t)slot * ovf_cap;
                for (uint32_t i = 0; i < c; ++i) {
                    const uint32_t o = row[i];
                    const uint32_t key = __float_as_uint(cQ[o].w);
                    uint32_t at = i;
                    while (at > 0 && __float_as_uint(cQ[out[at - 1]].w) > key) {
                        out[at] = out[at - 1];
                        --at;
                    }
                    out[at] = o;
                }
            }
        }
        wave_sync();  // LDS is reused by the next slots
    }
}

// ---------------------------------------------------------------------------
// Contact projection (find_segment_corrections, contact.rs:232-902, round
// sections). Terms of one candidate: the correction each endpoint would take
// and its aggregation weight, exactly as the sequential loop computes them.

struct SelfSeg {
    float4 p, q;  // endpoints (.w: rest arc length)
    float d1x, d1y, d1z;
    float radius, half_length;
    uint32_t seg, owner;
};

struct Term {
    float pen;
    float t0x, t0y, t0z, t1x, t1y, t1z;
    float aw;
};

// The term of one candidate, from its endpoints (.xyz of P2 and Q2), radius
// and segment: the capsule closest points, penetration and correction of
// find_segment_corrections (contact.rs:453-891), in its operation order.
__device__ __forceinline__ bool contact_term(const Image &g, const SelfSeg &me, const float4 &P2, const float4 &Q2,
                                             float other_radius, uint32_t other, float cf, uint32_t aggregation,
                                             Term &t) {
    const float eps_len = 1.0e-20f, eps_par = 1.0e-6f;
    const float p1x = me.p.x, p1y = me.p.y, p1z = me.p.z, q1x = me.q.x, q1y = me.q.y, q1z = me.q.z;
    const float d1x = me.d1x, d1y = me.d1y, d1z = me.d1z;
    float p2x = P2.x, p2y = P2.y, p2z = P2.z, q2x = Q2.x, q2y = Q2.y, q2z = Q2.z;
    float idx = 0.5f * (p2x + q2x - p1x - q1x);
    float idy = 0.5f * (p2y + q2y - p1y - q1y);
    float idz = 0.5f * (p2z + q2z - p1z - q1z);
    if (g.periodic[0]) {
        float sh = image_round(idx, g.length[0], g.inv_length[0]) * g.length[0];
        p2x -= sh;
        q2x -= sh;
    }
    if (g.periodic[1]) {
        float sh = image_round(idy, g.length[1], g.inv_length[1]) * g.length[1];
        p2y -= sh;
        q2y -= sh;
    }
    if (g.periodic[2]) {
        float sh = image_round(idz, g.length[2], g.inv_length[2]) * g.length[2];
        p2z -= sh;
        q2z -= sh;
    }
    float d2x = q2x - p2x, d2y = q2y - p2y, d2z = q2z - p2z;
    float mdx = 0.5f * (p2x + q2x - p1x - q1x);
    float mdy = 0.5f * (p2y + q2y - p1y - q1y);
    float mdz = 0.5f * (p2z + q2z - p1z - q1z);
    float reach = (me.half_length + 0.5f * sqrtf(d2x * d2x + d2y * d2y + d2z * d2z) + me.radius + other_radius) *
                  (1.0f + 1.0e-4f);
    if (!(mdx * mdx + mdy * mdy + mdz * mdz <= reach * reach)) return false;
    float rx = p1x - p2x, ry = p1y - p2y, rz = p1z - p2z;
    float a = d1x * d1x + d1y * d1y + d1z * d1z;
    float e = d2x * d2x + d2y * d2y + d2z * d2z;
    float f = d2x * rx + d2y * ry + d2z * rz;
    float s = 0.f, tt = 0.f;
    if (a <= eps_len && e > eps_len) {
        tt = rclamp(f / e, 0.f, 1.f);
    } else if (a > eps_len) {
        float cc = d1x * rx + d1y * ry + d1z * rz;
        if (e <= eps_len) {
            s = rclamp(-cc / a, 0.f, 1.f);
        } else {
            float b = d1x * d2x + d1y * d2y + d1z * d2z;
            float den = a * e - b * b;
            bool parallel = fabsf(den) <= eps_par * a * e;
            if (!parallel) s = rclamp((b * f - cc * e) / den, 0.f, 1.f);
            float proj = (b * s + f) / e;
            if (proj < 0.f) {
                s = rclamp(-cc / a, 0.f, 1.f);
            } else if (proj > 1.f) {
                tt = 1.f;
                s = rclamp((b - cc) / a, 0.f, 1.f);
            } else {
                tt = proj;
            }
            if (parallel) {  // contact.rs:695-726
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
                float gx = rx + d1x * s - d2x * tt, gy = ry + d1y * s - d2y * tt, gz = rz + d1z * s - d2z * tt;
                float ex = rx + d1x * es - d2x * et, ey = ry + d1y * es - d2y * et, ez = rz + d1z * es - d2z * et;
                if (ex * ex + ey * ey + ez * ez < gx * gx + gy * gy + gz * gz) {
                    s = es;
                    tt = et;
                }
            }
        }
    }
    float dx = (p2x + d2x * tt) - (p1x + d1x * s);
    float dy = (p2y + d2y * tt) - (p1y + d1y * s);
    float dz = (p2z + d2z * tt) - (p1z + d1z * s);
    float distance = sqrtf(dx * dx + dy * dy + dz * dz);
    float pen = me.radius + other_radius - distance;
    if (!(pen > 0.f)) return false;
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
        float nl = sqrtf(nx * nx + ny * ny + nz * nz);
        if (nl <= 1.0e-7f) {
            if (fabsf(d1x) <= fabsf(d1y) && fabsf(d1x) <= fabsf(d1z)) {
                nx = 0.f; ny = d1z; nz = -d1y;
            } else if (fabsf(d1y) <= fabsf(d1z)) {
                nx = -d1z; ny = 0.f; nz = d1x;
            } else {
                nx = d1y; ny = -d1x; nz = 0.f;
            }
            nl = sqrtf(nx * nx + ny * ny + nz * nz);
        }
        if (nl > 1.0e-7f) {
            nx /= nl;
            ny /= nl;
            nz /= nl;
        } else {
            nx = 1.0f;
        }
        float dom = (fabsf(nx) >= fabsf(ny) && fabsf(nx) >= fabsf(nz)) ? nx : (fabsf(ny) >= fabsf(nz) ? ny : nz);
        if (dom < 0.f) { nx = -nx; ny = -ny; nz = -nz; }
        if (me.seg > other) { nx = -nx; ny = -ny; nz = -nz; }
    }
    float w0 = 1.0f - s, w1 = s, ow0 = 1.0f - tt;
    // round sections: the director terms of the denominator are exactly zero
    float den = w0 * w0 + w1 * w1 + ow0 * ow0 + tt * tt + 0.0f;
    float mag = -cf * pen / den;
    float aw = aggregation == 1u ? pen : 1.0f;
    t.pen = pen;
    t.aw = aw;
    t.t0x = nx * mag * w0 * aw;
    t.t0y = ny * mag * w0 * aw;
    t.t0z = nz * mag * w0 * aw;
    t.t1x = nx * mag * w1 * aw;
    t.t1y = ny * mag * w1 * aw;
    t.t1z = nz * mag * w1 * aw;
    return true;
}

// Running sums of one segment, added in candidate order (contact.rs:830-891).
struct Accum {
    float c0x = 0.f, c0y = 0.f, c0z = 0.f, c1x = 0.f, c1y = 0.f, c1z = 0.f, weight = 0.f, max_pen = 0.f;

    __device__ __forceinline__ void add(const Term &t, uint32_t aggregation) {
        bool deepest = t.pen > max_pen;
        if (deepest) max_pen = t.pen;
        if (aggregation == 2u) {
            if (deepest) {
                c0x = t.t0x; c0y = t.t0y; c0z = t.t0z;
                c1x = t.t1x; c1y = t.t1y; c1z = t.t1z;
                weight = 1.0f;
            }
        } else {
            c0x += t.t0x; c0y += t.t0y; c0z += t.t0z;
            c1x += t.t1x; c1y += t.t1y; c1z += t.t1z;
            weight += t.aw;
        }
    }
};

__device__ __forceinline__ uint32_t block_max_256(uint32_t v, uint32_t *scratch4) {
    uint32_t m = wave_max_u32(v);
    if ((threadIdx.x & 63u) == 0) scratch4[threadIdx.x >> 6] = m;
    __syncthreads();
    uint32_t b = max(max(scratch4[0], scratch4[1]), max(scratch4[2], scratch4[3]));
    __syncthreads();
    return b;
}

// Slot of (specimen, metric) that workgroup `block` adds its maximum to.
__device__ __forceinline__ uint32_t *red_slot(uint32_t *red, uint32_t spec, uint32_t metric, uint32_t red_slots,
                                              uint32_t block) {
    return red + ((size_t)spec * 3u + metric) * red_slots + block % red_slots;
}

// Workgroup maximum of `bits` (256 threads) for the specimen of thread 0's
// element, added to one of that specimen's slots with one atomicMax; the
// elements of a later specimen (a workgroup straddling a specimen boundary)
// add theirs directly. Elements are ordered by specimen, so thread 0 holds
// the workgroup's first; inactive elements pass 0.
__device__ __forceinline__ void spec_block_max(uint32_t bits, uint32_t spec, uint32_t *red, uint32_t metric,
                                               uint32_t red_slots, uint32_t *scratch5) {
    if (threadIdx.x == 0) scratch5[4] = spec;
    __syncthreads();
    const uint32_t first = scratch5[4];
    if (spec != first && bits) {
        atomicMax(red_slot(red, spec, metric, red_slots, blockIdx.x), bits);
        __threadfence();
    }
    const uint32_t b = block_max_256(spec == first ? bits : 0u, scratch5);
    if (threadIdx.x == 0 && b) atomicMax(red_slot(red, first, metric, red_slots, blockIdx.x), b);
}

// Logical block for an XCD-aware dispatch: the hardware deals workgroups to
// the eight XCDs round-robin, so physical block b runs on XCD b % 8. Giving
// each XCD a contiguous range of logical blocks keeps spatially adjacent
// slots, which gather the same candidates, on one L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
    const uint32_t q = nb / 8u, r = nb % 8u, xcd = b % 8u, idx = b / 8u;
    return xcd < r ? xcd * (q + 1u) + idx : r * (q + 1u) + (xcd - r) * q + idx;
}

// A group of G lanes serves one cell-list slot: lane l evaluates candidates
// l, l+G, ... of the segment's list in parallel, then every lane of the
// group adds the contacting candidates' terms in list order (a ballot picks
// them, shuffles broadcast their terms), so the sums are the sequential
// loop's, bit for bit, with no float atomics. Overflowed lists read their
// stored stencil-order row; a list longer than that row scans the 27 cells
// sequentially in lane 0 of its group. Thread 0 of the grid first closes a
// pending rebuild (finish_neighbor_list_rebuild, neighbor.rs:424-432).
// Lists hold slots, and a listed candidate's endpoints, radius and segment
// are one 32-byte compact record in cell order, so a candidate costs one
// gather (list entry, then its record) and neighbouring groups share cache
// lines. The list build already left out the segment itself and its
// along-fiber neighbours (with the same static arc lengths), so listed
// candidates skip that test; the 27-cell scan applies it. Each lane issues
// the list entries and records of up to kContactRounds rounds before it
// evaluates any, to keep more gathers in flight; the terms are still added
// in list order.
constexpr uint32_t kContactRounds = 3;
template <int G>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_contact(DevState *st, SpecState *ss_all, uint32_t nspec, const GridParams *grids,
                                                 uint32_t red_slots, const float4 *cP, const float4 *cQ,
                                                 const float4 *slotA, const float4 *slotB,
                                                 const uint4 *slotC, const uint32_t *seg_cell,
                                                 const uint32_t *cell_counts, const uint32_t *cell_offsets,
                                                 const uint32_t *ncount, const uint32_t *nbr,
                                                 const uint32_t *nbr_off, const uint32_t *ovf,
                                                 float4 *corr0, float4 *corr1, uint32_t *red, float cf,
                                                 uint32_t aggregation, uint32_t block, uint32_t ovf_cap,
                                                 uint32_t S) {
    if (blockIdx.x == 0) {  // the rebuilds (if any) ran before this pass
        const bool pending = st->nstate[0] != 0u;
        __syncthreads();
        if (pending) {
            for (uint32_t s = threadIdx.x; s < nspec; s += 256) ss_all[s].rebuild = 0u;
            if (threadIdx.x == 0) {
                st->nstate[0] = 0;
                st->nstate[1] += 1;
            }
        }
    }
    constexpr uint32_t groups = 256 / G;
    const uint32_t lane = threadIdx.x % G, group = threadIdx.x / G;
    const uint32_t wave_lane = threadIdx.x & 63u, group_base = wave_lane - lane;
    const uint32_t slot = xcd_block(blockIdx.x, gridDim.x) * groups + group;
    // The kernel is bound by chains of dependent gathers, so every load that
    // depends only on the slot is issued at once: the slot's records, its
    // list count and offset, then the first entries of its listed row (read
    // only as far as the packed rows of the slot's 8-slot group reach;
    // entries past the count are never used as addresses).
    const uint32_t sl = slot < S ? slot : 0u;
    const uint4 sc = slotC[sl];
    const float4 P1 = cP[sl], Q1 = cQ[sl];
    const uint32_t listed_count = ncount[sl];
    const uint32_t row_off = nbr_off[sl];
    const uint32_t *row = nbr + row_off;
    // entries left in the slot's 8-slot group (0 for a slot never built)
    const uint32_t region_end = min((sl | 7u) + 1u, S) * block;
    const uint32_t row_room = row_off < region_end ? region_end - row_off : 0u;
    uint32_t first[kContactRounds];
#pragma unroll
    for (uint32_t r = 0; r < kContactRounds; ++r) {
        const uint32_t k = r * G + lane;
        first[r] = k < block && k < row_room ? row[k] : 0u;
    }
    const bool in_list = slot < S && sc.x != kNoSegment;
    const uint32_t spec = in_list ? sc.w : 0u;
    const bool active = in_list && running(ss_all, spec);
    const GridParams &g = grids[spec];
    const Image im = image_of(g);
    SelfSeg me;
    me.seg = sc.x;
    me.owner = sc.y;
    me.radius = __uint_as_float(sc.z);
    me.p = P1;  // .w: radius, unused here
    me.q = Q1;
    me.d1x = me.q.x - me.p.x;
    me.d1y = me.q.y - me.p.y;
    me.d1z = me.q.z - me.p.z;
    me.half_length = 0.5f * sqrtf(me.d1x * me.d1x + me.d1y * me.d1y + me.d1z * me.d1z);
    const uint32_t count = active ? listed_count : 0u;
    const bool listed = count <= block;
    const bool slow = count > ovf_cap;
    const uint32_t *list = listed ? row : ovf + (size_t)sl * ovf_cap;
    Accum acc;
    if (active && !slow) {
        for (uint32_t base = 0; base < count; base += G * kContactRounds) {
            uint32_t o[kContactRounds];
            float4 P2[kContactRounds], Q2[kContactRounds];
            if (base == 0 && listed) {
#pragma unroll
                for (uint32_t r = 0; r < kContactRounds; ++r) o[r] = r * G + lane < count ? first[r] : sl;
            } else {
#pragma unroll
                for (uint32_t r = 0; r < kContactRounds; ++r) {
                    const uint32_t k = base + r * G + lane;
                    o[r] = k < count ? list[k] : sl;
                }
            }
#pragma unroll
            for (uint32_t r = 0; r < kContactRounds; ++r) {
                P2[r] = cP[o[r]];
                Q2[r] = cQ[o[r]];
            }
#pragma unroll
            for (uint32_t r = 0; r < kContactRounds; ++r) {
                const uint32_t k = base + r * G + lane;
                Term t;
                bool hit = k < count && contact_term(im, me, P2[r], Q2[r], P2[r].w, __float_as_uint(Q2[r].w), cf,
                                                     aggregation, t);
                uint64_t mask = (uint64_t)__ballot(hit);
                uint32_t mine = (uint32_t)((mask >> group_base) & (G == 64 ? ~0ull : ((1ull << G) - 1ull)));
                while (mine) {
                    const int j = __builtin_ctz(mine);
                    mine &= mine - 1u;
                    const int src = (int)group_base + j;
                    Term u;
                    u.pen = __shfl(t.pen, src, 64);
                    u.t0x = __shfl(t.t0x, src, 64);
                    u.t0y = __shfl(t.t0y, src, 64);
                    u.t0z = __shfl(t.t0z, src, 64);
                    u.t1x = __shfl(t.t1x, src, 64);
                    u.t1y = __shfl(t.t1y, src, 64);
                    u.t1z = __shfl(t.t1z, src, 64);
                    u.aw = __shfl(t.aw, src, 64);
                    acc.add(u, aggregation);
                }
            }
        }
    } else if (active && slow && lane == 0) {
        // the reference's overflow scan: 27 stencil cells of the home cell at
        // the list-build positions, each cell in ascending segment order,
        // leaving out the segment itself and its along-fiber neighbours
        const uint32_t home = seg_cell[me.seg] - g.cell_base;
        const float arc_p = slotA[slot].w, arc_q = slotB[slot].w;
        for (uint32_t n = 0; n < 27; ++n) {
            uint32_t cell;
            if (!stencil_cell(g, home, n, cell)) continue;
            cell += g.cell_base;
            const uint32_t cs = cell_offsets[cell], cn = cell_counts[cell];
            for (uint32_t k = 0; k < cn; ++k) {
                const uint32_t o = cs + k;
                const uint4 ot = slotC[o];
                const float other_radius = __uint_as_float(ot.z);
                const float gap = fmaxf(slotA[o].w - arc_q, arc_p - slotB[o].w);
                const bool near = ot.y == me.owner && gap < me.radius + other_radius;
                if (ot.x == me.seg || near) continue;
                Term t;
                if (contact_term(im, me, cP[o], cQ[o], other_radius, ot.x, cf, aggregation, t)) acc.add(t, aggregation);
            }
        }
    }
    if (active && lane == 0) {
        if (aggregation != 2u && acc.weight > 1.0e-20f) {
            acc.c0x /= acc.weight; acc.c0y /= acc.weight; acc.c0z /= acc.weight;
            acc.c1x /= acc.weight; acc.c1y /= acc.weight; acc.c1z /= acc.weight;
        }
        corr0[me.seg] = make_float4(acc.c0x, acc.c0y, acc.c0z, acc.max_pen);
        corr1[me.seg] = make_float4(acc.c1x, acc.c1y, acc.c1z, 0.f);
    }
    // segment max penetration, max(0) as bits, reduced per workgroup and specimen
    __shared__ uint32_t scratch[5];
    spec_block_max(active && lane == 0 ? __float_as_uint(fmaxf(acc.max_pen, 0.f)) : 0u, spec, red, kMetricPen,
                   red_slots, scratch);
}

// ---------------------------------------------------------------------------
// curvature ratio (measure_curvature_ratio, kernels.rs:925-980) with the
// metric reductions (kernels.rs:95-127), then the convergence assessment
// (assess_reduced_metrics, kernels.rs:129-162).


__device__ __forceinline__ void assess_specimen(SpecState &sp, uint32_t pen_bits, uint32_t step_bits,
                                                uint32_t curv_bits, float penetration_tolerance,
                                                float curvature_ratio_tolerance, uint32_t iterations,
                                                uint32_t force_full, bool &keep_going) {
    const float pen = __uint_as_float(pen_bits), curv = __uint_as_float(curv_bits);
    sp.metrics[0] = pen;
    sp.metrics[1] = __uint_as_float(step_bits);
    sp.metrics[2] = curv;
    const bool stop = (force_full == 0u && pen <= penetration_tolerance && curv <= 1.0f + curvature_ratio_tolerance) ||
                      sp.control[2] >= iterations;
    if (stop) {
        sp.control[0] = 0;
    } else {
        sp.control[1] += 1;
        sp.control[2] += 1;
    }
    keep_going = !stop;
}

// Curvature ratio of every vertex of a running specimen, reduced per
// workgroup into the specimen's slots.
__global__ __launch_bounds__(256) void k_measure(const SpecState *ss, uint32_t red_slots, const float4 *pos,
                                                 const float4 *vA, const float4 *vB, uint32_t *red, float *ratios,
                                                 uint32_t V) {
    __shared__ uint32_t scratch[5];
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    float ratio = 0.f;
    uint32_t spec = 0;
    bool measured = false;
    if (v < V) {
        const float4 b4 = vB[v];
        const uint32_t flags = __float_as_uint(b4.z);
        spec = flags >> kSpecShift;
        const float kmax = vA[v].x;
        measured = running(ss, spec) && !(flags & kInactive);
        if (measured && kmax > 0.f && (flags & kHasPrev) && (flags & kHasNext)) {
            float4 a = pos[v - 1], b = pos[v], c = pos[v + 1];
            float ax = b.x - a.x, ay = b.y - a.y, az = b.z - a.z;
            float bx = c.x - b.x, by = c.y - b.y, bz = c.z - b.z;
            float l1 = sqrtf(ax * ax + ay * ay + az * az);
            float l2 = sqrtf(bx * bx + by * by + bz * bz);
            if (l1 > 1.0e-7f && l2 > 1.0e-7f) {
                float cosine = rclamp((ax * bx + ay * by + az * bz) / (l1 * l2), -1.f, 1.f);
                float dual = 0.5f * (l1 + l2);
                float kappa = 2.0f * sqrtf(0.5f * (1.0f - cosine)) / dual;
                ratio = kappa / kmax;
            }
        }
    }
    // every measured vertex keeps its ratio (the reference's curvature_ratio
    // buffer), which the compaction metrics read
    if (measured) ratios[v] = ratio;
    spec_block_max(__float_as_uint(fmaxf(ratio, 0.f)), spec, red, kMetricCurv, red_slots, scratch);
}

// One workgroup of 256 threads: folds the slots of every running specimen
// and assesses it (assess_reduced_metrics, kernels.rs:129-162). Runs after
// the measuring kernel in stream order, so every slot is final.
__global__ __launch_bounds__(256) void k_assess(SpecState *ss, uint32_t nspec, uint32_t red_slots, uint32_t *red,
                                                float penetration_tolerance, float curvature_ratio_tolerance,
                                                uint32_t iterations, uint32_t force_full) {
    __shared__ uint32_t scratch[5];
    const uint32_t t = threadIdx.x;
    if (nspec == 1u && red_slots == 256u) {
        const bool run = running(ss, 0);
        if (!run) return;  // uniform across the workgroup
        uint32_t pen_bits = block_max_256(*red_slot(red, 0, kMetricPen, 256u, t), scratch);
        uint32_t step_bits = block_max_256(*red_slot(red, 0, kMetricStep, 256u, t), scratch);
        uint32_t curv_bits = block_max_256(*red_slot(red, 0, kMetricCurv, 256u, t), scratch);
        __shared__ uint32_t keep;
        if (t == 0) {
            bool keep_going = false;
            assess_specimen(ss[0], pen_bits, step_bits, curv_bits, penetration_tolerance, curvature_ratio_tolerance,
                            iterations, force_full, keep_going);
            keep = keep_going ? 1u : 0u;
        }
        __syncthreads();
        *red_slot(red, 0, kMetricPen, 256u, t) = 0u;
        *red_slot(red, 0, kMetricCurv, 256u, t) = 0u;
        // the step maxima are recomputed by this iteration's contact
        // application; a stopping pass leaves them for the next batch's
        // first assessment
        if (keep) *red_slot(red, 0, kMetricStep, 256u, t) = 0u;
        return;
    }
    for (uint32_t sp = t; sp < nspec; sp += blockDim.x) {
        if (!running(ss, sp)) continue;
        uint32_t m[3] = {0u, 0u, 0u};
        for (uint32_t metric = 0; metric < 3; ++metric)
            for (uint32_t k = 0; k < red_slots; ++k) m[metric] = max(m[metric], *red_slot(red, sp, metric, red_slots, k));
        bool keep_going = false;
        assess_specimen(ss[sp], m[kMetricPen], m[kMetricStep], m[kMetricCurv], penetration_tolerance,
                        curvature_ratio_tolerance, iterations, force_full, keep_going);
        for (uint32_t k = 0; k < red_slots; ++k) {
            *red_slot(red, sp, kMetricPen, red_slots, k) = 0u;
            *red_slot(red, sp, kMetricCurv, red_slots, k) = 0u;
            if (keep_going) *red_slot(red, sp, kMetricStep, red_slots, k) = 0u;
        }
    }
}

// ---------------------------------------------------------------------------
// One Jacobi pass of stretch + rest-chord bend (find_internal_corrections and
// apply_internal_corrections, kernels.rs:1027-1163, 1417-1453), read from
// `in`, written to `out`; the last pass also applies the contact corrections
// (apply_contact_corrections, kernels.rs:1455-1537).

__device__ __forceinline__ void accumulate_pull(float4 o, float x, float y, float z, float rest, float k, float &cx,
                                                float &cy, float &cz) {
    float dx = o.x - x, dy = o.y - y, dz = o.z - z;
    float len = sqrtf(dx * dx + dy * dy + dz * dz);
    if (len > 1.0e-7f) {
        float mag = 0.5f * k * (len - rest) / len;
        cx += dx * mag;
        cy += dy * mag;
        cz += dz * mag;
    }
}

__device__ __forceinline__ float wall_move(float moved, uint32_t periodic, float lo, float hi, float extent) {
    return periodic ? moved : rclamp(moved, lo + extent, hi - extent);
}

__global__ __launch_bounds__(256) void k_vertex_pass(const SpecState *ss, const GridParams *grids,
                                                     uint32_t red_slots, const float4 *in, float4 *out,
                                                     const float4 *vA, const float4 *vB, const float4 *corr0,
                                                     const float4 *corr1, float4 *wall, uint32_t *red, uint32_t V,
                                                     float stretch, float bend, float max_step, uint32_t do_internal,
                                                     uint32_t do_contact, uint32_t first_pass, uint32_t track_walls) {
    __shared__ uint32_t scratch[5];
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    float step = 0.f;
    uint32_t spec = 0;
    float4 sb = make_float4(0.f, 0.f, 0.f, 0.f);
    if (v < V) {
        sb = vB[v];
        spec = __float_as_uint(sb.z) >> kSpecShift;
    }
    if (v < V && running(ss, spec) && !(__float_as_uint(sb.z) & kInactive)) {
        const GridParams &g = grids[spec];
        float4 me = in[v];
        float4 sa = vA[v];
        uint32_t flags = __float_as_uint(sb.z);
        float4 result = me;
        // cleared once per iteration (clear_wall_reactions, kernels.rs:14-24)
        float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
        if (track_walls && !first_pass) w = wall[v];
        if (!(flags & kPinned)) {
            float x = me.x, y = me.y, z = me.z;
            if (do_internal) {
                float cx = 0.f, cy = 0.f, cz = 0.f;
                if (flags & kHasPrev) accumulate_pull(in[v - 1], x, y, z, sa.y, stretch, cx, cy, cz);
                if (flags & kHasNext) accumulate_pull(in[v + 1], x, y, z, sa.z, stretch, cx, cy, cz);
                if (flags & kHasPrev2) accumulate_pull(in[v - 2], x, y, z, sb.x, bend, cx, cy, cz);
                if (flags & kHasNext2) accumulate_pull(in[v + 2], x, y, z, sb.y, bend, cx, cy, cz);
                float mx = x + cx, my = y + cy, mz = z + cz;
                x = wall_move(mx, g.periodic[0], g.lower[0], g.upper[0], sa.w);
                y = wall_move(my, g.periodic[1], g.lower[1], g.upper[1], sa.w);
                z = wall_move(mz, g.periodic[2], g.lower[2], g.upper[2], sa.w);
                if (!g.periodic[0]) w.x += mx - x;
                if (!g.periodic[1]) w.y += my - y;
                if (!g.periodic[2]) w.z += mz - z;
            }
            if (do_contact) {
                uint32_t next_seg = v - __float_as_uint(sb.w);  // a vertex's next segment index is v - fiber
                float cx = 0.f, cy = 0.f, cz = 0.f;
                if (flags & kHasPrev) {
                    float4 c = corr1[next_seg - 1];
                    cx += c.x;
                    cy += c.y;
                    cz += c.z;
                }
                if (flags & kHasNext) {
                    float4 c = corr0[next_seg];
                    cx += c.x;
                    cy += c.y;
                    cz += c.z;
                }
                float len = sqrtf(cx * cx + cy * cy + cz * cz);
                if (len > max_step) {
                    float sc = max_step / len;
                    cx *= sc;
                    cy *= sc;
                    cz *= sc;
                }
                float mx = x + cx, my = y + cy, mz = z + cz;
                x = wall_move(mx, g.periodic[0], g.lower[0], g.upper[0], sa.w);
                y = wall_move(my, g.periodic[1], g.lower[1], g.upper[1], sa.w);
                z = wall_move(mz, g.periodic[2], g.lower[2], g.upper[2], sa.w);
                if (!g.periodic[0]) w.x += mx - x;
                if (!g.periodic[1]) w.y += my - y;
                if (!g.periodic[2]) w.z += mz - z;
                step = sqrtf(cx * cx + cy * cy + cz * cz);
            }
            result.x = x;
            result.y = y;
            result.z = z;
        }
        out[v] = result;
        if (track_walls) wall[v] = w;
    }
    if (do_contact) spec_block_max(__float_as_uint(fmaxf(step, 0.f)), spec, red, kMetricStep, red_slots, scratch);
}

// Stretch and rest-chord bend of one vertex in one Jacobi pass, with its
// wall clamp (the internal half of k_vertex_pass): x, y, z in and out, the
// reactions added to w.
__device__ __forceinline__ void internal_move(const GridParams &g, uint32_t flags, float4 sa, float4 sb,
                                              float4 pm2, float4 pm1, float4 pp1, float4 pp2, float stretch,
                                              float bend, float &x, float &y, float &z, float4 &w) {
    float cx = 0.f, cy = 0.f, cz = 0.f;
    if (flags & kHasPrev) accumulate_pull(pm1, x, y, z, sa.y, stretch, cx, cy, cz);
    if (flags & kHasNext) accumulate_pull(pp1, x, y, z, sa.z, stretch, cx, cy, cz);
    if (flags & kHasPrev2) accumulate_pull(pm2, x, y, z, sb.x, bend, cx, cy, cz);
    if (flags & kHasNext2) accumulate_pull(pp2, x, y, z, sb.y, bend, cx, cy, cz);
    float mx = x + cx, my = y + cy, mz = z + cz;
    x = wall_move(mx, g.periodic[0], g.lower[0], g.upper[0], sa.w);
    y = wall_move(my, g.periodic[1], g.lower[1], g.upper[1], sa.w);
    z = wall_move(mz, g.periodic[2], g.lower[2], g.upper[2], sa.w);
    if (!g.periodic[0]) w.x += mx - x;
    if (!g.periodic[1]) w.y += my - y;
    if (!g.periodic[2]) w.z += mz - z;
}

// All Jacobi passes of an iteration in one launch. A workgroup owns 256
// vertices and stages them with 2 x passes more on each side (positions and
// static records in LDS); pass p (0-based) computes its own vertices and
// 2 (passes - 1 - p) more on each side, which is what the passes after it
// read (a vertex reads up to two neighbours on each side), so the last pass
// has every owned vertex exact without any exchange between workgroups. Each
// vertex takes k_vertex_pass's operations in its order (the last pass adds
// the contact corrections and the step metric; a vertex's wall reactions
// accumulate in a register across the passes, as k_vertex_pass accumulates
// them in memory). Reads `in`, writes `out` (never in place: neighbouring
// workgroups read `in`'s halo).
constexpr uint32_t kPassOwned = 256;
constexpr uint32_t kMaxFusedPasses = 32;
__host__ __device__ inline size_t vertex_passes_lds_bytes(uint32_t passes) {
    return (size_t)(kPassOwned + 4u * passes) * 4u * sizeof(float4);  // two position buffers, vA, vB
}
__global__ __launch_bounds__(256) void k_vertex_passes(const SpecState *ss, const GridParams *grids,
                                                       uint32_t red_slots, const float4 *in, float4 *out,
                                                       const float4 *vA, const float4 *vB, const float4 *corr0,
                                                       const float4 *corr1, float4 *wall, uint32_t *red,
                                                       uint32_t V, float stretch, float bend, float max_step,
                                                       uint32_t passes, uint32_t track_walls) {
    extern __shared__ float4 lds4[];
    __shared__ uint32_t scratch[5];
    const uint32_t H = 2u * passes, n = kPassOwned + 2u * H;
    float4 *X0 = lds4, *X1 = lds4 + n, *A = lds4 + 2u * n, *B = lds4 + 3u * n;
    const uint32_t tid = threadIdx.x;
    const uint32_t b0 = blockIdx.x * kPassOwned;
    // local index i holds vertex b0 - H + i; a vertex outside [0, V) is never
    // read (a vertex reads only neighbours its flags name)
    for (uint32_t i = tid; i < n; i += 256u) {
        const int64_t v = (int64_t)b0 - (int64_t)H + (int64_t)i;
        if (v >= 0 && v < (int64_t)V) {
            X0[i] = in[v];
            A[i] = vA[v];
            B[i] = vB[v];
        } else {
            B[i] = make_float4(0.f, 0.f, __uint_as_float(kInactive), 0.f);
        }
    }
    __syncthreads();
    const uint32_t own = H + tid;
    const uint32_t vo = b0 + tid;
    float4 w = make_float4(0.f, 0.f, 0.f, 0.f);  // cleared once per iteration (kernels.rs:14-24)
    float step = 0.f;
    uint32_t spec_o = 0;
    float4 *src = X0, *dst = X1;
    for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t r = 2u * (passes - 1u - p);
        const bool last = p + 1u == passes;
        // this thread's vertices of the pass: its own, and one of the 2r
        // halo vertices when tid < 2r
        for (uint32_t k = 0; k < 2u; ++k) {
            uint32_t i;
            if (k == 0) {
                i = own;
            } else {
                if (tid >= 2u * r) break;
                i = tid < r ? H - r + tid : H + kPassOwned + (tid - r);
            }
            const float4 sb = B[i];
            const uint32_t flags = __float_as_uint(sb.z);
            const uint32_t spec = flags >> kSpecShift;
            const float4 me = src[i];
            if (k == 0) spec_o = spec;
            if ((flags & kInactive) || !running(ss, spec)) {
                dst[i] = me;
                continue;
            }
            const GridParams &g = grids[spec];
            const float4 sa = A[i];
            float4 result = me;
            float4 wr = k == 0 ? w : make_float4(0.f, 0.f, 0.f, 0.f);
            if (!(flags & kPinned)) {
                float x = me.x, y = me.y, z = me.z;
                internal_move(g, flags, sa, sb, (flags & kHasPrev2) ? src[i - 2u] : me,
                              (flags & kHasPrev) ? src[i - 1u] : me, (flags & kHasNext) ? src[i + 1u] : me,
                              (flags & kHasNext2) ? src[i + 2u] : me, stretch, bend, x, y, z, wr);
                if (last) {  // only owned vertices run the last pass
                    const uint32_t next_seg = vo - __float_as_uint(sb.w);
                    float cx = 0.f, cy = 0.f, cz = 0.f;
                    if (flags & kHasPrev) {
                        float4 c = corr1[next_seg - 1];
                        cx += c.x;
                        cy += c.y;
                        cz += c.z;
                    }
                    if (flags & kHasNext) {
                        float4 c = corr0[next_seg];
                        cx += c.x;
                        cy += c.y;
                        cz += c.z;
                    }
                    float len = sqrtf(cx * cx + cy * cy + cz * cz);
                    if (len > max_step) {
                        float sc = max_step / len;
                        cx *= sc;
                        cy *= sc;
                        cz *= sc;
                    }
                    float mx = x + cx, my = y + cy, mz = z + cz;
                    x = wall_move(mx, g.periodic[0], g.lower[0], g.upper[0], sa.w);
                    y = wall_move(my, g.periodic[1], g.lower[1], g.upper[1], sa.w);
                    z = wall_move(mz, g.periodic[2], g.lower[2], g.upper[2], sa.w);
                    if (!g.periodic[0]) wr.x += mx - x;
                    if (!g.periodic[1]) wr.y += my - y;
                    if (!g.periodic[2]) wr.z += mz - z;
                    step = sqrtf(cx * cx + cy * cy + cz * cz);
                }
                result.x = x;
                result.y = y;
                result.z = z;
            }
            dst[i] = result;
            if (k == 0) w = wr;
        }
        __syncthreads();
        float4 *t = src;
        src = dst;
        dst = t;
    }
    if (vo < V) {
        const uint32_t flags = __float_as_uint(B[own].z);
        if (!(flags & kInactive) && running(ss, spec_o)) {
            out[vo] = src[own];
            if (track_walls) wall[vo] = w;
        }
    }
    spec_block_max(__float_as_uint(fmaxf(step, 0.f)), vo < V ? spec_o : 0u, red, kMetricStep, red_slots, scratch);
}

// ---------------------------------------------------------------------------
// Per-fiber curvature Gauss–Seidel (project_fiber_curvature_in_place and
// project_curvature_triplet_in_place, kernels.rs:1165-1415), then the
// neighbour-list displacement trigger (flag_neighbor_list_displacement,
// neighbor.rs:443-465) for the fiber's vertices.

// Projection of the triplet (previous, center, next) in material order.
// Returns true when the projection branch runs; the moved positions and the
// wall reactions of each vertex are returned in mp/mc/mn and rp/rc/rn (zero
// for a pinned vertex, which keeps its position).
__device__ __forceinline__ bool triplet_projection(float3 pp, float3 pc, float3 pn, float kmax_center, bool pin_p,
                                                   bool pin_c, bool pin_n, float ext_p, float ext_c, float ext_n,
                                                   const GridParams &g, float stiffness, float margin, float3 &mp,
                                                   float3 &mc, float3 &mn, float3 &rp, float3 &rc, float3 &rn) {
    float kmax = kmax_center * (1.0f - margin);
    if (!(kmax > 0.f)) return false;
    float ux = pc.x - pp.x, uy = pc.y - pp.y, uz = pc.z - pp.z;
    float vx = pn.x - pc.x, vy = pn.y - pc.y, vz = pn.z - pc.z;
    float l1 = sqrtf(ux * ux + uy * uy + uz * uz);
    float l2 = sqrtf(vx * vx + vy * vy + vz * vz);
    if (!(l1 > 1.0e-7f && l2 > 1.0e-7f)) return false;
    float unx = ux / l1, uny = uy / l1, unz = uz / l1;
    float vnx = vx / l2, vny = vy / l2, vnz = vz / l2;
    float cosine = rclamp(unx * vnx + uny * vny + unz * vnz, -1.f, 1.f);
    float sine_half = 0.25f * kmax * (l1 + l2);
    if (!(sine_half < 1.0f)) return false;
    float min_cos = 1.0f - 2.0f * sine_half * sine_half;
    float violation = min_cos - cosine;
    if (!(violation > 0.f)) return false;
    float sine = sqrtf(fmaxf(1.0f - cosine * cosine, 0.f));
    if (sine > 1.0e-3f) violation = (acosf_f32(cosine) - acosf_f32(min_cos)) * sine;
    if (!(violation > 0.f)) return false;
    float gux = (vnx - cosine * unx) / l1, guy = (vny - cosine * uny) / l1, guz = (vnz - cosine * unz) / l1;
    float gvx = (unx - cosine * vnx) / l2, gvy = (uny - cosine * vny) / l2, gvz = (unz - cosine * vnz) / l2;
    float gcx = gux - gvx, gcy = guy - gvy, gcz = guz - gvz;
    float fwp = pin_p ? 0.f : 1.f, fwc = pin_c ? 0.f : 1.f, fwn = pin_n ? 0.f : 1.f;
    float den = fwp * (gux * gux + guy * guy + guz * guz) + fwc * (gcx * gcx + gcy * gcy + gcz * gcz) +
                fwn * (gvx * gvx + gvy * gvy + gvz * gvz);
    if (!(den > 1.0e-12f)) return false;
    float scale = stiffness * violation / den;
    const float gp[3] = {gux, guy, guz}, gc[3] = {gcx, gcy, gcz}, gn[3] = {gvx, gvy, gvz};
    const float P[3] = {pp.x, pp.y, pp.z}, Cc[3] = {pc.x, pc.y, pc.z}, N[3] = {pn.x, pn.y, pn.z};
    float oP[3], oC[3], oN[3], wP[3], wC[3], wN[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float m_p = P[k] - fwp * scale * gp[k];
        float m_c = Cc[k] + fwc * scale * gc[k];
        float m_n = N[k] + fwn * scale * gn[k];
        float v_p = m_p, v_c = m_c, v_n = m_n;
        wP[k] = wC[k] = wN[k] = 0.f;
        if (!g.periodic[k]) {
            v_p = rclamp(m_p, g.lower[k] + ext_p, g.upper[k] - ext_p);
            v_c = rclamp(m_c, g.lower[k] + ext_c, g.upper[k] - ext_c);
            v_n = rclamp(m_n, g.lower[k] + ext_n, g.upper[k] - ext_n);
            wP[k] = m_p - v_p;
            wC[k] = m_c - v_c;
            wN[k] = m_n - v_n;
        }
        oP[k] = fwp > 0.f ? v_p : P[k];
        oC[k] = fwc > 0.f ? v_c : Cc[k];
        oN[k] = fwn > 0.f ? v_n : N[k];
    }
    mp = make_float3(oP[0], oP[1], oP[2]);
    mc = make_float3(oC[0], oC[1], oC[2]);
    mn = make_float3(oN[0], oN[1], oN[2]);
    rp = make_float3(wP[0], wP[1], wP[2]);
    rc = make_float3(wC[0], wC[1], wC[2]);
    rn = make_float3(wN[0], wN[1], wN[2]);
    return true;
}

// Curvature sweep with 4 lanes per fiber (16 fibers per wave), each fiber
// staged in LDS as float4 (x, y, z, wall extent) plus (kmax, weight).
//
// A group alternates between two modes. Screening: the next L triplets of
// the sweep are tested from the current positions with a cheap bound; a
// triplet is passed over only when its cosine exceeds the limit cosine by
// more than 1e-5, far beyond the few-ulp error of the bound, so the exact
// test would find no violation either. Exact: the triplet at the cursor is
// evaluated with the sequential sweep's f32 operations in its order by three
// lanes, one per axis (cross-axis sums in x, y, z order through quad
// broadcasts); when its projection runs, those lanes move its three vertices.
// A projection at sweep position o changes the triplets at o + 1 and o + 2,
// which are then evaluated exactly in turn with the positions carried in
// registers; screening results for positions past o + 2 stay valid and pick
// the next exact candidate. Every triplet the sweep would project is thus
// projected in the same order from the same values: the result is the
// sequential Gauss–Seidel sweep's, bit for bit. A sweep direction that
// projects nothing ends the cleanup, since every later pass would revisit
// the same unchanged triplets.
__device__ __forceinline__ float quad_bcast(float x, int k) {
    // v_mov_b32 with a DPP quad permutation: every lane of a quad reads lane k
    switch (k) {
        case 0: return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x00, 0xF, 0xF, false));
        case 1: return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x55, 0xF, 0xF, false));
        default: return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xAA, 0xF, 0xF, false));
    }
}
__device__ __forceinline__ float axis_sum(float x) {  // (x + y) + z over lanes 0, 1, 2 of the quad
    return (quad_bcast(x, 0) + quad_bcast(x, 1)) + quad_bcast(x, 2);
}
__device__ __forceinline__ void lds_order() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

__host__ __device__ inline size_t curvature_group_stride_floats(uint32_t max_vertices, uint32_t track_walls) {
    const size_t n = (max_vertices + 1u) & ~1u;  // keeps every region 16-byte aligned
    return n * (4u + 2u + (track_walls ? 4u : 0u)) + 4u;  // + a sink slot
}

template <int L>
__global__ __launch_bounds__(64) void k_curvature_gs_group(DevState *st, SpecState *ss, const GridParams *grids,
                                                           const float4 *in, float4 *pos, const float4 *refpos,
                                                           const float4 *vA, const float4 *vB, const uint4 *fspan,
                                                           float4 *wall, uint32_t F, uint32_t sweeps, float stiffness,
                                                           float margin, uint32_t track_walls,
                                                           uint32_t max_vertices) {
    static_assert(L == 4, "group width (16 and 32 lanes: k_curvature_gs_lean; 8: k_curvature_gs_seq)");
    constexpr uint32_t G = 64 / L;
    extern __shared__ float4 lds4[];
    const uint32_t lane = threadIdx.x, group = lane / L, gl = lane % L, gbase = group * L;
    const uint32_t nv = (max_vertices + 1u) & ~1u;
    float4 *P = (float4 *)((float *)lds4 + group * curvature_group_stride_floats(max_vertices, track_walls));
    float4 *WR = P + nv;                                     // wall reactions (when tracked)
    float2 *KW = (float2 *)(track_walls ? WR + nv : WR);     // (kmax, 1 - pinned)
    float *Pf = (float *)P;
    const uint32_t fiber = blockIdx.x * G + group;
    uint4 span = make_uint4(0u, 0u, 0u, 0u);
    if (fiber < F) span = fspan[fiber];
    // a specimen that stopped, or an inactive fiber, keeps its positions
    if (!running(ss, span.z) || (span.y && (__float_as_uint(vB[span.x].z) & kInactive))) span.y = 0u;
    const uint32_t start = span.x, n = span.y;
    const GridParams &g = grids[span.z];
    const float half_skin = 0.5f * g.skin;
    for (uint32_t i = gl; i < n; i += L) {
        const float4 p = in[start + i];
        const float4 a = vA[start + i];
        P[i] = make_float4(p.x, p.y, p.z, a.w);
        KW[i] = make_float2(a.x, (__float_as_uint(vB[start + i].z) & kPinned) ? 0.f : 1.f);
        if (track_walls) WR[i] = wall[start + i];
    }
    // the axis of an exact lane (lanes 0..2, and 4..6 of the two-quad form)
    // and that axis's walls, held in registers
    const uint32_t k = (gl & 3u) < 3u ? (gl & 3u) : 0u;
    const bool walled = !g.periodic[k];
    const float lo = g.lower[k], hi = g.upper[k];
    __syncthreads();
    const uint32_t trip = n >= 3 ? n - 2 : 0;  // triplet t has center t + 1
    const uint64_t gmask = (1ull << (L % 64)) - 1ull;
    bool done = trip == 0 || sweeps == 0;
    bool any = false;
    uint32_t sweep = 0, dir = 0, cursor = 0;
    while (true) {
        if (!done && cursor >= trip) {
            if (!any || (dir == 1 && sweep + 1 >= sweeps)) {
                done = true;
            } else {
                any = false;
                cursor = 0;
                if (dir == 0) {
                    dir = 1;
                } else {
                    dir = 0;
                    ++sweep;
                }
            }
        }
        if (__ballot(!done) == 0) break;
        // screen the next L triplets
        bool maybe = false;
        const uint32_t order = cursor + gl;
        if (!done && order < trip) {
            const uint32_t t = dir == 0 ? order : trip - 1 - order;
            const float4 a = P[t], b = P[t + 1], c = P[t + 2];
            const float kmax = KW[t + 1].x * (1.0f - margin);
            const float ux = b.x - a.x, uy = b.y - a.y, uz = b.z - a.z;
            const float vx = c.x - b.x, vy = c.y - b.y, vz = c.z - b.z;
            const float uu = ux * ux + uy * uy + uz * uz, vv = vx * vx + vy * vy + vz * vz;
            const float uv = ux * vx + uy * vy + uz * vz;
            const float ru = __builtin_amdgcn_rsqf(uu), rv = __builtin_amdgcn_rsqf(vv);
            const float cos_est = uv * ru * rv;
            const float half_est = 0.25f * kmax * (uu * ru + vv * rv);
            const float min_cos_est = 1.0f - 2.0f * half_est * half_est;
            maybe = (kmax > 0.f) & !(cos_est - min_cos_est > 1.0e-5f);
        }
        const uint64_t ballot = (uint64_t)__ballot(maybe);
        const uint64_t mine = (ballot >> gbase) & gmask;
        if (!done && mine == 0) cursor += L;
        if (ballot == 0) continue;
        // exact evaluation of each group's first candidate
        const uint32_t j = mine ? (uint32_t)__builtin_ctzll(mine) : 0u;
        const uint32_t t = dir == 0 ? cursor + j : trip - 1 - (cursor + j);
        bool hit = false;
        if (mine != 0 && gl < 3) {
            float pp = Pf[4 * t + k], pc = Pf[4 * (t + 1) + k], pn = Pf[4 * (t + 2) + k];
            const float ep = P[t].w, ec = P[t + 1].w, en = P[t + 2].w;
            const float2 kwp = KW[t], kwc = KW[t + 1], kwn = KW[t + 2];
            const float kmax = kwc.x * (1.0f - margin);
            const float u = pc - pp, v = pn - pc;
            const float l1 = sqrtf(axis_sum(u * u));
            const float l2 = sqrtf(axis_sum(v * v));
            const float un = u / l1, vn = v / l2;
            const float cosine = rclamp(axis_sum(un * vn), -1.f, 1.f);
            const float sine_half = 0.25f * kmax * (l1 + l2);
            const float min_cos = 1.0f - 2.0f * sine_half * sine_half;
            float violation = min_cos - cosine;
            bool ok = kmax > 0.f && l1 > 1.0e-7f && l2 > 1.0e-7f && sine_half < 1.0f && violation > 0.f;
            const float sine = sqrtf(fmaxf(1.0f - cosine * cosine, 0.f));
            // one arc cosine per lane: lane 0 the turning angle, lane 1 the limit
            const float angle = acosf_f32(gl == 1 ? min_cos : cosine);
            const float turn = quad_bcast(angle, 0), limit_angle = quad_bcast(angle, 1);
            if (sine > 1.0e-3f) violation = (turn - limit_angle) * sine;
            ok = ok && violation > 0.f;
            const float gu = (vn - cosine * un) / l1;
            const float gv = (un - cosine * vn) / l2;
            const float gc = gu - gv;
            const float fwp = kwp.y, fwc = kwc.y, fwn = kwn.y;
            const float den = fwp * axis_sum(gu * gu) + fwc * axis_sum(gc * gc) + fwn * axis_sum(gv * gv);
            ok = ok && den > 1.0e-12f;
            if (ok) {
                const float scale = stiffness * violation / den;
                const float m_p = pp - fwp * scale * gu;
                const float m_c = pc + fwc * scale * gc;
                const float m_n = pn + fwn * scale * gv;
                float v_p = m_p, v_c = m_c, v_n = m_n;
                if (walled) {
                    v_p = rclamp(m_p, lo + ep, hi - ep);
                    v_c = rclamp(m_c, lo + ec, hi - ec);
                    v_n = rclamp(m_n, lo + en, hi - en);
                    if (track_walls) {
                        float *W = (float *)WR;
                        if (fwp > 0.f) W[4 * t + k] += m_p - v_p;
                        if (fwc > 0.f) W[4 * (t + 1) + k] += m_c - v_c;
                        if (fwn > 0.f) W[4 * (t + 2) + k] += m_n - v_n;
                    }
                }
                if (fwp > 0.f) Pf[4 * t + k] = v_p;
                if (fwc > 0.f) Pf[4 * (t + 1) + k] = v_c;
                if (fwn > 0.f) Pf[4 * (t + 2) + k] = v_n;
            }
            hit = ok;
        }
        const uint64_t hits = (uint64_t)__ballot(hit && gl == 0);
        if (mine != 0) {
            cursor += j + 1;
            if ((hits >> gbase) & 1ull) any = true;
        }
        // one wave: its LDS accesses complete in order; keep the compiler's
        lds_order();
    }
    __syncthreads();
    const float limit = half_skin * half_skin;
    bool flag = false;
    for (uint32_t i = gl; i < n; i += L) {
        const uint32_t v = start + i;
        const float4 old = in[v], q = P[i];
        pos[v] = make_float4(q.x, q.y, q.z, old.w);
        if (track_walls) wall[v] = make_float4(WR[i].x, WR[i].y, WR[i].z, 0.f);
        const float4 r = refpos[v];
        const float dx = q.x - r.x, dy = q.y - r.y, dz = q.z - r.z;
        if (dx * dx + dy * dy + dz * dz > limit) flag = true;
    }
    if (flag) ss[span.z].rebuild = 1u;
    if (__ballot(flag) != 0 && lane == 0) st->nstate[0] = 1u;
}

// Curvature sweep, straight-line form (L = 16 or 32 lanes per fiber).
//
// The same screening and exact two-quad projection as k_curvature_gs_group,
// in the same order, so the result is again the sequential sweep's, bit for
// bit; what changes is the control flow. The step is issue-bound with about
// one wave per SIMD, and divergent branches (each an exec-mask save and
// restore plus an instruction-buffer refill) cost as much as the arithmetic,
// so the loop is written without them: the sweep bookkeeping is a set of
// selects, screening reads a clamped position instead of branching on its
// validity, and every lane of a wave runs the exact evaluation whenever any
// group has a candidate; groups without one evaluate a straight dummy
// triplet with no bend limit, which never projects. Stores that must not
// land are sent to a per-group sink slot in LDS.
//
// Cross-lane sums use DPP with bound_ctrl set: every lane of these patterns
// reads an active lane of its own row, so the flag never supplies a value,
// and it lets the compiler fold the lane move into the add that uses it.
__device__ __forceinline__ float quad_lane(float x, int k) {
    switch (k) {
        case 0: return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x00, 0xF, 0xF, true));
        case 1: return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x55, 0xF, 0xF, true));
        default: return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0xAA, 0xF, 0xF, true));
    }
}
__device__ __forceinline__ float quad_axis_sum(float x) {  // (x + y) + z over lanes 0, 1, 2 of the quad
    return (quad_lane(x, 0) + quad_lane(x, 1)) + quad_lane(x, 2);
}
// lanes 0..3 of an 8-lane group read lanes 4..7 and lanes 4..7 read 0..3
__device__ __forceinline__ float other_quad(float x, bool second) {
    const int from_next = __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x104, 0xF, 0xF, true);
    const int from_prev = __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x114, 0xF, 0xF, true);
    return __int_as_float(second ? from_prev : from_next);
}

// Correctly rounded square root for x = 0, x = +inf or x >= 2^-96: the IEEE
// expansion (the hardware root s, then s - 1 ulp or s + 1 ulp where an exact
// fused residual shows the root is past it) without the scaling that only
// inputs below 2^-96 take, so for these inputs it is sqrtf bit for bit. The
// curvature sweep's inputs meet it: 1 - c^2 for f32 c in [-1, 1] is 0 or at
// least 2^-24, acos's z is 0 or at least 2^-25 where its root is used, and a
// squared segment length below 2^-96 belongs to a segment far shorter than
// the 1e-7 below which the projection is not taken, whatever the root.
__device__ __forceinline__ float sqrt_rn_normal(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float down = __uint_as_float(__float_as_uint(s) - 1u);
    const float up = __uint_as_float(__float_as_uint(s) + 1u);
    const float r = __builtin_fmaf(-down, s, x) <= 0.f ? down : s;
    const float q = __builtin_fmaf(-up, s, x) > 0.f ? up : r;
    return __builtin_amdgcn_class(x, 0x260) ? x : q;  // +-0 and +inf are their own roots
}
struct SqrtRnNormal {
    __device__ float operator()(float z) const { return sqrt_rn_normal(z); }
};
// f32::clamp for non-NaN x: when lo <= hi the median of (x, lo, hi); when
// lo > hi clamp's second test takes hi whatever x, which the median of
// (x, hi, hi) also gives
__device__ __forceinline__ float clamp_med3(float x, float lo, float hi) {
    return __builtin_amdgcn_fmed3f(x, fminf(lo, hi), hi);
}

template <int L>
__global__ __launch_bounds__(64) void k_curvature_gs_lean(DevState *st, SpecState *ss, const GridParams *grids,
                                                          const float4 *in, float4 *pos, const float4 *refpos,
                                                          const float4 *vA, const float4 *vB, const uint4 *fspan,
                                                          float4 *wall, uint32_t F, uint32_t sweeps, float stiffness,
                                                          float margin, uint32_t track_walls,
                                                          uint32_t max_vertices) {
    static_assert(L == 16 || L == 32, "group width");
    constexpr uint32_t G = 64 / L;
    extern __shared__ float4 lds4[];
    const uint32_t lane = threadIdx.x, group = lane / L, gl = lane % L, gbase = group * L;
    const uint32_t nv = (max_vertices + 1u) & ~1u;
    float4 *P = (float4 *)((float *)lds4 + group * curvature_group_stride_floats(max_vertices, track_walls));
    float4 *WR = P + nv;                                     // wall reactions (when tracked)
    float2 *KW = (float2 *)(track_walls ? WR + nv : WR);     // (kmax, 1 - pinned)
    float *Pf = (float *)P;
    float *W = (float *)WR;
    float *sink = (float *)(KW + nv);                        // absorbs stores that must not land
    const uint32_t fiber = blockIdx.x * G + group;
    uint4 span = make_uint4(0u, 0u, 0u, 0u);
    if (fiber < F) span = fspan[fiber];
    // a specimen that stopped, or an inactive fiber, keeps its positions
    if (!running(ss, span.z) || (span.y && (__float_as_uint(vB[span.x].z) & kInactive))) span.y = 0u;
    const uint32_t start = span.x, n = span.y;
    const GridParams &g = grids[span.z];
    const float half_skin = 0.5f * g.skin;
    for (uint32_t i = gl; i < n; i += L) {
        const float4 p = in[start + i];
        const float4 a = vA[start + i];
        P[i] = make_float4(p.x, p.y, p.z, a.w);
        KW[i] = make_float2(a.x, (__float_as_uint(vB[start + i].z) & kPinned) ? 0.f : 1.f);
        if (track_walls) WR[i] = wall[start + i];
    }
    // Lane roles, repeated in every 8 lanes of the group: lanes 0..2 hold the
    // axes of the triplet's first segment (previous -> center), lanes 4..6 those
    // of its second (center -> next); lanes 3 and 7 (and the lanes past 8 of a
    // wider group) compute the same values as another lane and store nothing.
    const bool second = (gl & 4u) != 0u;
    const uint32_t k = (gl & 3u) < 3u ? (gl & 3u) : 0u;
    const uint32_t o = second ? 1u : 0u;
    const bool writer = gl < 8u && (gl & 3u) < 3u;
    const bool walled = !g.periodic[k];
    const float lo = g.lower[k], hi = g.upper[k];
    __syncthreads();
    const uint32_t trip = n >= 3 ? n - 2 : 0;  // triplet t has center t + 1
    const uint64_t gmask = (1ull << (L % 64)) - 1ull;
    bool done = trip == 0 || sweeps == 0;
    bool any = false;
    uint32_t sweep = 0, dir = 0, cursor = 0;
    while (true) {
        // the end of a sweep direction: finish, or turn around
        const bool end = !done && cursor >= trip;
        const bool finish = end && (!any || (dir == 1 && sweep + 1 >= sweeps));
        const bool turn = end && !finish;
        done = done || finish;
        sweep = (turn && dir == 1u) ? sweep + 1u : sweep;
        dir = turn ? 1u - dir : dir;
        cursor = turn ? 0u : cursor;
        any = any && !turn;
        if (__ballot(!done) == 0) break;
        // the operands of the block's first position, read with the screen:
        // after a projection the next candidate is most often that one
        const uint32_t tf = done ? 0u : (dir == 0 ? cursor : trip - 1 - cursor);
        float a0 = Pf[4 * (tf + o) + k], a1 = Pf[4 * (tf + o + 1) + k];
        float e0 = P[tf + o].w, e1 = P[tf + o + 1].w;
        float2 kwp = KW[tf], kwc = KW[tf + 1], kwn = KW[tf + 2];
        // screen the next L positions of every group still sweeping
        const uint32_t order = cursor + gl;
        const bool valid = !done && order < trip;
        const uint32_t ts = valid ? (dir == 0 ? order : trip - 1 - order) : 0u;
        bool maybe;
        {
            const float4 a = P[ts], b = P[ts + 1], c = P[ts + 2];
            const float kmax = KW[ts + 1].x * (1.0f - margin);
            const float ux = b.x - a.x, uy = b.y - a.y, uz = b.z - a.z;
            const float vx = c.x - b.x, vy = c.y - b.y, vz = c.z - b.z;
            const float uu = ux * ux + uy * uy + uz * uz, vv = vx * vx + vy * vy + vz * vz;
            const float uv = ux * vx + uy * vy + uz * vz;
            const float ru = __builtin_amdgcn_rsqf(uu), rv = __builtin_amdgcn_rsqf(vv);
            const float cos_est = uv * ru * rv;
            const float half_est = 0.25f * kmax * (uu * ru + vv * rv);
            const float min_cos_est = 1.0f - 2.0f * half_est * half_est;
            maybe = valid & (kmax > 0.f) & !(cos_est - min_cos_est > 1.0e-5f);
        }
        const uint64_t mine = ((uint64_t)__ballot(maybe) >> gbase) & gmask;
        const bool found = mine != 0;
        const uint32_t j = found ? (uint32_t)__builtin_ctzll(mine) : 0u;
        cursor = (!done && !found) ? cursor + L : cursor;
        if (__ballot(found) == 0) continue;
        // exact evaluation of each group's first candidate
        const uint32_t ord = cursor + j;
        const uint32_t t = found ? (dir == 0 ? ord : trip - 1 - ord) : 0u;
        if (__ballot(found && j != 0u) != 0) {
            if (found && j != 0u) {
                a0 = Pf[4 * (t + o) + k];
                a1 = Pf[4 * (t + o + 1) + k];
                e0 = P[t + o].w;
                e1 = P[t + o + 1].w;
                kwp = KW[t];
                kwc = KW[t + 1];
                kwn = KW[t + 2];
            }
        }
        if (!found) {  // a straight unit segment pair along x with no limit: never projects
            a0 = 0.f;
            a1 = k == 0u ? 1.f : 0.f;
            kwc.x = 0.f;
        }
        // the sequential projection's operations, in its order
        // The sequential projection's operations, in its order. Values the
        // formula pairs (l1 and l2, un and vn, ...) sit on the two quads; where
        // the formula combines them symmetrically (l1 + l2, un . vn) each quad
        // uses its own value first, which commutes exactly.
        const float kmax = kwc.x * (1.0f - margin);
        const float d = a1 - a0;
        const float len = sqrt_rn_normal(quad_axis_sum(d * d));
        const float dn = d / len;
        const float dn_o = other_quad(dn, second), len_o = other_quad(len, second);
        const float cosine = __builtin_amdgcn_fmed3f(quad_axis_sum(dn * dn_o), -1.f, 1.f);  // finite when used
        const float sine_half = 0.25f * kmax * (len + len_o);
        const float min_cos = 1.0f - 2.0f * sine_half * sine_half;
        float violation = min_cos - cosine;
        bool ok = kmax > 0.f && len > 1.0e-7f && len_o > 1.0e-7f && sine_half < 1.0f && violation > 0.f;
        if (__ballot(ok) != 0) {
            const float sine = sqrt_rn_normal(fmaxf(1.0f - cosine * cosine, 0.f));
            // the first quad takes the turning angle, the second the limit;
            // turn - limit is the negated difference on the second quad
            const float angle = acosf_f32_with(second ? min_cos : cosine, SqrtRnNormal{});
            const float diff = angle - other_quad(angle, second);
            if (sine > 1.0e-3f) violation = (second ? -diff : diff) * sine;
            ok = ok && violation > 0.f;
            // (vn - cos un) / l1 in the first quad (gu), (un - cos vn) / l2 in
            // the second (gv); gu - gv is exact on the first quad, which alone
            // moves the center, and only its square is used on the second
            const float gq = (dn_o - cosine * dn) / len;
            const float gq_o = other_quad(gq, second);
            const float gc = gq - gq_o;
            const float fwp = kwp.y, fwc = kwc.y, fwn = kwn.y;
            const float sg = quad_axis_sum(gq * gq), sg_o = other_quad(sg, second);
            const float suu = second ? sg_o : sg, svv = second ? sg : sg_o;
            const float den = fwp * suu + fwc * quad_axis_sum(gc * gc) + fwn * svv;
            ok = ok && den > 1.0e-12f;
            const float scale = stiffness * violation / den;
            // first quad: previous (- gu) and center (+ gc); second quad: next (+ gv)
            const float f0 = fwp, f1 = second ? fwn : fwc;
            const float m0 = a0 - fwp * scale * gq;
            const float m1 = a1 + f1 * scale * (second ? gq : gc);
            const float v0 = walled ? clamp_med3(m0, lo + e0, hi - e0) : m0;
            const float v1 = walled ? clamp_med3(m1, lo + e1, hi - e1) : m1;
            const bool s0 = writer && ok && !second && f0 > 0.f, s1 = writer && ok && f1 > 0.f;
            const uint32_t i0 = t + o, i1 = t + o + 1;
            if (track_walls) {
                float *w0 = (s0 && walled) ? &W[4 * i0 + k] : &sink[k];
                float *w1 = (s1 && walled) ? &W[4 * i1 + k] : &sink[k];
                *w0 = *w0 + (m0 - v0);
                *w1 = *w1 + (m1 - v1);
            }
            *(s0 ? &Pf[4 * i0 + k] : &sink[k]) = v0;
            *(s1 ? &Pf[4 * i1 + k] : &sink[k]) = v1;
        }
        const bool hit = ((uint64_t)__ballot(ok && gl == 0) >> gbase) & 1ull;
        cursor = found ? ord + 1u : cursor;
        any = any || hit;
        // one wave: its LDS accesses complete in order; keep the compiler's
        lds_order();
    }
    __syncthreads();
    const float limit = half_skin * half_skin;
    bool flag = false;
    for (uint32_t i = gl; i < n; i += L) {
        const uint32_t v = start + i;
        const float4 old = in[v], q = P[i];
        pos[v] = make_float4(q.x, q.y, q.z, old.w);
        if (track_walls) wall[v] = make_float4(WR[i].x, WR[i].y, WR[i].z, 0.f);
        const float4 r = refpos[v];
        const float dx = q.x - r.x, dy = q.y - r.y, dz = q.z - r.z;
        if (dx * dx + dy * dy + dz * dz > limit) flag = true;
    }
    if (flag) ss[span.z].rebuild = 1u;
    if (__ballot(flag) != 0 && lane == 0) st->nstate[0] = 1u;
}

// Curvature sweep with 8 lanes per fiber walking every triplet in sweep
// order (no screen). On the felt's critical fibers about nine visits in ten
// project, so a screen saves few exact evaluations and costs a loop
// iteration of its own; the sequential walk instead keeps the triplet in
// registers: a step's three vertices pass to the next step between the two
// quads, and the one vertex new to the next triplet is read from LDS while
// the step runs (no step before it changes that vertex).
//
// Lane roles as in k_curvature_gs_lean. The exact step is the sequential
// projection's operations in its order, with the scalar part spread over
// the quad's lanes: lane 3 takes the square root of 1 - cos^2 while lanes
// 0..2 take acos's sqrt(z); lane 1 evaluates acos's Q(z) while the others
// evaluate P(z) (one Horner chain with per-lane coefficients, where a zero
// start and the constant term 1 reproduce Q's own evaluation exactly); and
// one division gives P/Q on lane 0 and acos's low-order correction c on
// lane 1. Each value is the one the sequential walk computes, bit for bit.
//
// LDS per fiber: positions (x, y, z, kmax (1 - margin)) and one weight per
// vertex (0 for a pinned vertex); with walls tracked, the weight is the
// fourth float of the vertex's wall reactions. The clamp bounds
// (min(lo + e, hi - e), hi - e) of a walled axis, or (-inf, inf), depend on
// the vertex only through its wall extent e, the fiber's radius, so each
// lane holds its axis's bounds in registers.
__host__ __device__ inline size_t curvature_seq_stride_floats(uint32_t max_vertices, uint32_t track_walls) {
    const size_t n = (max_vertices + 1u) & ~1u;
    if (track_walls) return n * 8u + 4u;                     // + a sink slot
    return n * 4u + ((n + 3u) & ~(size_t)3u) + 4u;
}

// The other quad of a lane's fiber in the sequential sweep, whose two quads
// sit 8 lanes apart in a row of 16 lanes: one rotation of each row by 8
// lanes (DPP row_ror:8) exchanges them.
__device__ __forceinline__ float swap_quads(float x) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x128, 0xF, 0xF, false));
}
// the IEEE square root for x = +-0, +inf, NaN and x >= 2^-96 (where the
// scaling of tiny inputs is not taken): the hardware root s, then s - 1 ulp
// or s + 1 ulp where an exact fused residual shows the root is past it.
// For x = 0 the residual at s - 1 ulp is NaN and at s + 1 ulp zero, so s
// stays; for +inf both residuals are NaN.
__device__ __forceinline__ float sqrt_rn_plain(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float down = __uint_as_float(__float_as_uint(s) - 1u);
    const float up = __uint_as_float(__float_as_uint(s) + 1u);
    const float r = __builtin_fmaf(-down, s, x) <= 0.f ? down : s;
    return __builtin_fmaf(-up, s, x) > 0.f ? up : r;
}

// The IEEE division's own instruction sequence (reciprocal, one refinement,
// two fused corrections, then the special-case fixup) without the operand
// scaling of v_div_scale / v_div_fmas. That scaling is taken only when the
// numerator is below 2^-103, the denominator or the quotient is subnormal,
// or their exponents differ by 96 or more; otherwise the two sequences are
// the same operations on the same values. The sweep's divisions stay clear of
// those ranges whenever their result is used: position differences and
// their numerators are 0 or above 2^-60, segment lengths above 1e-7, acos's
// Q(z) near 1, its correction's denominator above 2^-13, and the projection
// denominator between 1e-12 and 2^60. A zero numerator goes through the
// fixup, which returns the correctly signed zero.
__device__ __forceinline__ float recip_refined(float b) {  // the sequence's refined reciprocal
    const float r = __builtin_amdgcn_rcpf(b);
    return __builtin_fmaf(__builtin_fmaf(-b, r, 1.0f), r, r);
}
// the sequence after its reciprocal: divisions by one denominator share it
__device__ __forceinline__ float div_rn_unscaled_with(float a, float b, float r) {
    float q = a * r;
    float rem = __builtin_fmaf(-b, q, a);
    q = __builtin_fmaf(rem, r, q);
    rem = __builtin_fmaf(-b, q, a);
    q = __builtin_fmaf(rem, r, q);
    return __builtin_amdgcn_div_fixupf(q, b, a);
}
__device__ __forceinline__ float div_rn_unscaled(float a, float b) {
    float r = recip_refined(b);
    float q = a * r;
    float rem = __builtin_fmaf(-b, q, a);
    q = __builtin_fmaf(rem, r, q);
    rem = __builtin_fmaf(-b, q, a);
    q = __builtin_fmaf(rem, r, q);
    return __builtin_amdgcn_div_fixupf(q, b, a);
}

// Per-lane constants and LDS views of one fiber's group in the sequential
// sweep.
struct SeqLanes {
    float4 *P;
    float *Pf, *W, *WT, *sink;
    float2 bound;  // this lane's axis: (lower, upper) clamp of a vertex
    bool second, qlane, walled;
    uint32_t ql, k, o;
    float hc4, hc3, hc2, hc1, hc0;
};

// One sweep direction of every group of the wave (DIR 0: forward, 1:
// backward), steps in lockstep up to the wave's longest chain; a group whose
// chain is shorter, or that is done, takes no step. Returns whether the
// group's direction projected anything.
//
// The triplet index t is the same for every group of the wave (forward 0 up,
// backward from the wave's last triplet down): a group whose chain is
// shorter idles while t lies past its last triplet and walks its own
// vertices' LDS records along (its region holds the longest fiber's count),
// so when t reaches its chain the registers hold its triplet. Indices stay
// inside the group's region: t + 3 <= the longest fiber's vertex count.
template <int DIR, bool WALLS>
__device__ __forceinline__ bool seq_direction(const SeqLanes &q, bool on, uint32_t trip, uint32_t maxtrip,
                                              float stiffness) {
    constexpr uint32_t WS = WALLS ? 4u : 1u;  // floats from one vertex's weight to the next
    const float pio2_hi = f32_from_bits(0x3fc90fdau), pio2_lo = f32_from_bits(0x33a22168u);
    const float pi = f32_from_bits(0x40490fdau);
    const float hc5 = f32_from_bits(0x3811ef08u);
    const bool second = q.second, qlane = q.qlane;
    const uint32_t k = q.k, o = q.o, ql = q.ql;
    float4 *P = q.P;
    float *Pf = q.Pf, *WT = q.WT;
    const float2 b0 = q.bound, b1 = q.bound;
    uint32_t t = DIR == 1 ? maxtrip - 1u : 0u;  // maxtrip >= 1
    float a0 = Pf[4u * (t + o) + k], a1 = Pf[4u * (t + o + 1u) + k];
    // the triplet's static records (center kmax, the three weights) are read
    // one step ahead, so no step waits on LDS for them
    float kmax_n = P[t + 1u].w, fwp_n = WT[WS * t], fwc_n = WT[WS * (t + 1u)], fwn_n = WT[WS * (t + 2u)];
    // these reads complete before the loop (s_waitcnt lgkmcnt(0)): with
    // nothing read pending on either way into the loop, its first step does
    // not wait for the stores of the step before
    __builtin_amdgcn_s_waitcnt(0xc07f);
    bool any = false;
    for (uint32_t c = 0; c < maxtrip; ++c) {
        const bool act = on && t < trip;
        const float kmax = kmax_n, fwp = fwp_n, fwc = fwc_n, fwn = fwn_n, qkmax = 0.25f * kmax;
        const uint32_t tn = DIR == 0 ? t + 1u : (t > 0u ? t - 1u : 0u);
        // the next triplet's static records, then the vertex it adds (read
        // last, so that the wait for it at the end of the step covers all
        // three reads and none is pending when the next step begins)
        kmax_n = P[tn + 1u].w;
        if (DIR == 0) {
            fwp_n = fwc;
            fwc_n = fwn;
            fwn_n = WT[WS * (tn + 2u)];
        } else {
            fwn_n = fwc;
            fwc_n = fwp;
            fwp_n = WT[WS * tn];
        }
        const float nextv = Pf[4u * (DIR == 0 ? tn + 2u : tn) + k];
        // the sequential projection's operations, in its order
        const float d = a1 - a0;
        const float len = sqrt_rn_plain(quad_axis_sum(d * d));
        const float rlen = recip_refined(len);  // shared by d / len and the gradient's / len
        const float dn = div_rn_unscaled_with(d, len, rlen);
        const float dn_o = swap_quads(dn), len_o = swap_quads(len);
        const float cosine = __builtin_amdgcn_fmed3f(quad_axis_sum(dn * dn_o), -1.f, 1.f);  // finite when used
        const float sine_half = qkmax * (len + len_o);  // (0.25 kmax) (l1 + l2)
        const float min_cos = 1.0f - 2.0f * sine_half * sine_half;
        float violation = min_cos - cosine;
        bool ok = act && kmax > 0.f && len > 1.0e-7f && len_o > 1.0e-7f && sine_half < 1.0f && violation > 0.f;
        float w0 = a0, w1 = a1;  // this lane's two positions after the step
        if (__ballot(ok) != 0) {
            // (vn - cos un) / l1 in the first quad (gu), (un - cos vn) / l2 in
            // the second (gv); gu - gv is exact on the first quad, which alone
            // moves the center, and only its square is used on the second.
            // None of it waits for the arc cosines below.
            const float gq = div_rn_unscaled_with(dn_o - cosine * dn, len, rlen);
            const float gq_o = swap_quads(gq);
            const float gc = gq - gq_o;
            const float sg = quad_axis_sum(gq * gq), sg_o = swap_quads(sg);
            const float suu = second ? sg_o : sg, svv = second ? sg : sg_o;
            const float den = fwp * suu + fwc * quad_axis_sum(gc * gc) + fwn * svv;
            // acos of the turning cosine (first quad) and of the limit (second)
            const float x = second ? min_cos : cosine;
            const uint32_t ix = __float_as_uint(x) & 0x7fffffffu;
            const bool negative = (__float_as_uint(x) >> 31) != 0u, small = ix < 0x3f000000u;
            const float z = small ? x * x : (1.0f - fabsf(x)) * 0.5f;
            const float sq = sqrt_rn_plain(ql == 3u ? 1.0f - cosine * cosine : z);
            const float sine = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(sq), 0xFF, 0xF, 0xF, true));
            const float s = quad_lane(sq, 0);
            float h = (qlane ? 0.f : z) * hc5;
            h = h + q.hc4;
            h = z * h;
            h = h + q.hc3;
            h = z * h;
            h = h + q.hc2;
            h = z * h;
            h = h + q.hc1;
            h = z * h;
            h = h + q.hc0;
            const float p = z * h;
            const float s_hi = __uint_as_float(__float_as_uint(s) & 0xfffff000u);
            const float qq = quad_lane(h, 1);  // read by every lane (a cross-lane read outside any branch)
            const float quo = div_rn_unscaled(qlane ? z - s_hi * s_hi : p, qlane ? s + s_hi : qq);
            const float r = quad_lane(quo, 0), cc = quad_lane(quo, 1);
            float near_zero = pio2_hi - (x - (pio2_lo - x * r));
            float below = pi - 2.0f * (s + (r * s - pio2_lo));
            float above = 2.0f * (s_hi + (r * s + cc));
            // the three branches are selected, not branched to
            __asm__ volatile("" : "+v"(near_zero), "+v"(below), "+v"(above));
            float angle = small ? near_zero : (negative ? below : above);
            angle = ix <= 0x32800000u ? pio2_hi + pio2_lo : angle;                        // |x| <= 2^-26
            angle = ix == 0x3f800000u ? (negative ? pi + 2.0f * pio2_lo : 0.0f) : angle;  // |x| = 1
            // turn - limit on the first quad; the second holds its negation
            const float diff = angle - swap_quads(angle);
            violation = sine > 1.0e-3f ? (second ? -diff : diff) * sine : violation;
            ok = ok && violation > 0.f && den > 1.0e-12f;
            const float scale = div_rn_unscaled(stiffness * violation, den);
            // first quad: previous (- gu) and center (+ gc); second quad: next (+ gv)
            const float f1 = second ? fwn : fwc;
            const float m0 = a0 - fwp * scale * gq;
            const float m1 = a1 + f1 * scale * (second ? gq : gc);
            const float v0 = __builtin_amdgcn_fmed3f(m0, b0.x, b0.y);
            const float v1 = __builtin_amdgcn_fmed3f(m1, b1.x, b1.y);
            const bool moved0 = ok && !second && fwp > 0.f, moved1 = ok && f1 > 0.f;
            if (WALLS) {
                float *wr0 = (moved0 && q.walled && ql < 3u) ? &q.W[4u * (t + o) + k] : &q.sink[k];
                float *wr1 = (moved1 && q.walled && ql < 3u) ? &q.W[4u * (t + o + 1u) + k] : &q.sink[k];
                *wr0 = *wr0 + (m0 - v0);
                *wr1 = *wr1 + (m1 - v1);
            }
            w0 = moved0 ? v0 : a0;
            w1 = moved1 ? v1 : a1;
            any = any || ok;  // outside this branch no lane's ok holds
        }
        // the first quad holds the triplet's previous and center vertices,
        // the second quad's w1 its next; each lane stores what it holds (the
        // second quad's w1 twice). A group that takes no step stores to its
        // sink instead: with no branch, every step issues the same LDS
        // operations, and the wait for the next vertex's read need not wait
        // for these stores.
        {
            float *s0 = act ? &Pf[4u * (second ? t + 2u : t) + k] : &q.sink[k];
            float *s1 = act ? &Pf[4u * (t + o + 1u) + k] : &q.sink[k];
            *s0 = second ? w1 : w0;
            *s1 = w1;
        }
        // the next triplet's operands: forward (center, next | next, new),
        // backward (new, previous | previous, center)
        const float x0 = swap_quads(w0), x1 = swap_quads(w1);
        if (DIR == 0) {
            a0 = w1;
            a1 = second ? nextv : x1;
        } else {
            a0 = second ? x0 : nextv;
            a1 = second ? x1 : w0;
        }
        t = tn;
        // one wave: its LDS accesses complete in order; keep the compiler's
        lds_order();
    }
    return any;
}

template <bool WALLS>
__global__ __launch_bounds__(64) void k_curvature_gs_seq(DevState *st, SpecState *ss, const GridParams *grids,
                                                         const float4 *in, float4 *pos, const float4 *refpos,
                                                         const float4 *vA, const float4 *vB, const uint4 *fspan,
                                                         float4 *wall, uint32_t F, uint32_t sweeps, float stiffness,
                                                         float margin, uint32_t track_walls,
                                                         uint32_t max_vertices) {
    constexpr uint32_t L = 8, G = 8;
    extern __shared__ float4 lds4[];
    // a row of 16 lanes serves two fibers: fiber h (0 or 1) of the row has
    // its first quad at lanes 4h..4h+3 and its second 8 lanes on, so that
    // swap_quads is one rotation; gl numbers a fiber's lanes 0..7 (quad lane,
    // plus 4 in the second quad)
    const uint32_t lane = threadIdx.x, within = lane & 15u;
    const uint32_t group = 2u * (lane >> 4) + ((within >> 2) & 1u);
    const uint32_t gl = (within & 3u) | ((within >> 1) & 4u);
    const uint32_t nv = (max_vertices + 1u) & ~1u;
    SeqLanes q;
    q.P = (float4 *)((float *)lds4 + group * curvature_seq_stride_floats(max_vertices, WALLS));
    float4 *WR = q.P + nv;                                 // wall reactions and weights (walls tracked)
    q.WT = WALLS ? (float *)WR + 3 : (float *)WR;          // weights
    q.sink = WALLS ? (float *)(WR + nv) : q.WT + ((nv + 3u) & ~3u);  // absorbs stores that must not land
    q.Pf = (float *)q.P;
    q.W = (float *)WR;
    float4 *P = q.P;
    const uint32_t fiber = blockIdx.x * G + group;
    uint4 span = make_uint4(0u, 0u, 0u, 0u);
    if (fiber < F) span = fspan[fiber];
    // a specimen that stopped, or an inactive fiber, keeps its positions
    if (!running(ss, span.z) || (span.y && (__float_as_uint(vB[span.x].z) & kInactive))) span.y = 0u;
    const uint32_t start = span.x, n = span.y;
    const GridParams &g = grids[span.z];
    const float half_skin = 0.5f * g.skin;
    const float inf = __builtin_inff();
    for (uint32_t i = gl; i < n; i += L) {
        const float4 p = in[start + i];
        const float4 a = vA[start + i];
        P[i] = make_float4(p.x, p.y, p.z, a.x * (1.0f - margin));
        const float weight = (__float_as_uint(vB[start + i].z) & kPinned) ? 0.f : 1.f;
        if (WALLS) {
            const float4 r = wall[start + i];
            WR[i] = make_float4(r.x, r.y, r.z, weight);
        } else {
            q.WT[i] = weight;
        }
    }
    // fiber lanes 0..2 hold the axes of the triplet's first segment
    // (previous -> center), lanes 4..6 those of its second (center -> next);
    // lanes 3 and 7 repeat lane 0's and lane 4's vector values
    q.second = (gl & 4u) != 0u;
    q.ql = gl & 3u;
    q.k = q.ql < 3u ? q.ql : 0u;
    q.o = q.second ? 1u : 0u;
    q.walled = !g.periodic[q.k];
    {
        // the wall extent of every vertex is the fiber's radius
        const float e = n > 0u ? vA[start].w : 0.f;
        const float lo = g.lower[q.k] + e, hi = g.upper[q.k] - e;
        q.bound = q.walled ? make_float2(fminf(lo, hi), hi) : make_float2(-inf, inf);
    }
    // acos (fdlibm e_acosf.c): R(z) = P(z) / Q(z), P on lanes 0, 2, 3 and Q on
    // lane 1 of each quad
    q.qlane = q.ql == 1u;
    q.hc4 = q.qlane ? f32_from_bits(0x3d9dc62eu) : f32_from_bits(0x3a4f7f04u);
    q.hc3 = q.qlane ? f32_from_bits(0xbf303361u) : f32_from_bits(0xbd241146u);
    q.hc2 = q.qlane ? f32_from_bits(0x4001572du) : f32_from_bits(0x3e4e0aa8u);
    q.hc1 = q.qlane ? f32_from_bits(0xc019d139u) : f32_from_bits(0xbea6b090u);
    q.hc0 = q.qlane ? 1.0f : f32_from_bits(0x3e2aaaabu);
    __syncthreads();
    const uint32_t trip = n >= 3 ? n - 2 : 0;  // triplet t has center t + 1
    const uint32_t maxtrip = (uint32_t)__builtin_amdgcn_readfirstlane((int)wave_max_u32(trip));  // wave-uniform
    // a sweep direction that projects nothing ends the group's cleanup: the
    // directions after it would visit the same unchanged triplets
    bool done = trip == 0;
    for (uint32_t sweep = 0; sweep < sweeps; ++sweep) {
        if (__ballot(!done) == 0) break;
        done = done || !seq_direction<0, WALLS>(q, !done, trip, maxtrip, stiffness);
        if (__ballot(!done) == 0) break;
        done = done || !seq_direction<1, WALLS>(q, !done, trip, maxtrip, stiffness);
    }
    __syncthreads();
    const float limit = half_skin * half_skin;
    bool flag = false;
    for (uint32_t i = gl; i < n; i += L) {
        const uint32_t v = start + i;
        const float4 old = in[v], qv = P[i];
        pos[v] = make_float4(qv.x, qv.y, qv.z, old.w);
        if (WALLS) wall[v] = make_float4(WR[i].x, WR[i].y, WR[i].z, 0.f);
        const float4 r = refpos[v];
        const float dx = qv.x - r.x, dy = qv.y - r.y, dz = qv.z - r.z;
        if (dx * dx + dy * dy + dz * dz > limit) flag = true;
    }
    if (flag) ss[span.z].rebuild = 1u;
    if (__ballot(flag) != 0 && lane == 0) st->nstate[0] = 1u;
}

// One fiber per wave: the same screening and exact evaluation as
// k_curvature_gs_group, with every decision wave-uniform. After a screen
// the wave's three axis lanes run the exact chain in a tight loop: the
// triplet positions slide through registers, the next vertex is loaded
// while the current triplet is evaluated, and the loop leaves only when the
// next position is neither changed by the last projection nor a screened
// candidate. LDS accesses of one wave complete in order, so the screen that
// follows reads the chain's writes without a barrier.
__global__ __launch_bounds__(64) void k_curvature_gs_wave64(DevState *st, SpecState *ss,
                                                            const GridParams *grids, const float4 *in, float4 *pos,
                                                            const float4 *refpos, const float4 *vA, const float4 *vB,
                                                            const uint4 *fspan, float4 *wall, uint32_t F,
                                                            uint32_t sweeps, float stiffness, float margin,
                                                            uint32_t track_walls, uint32_t max_vertices) {
    extern __shared__ float4 lds4[];
    const uint32_t lane = threadIdx.x;
    const uint32_t nv = (max_vertices + 1u) & ~1u;
    float4 *P = lds4;
    float4 *WR = P + nv;
    float2 *KW = (float2 *)(track_walls ? WR + nv : WR);
    float *Pf = (float *)P;
    const uint32_t fiber = blockIdx.x;
    if (fiber >= F) return;
    const uint4 span = fspan[fiber];
    if (!running(ss, span.z) || (__float_as_uint(vB[span.x].z) & kInactive)) return;
    const uint32_t start = span.x, n = span.y;
    const GridParams &g = grids[span.z];
    const float half_skin = 0.5f * g.skin;
    for (uint32_t i = lane; i < n; i += 64) {
        const float4 p = in[start + i];
        const float4 a = vA[start + i];
        P[i] = make_float4(p.x, p.y, p.z, a.w);
        KW[i] = make_float2(a.x, (__float_as_uint(vB[start + i].z) & kPinned) ? 0.f : 1.f);
        if (track_walls) WR[i] = wall[start + i];
    }
    __syncthreads();
    const uint32_t trip = n >= 3 ? n - 2 : 0;
    const uint32_t k = lane < 3 ? lane : 0u;
    const bool walled = !g.periodic[k];
    const float lo = g.lower[k], hi = g.upper[k];
    for (uint32_t sweep = 0; sweep < sweeps && trip > 0; ++sweep) {
        bool stop = false;
        for (uint32_t dir = 0; dir < 2; ++dir) {
            bool any = false;
            uint32_t cursor = 0;
            while (cursor < trip) {
                // screen the next 64 positions
                bool maybe = false;
                const uint32_t order = cursor + lane;
                if (order < trip) {
                    const uint32_t t = dir == 0 ? order : trip - 1 - order;
                    const float4 a = P[t], b = P[t + 1], c = P[t + 2];
                    const float kmax = KW[t + 1].x * (1.0f - margin);
                    const float ux = b.x - a.x, uy = b.y - a.y, uz = b.z - a.z;
                    const float vx = c.x - b.x, vy = c.y - b.y, vz = c.z - b.z;
                    const float uu = ux * ux + uy * uy + uz * uz, vv = vx * vx + vy * vy + vz * vz;
                    const float uv = ux * vx + uy * vy + uz * vz;
                    const float ru = __builtin_amdgcn_rsqf(uu), rv = __builtin_amdgcn_rsqf(vv);
                    const float cos_est = uv * ru * rv;
                    const float half_est = 0.25f * kmax * (uu * ru + vv * rv);
                    const float min_cos_est = 1.0f - 2.0f * half_est * half_est;
                    maybe = (kmax > 0.f) & !(cos_est - min_cos_est > 1.0e-5f);
                }
                const uint64_t cand = (uint64_t)__ballot(maybe);
                if (cand == 0) {
                    cursor += 64;
                    continue;
                }
                const uint32_t wbase = cursor;
                cursor += (uint32_t)__builtin_ctzll(cand);
                // exact chain on the three axis lanes; control stays uniform
                uint32_t next_cursor = cursor;
                bool chain_any = false;
                if (lane < 3) {
                    uint32_t t = dir == 0 ? cursor : trip - 1 - cursor;
                    float pp = Pf[4 * t + k], pc = Pf[4 * (t + 1) + k], pn = Pf[4 * (t + 2) + k];
                    float ep = P[t].w, ec = P[t + 1].w, en = P[t + 2].w;
                    float2 kwp = KW[t], kwc = KW[t + 1], kwn = KW[t + 2];
                    int32_t last_hit = -3;
                    while (true) {
                        // the vertex that enters when the triplet slides by one
                        const bool more = cursor + 1 < trip;
                        const uint32_t tn = more ? (dir == 0 ? t + 3 : t - 1) : t;
                        const float pnext = Pf[4 * tn + k];
                        const float enext = P[tn].w;
                        const float2 kwnext = KW[tn];
                        const float kmax = kwc.x * (1.0f - margin);
                        const float u = pc - pp, v = pn - pc;
                        const float l1 = sqrtf(axis_sum(u * u));
                        const float l2 = sqrtf(axis_sum(v * v));
                        const float un = u / l1, vn = v / l2;
                        const float cosine = rclamp(axis_sum(un * vn), -1.f, 1.f);
                        const float sine_half = 0.25f * kmax * (l1 + l2);
                        const float min_cos = 1.0f - 2.0f * sine_half * sine_half;
                        float violation = min_cos - cosine;
                        bool ok = kmax > 0.f && l1 > 1.0e-7f && l2 > 1.0e-7f && sine_half < 1.0f && violation > 0.f;
                        const float sine = sqrtf(fmaxf(1.0f - cosine * cosine, 0.f));
                        const float angle = acosf_f32(lane == 1 ? min_cos : cosine);
                        const float turn = quad_bcast(angle, 0), limit_angle = quad_bcast(angle, 1);
                        if (sine > 1.0e-3f) violation = (turn - limit_angle) * sine;
                        ok = ok && violation > 0.f;
                        const float gu = (vn - cosine * un) / l1;
                        const float gv = (un - cosine * vn) / l2;
                        const float gc = gu - gv;
                        const float fwp = kwp.y, fwc = kwc.y, fwn = kwn.y;
                        const float den =
                            fwp * axis_sum(gu * gu) + fwc * axis_sum(gc * gc) + fwn * axis_sum(gv * gv);
                        ok = ok && den > 1.0e-12f;
                        // lanes 0..2 hold identical scalars: the decision is uniform
                        const bool hit = __builtin_amdgcn_readfirstlane(ok ? 1 : 0) != 0;
                        if (hit) {
                            const float scale = stiffness * violation / den;
                            const float m_p = pp - fwp * scale * gu;
                            const float m_c = pc + fwc * scale * gc;
                            const float m_n = pn + fwn * scale * gv;
                            float v_p = m_p, v_c = m_c, v_n = m_n;
                            if (walled) {
                                v_p = rclamp(m_p, lo + ep, hi - ep);
                                v_c = rclamp(m_c, lo + ec, hi - ec);
                                v_n = rclamp(m_n, lo + en, hi - en);
                                if (track_walls) {
                                    float *W = (float *)WR;
                                    if (fwp > 0.f) W[4 * t + k] += m_p - v_p;
                                    if (fwc > 0.f) W[4 * (t + 1) + k] += m_c - v_c;
                                    if (fwn > 0.f) W[4 * (t + 2) + k] += m_n - v_n;
                                }
                            }
                            if (fwp > 0.f) pp = v_p;
                            if (fwc > 0.f) pc = v_c;
                            if (fwn > 0.f) pn = v_n;
                            Pf[4 * t + k] = pp;
                            Pf[4 * (t + 1) + k] = pc;
                            Pf[4 * (t + 2) + k] = pn;
                            chain_any = true;
                            last_hit = (int32_t)cursor;
                        }
                        ++cursor;
                        if (!more) break;
                        // continue while the next position changed with a
                        // projection or was screened as a candidate
                        const uint32_t rel = cursor - wbase;
                        const bool screened = rel < 64 && ((cand >> rel) & 1ull);
                        if (!((int32_t)cursor <= last_hit + 2 || screened)) break;
                        if (dir == 0) {
                            pp = pc; pc = pn; pn = pnext;
                            ep = ec; ec = en; en = enext;
                            kwp = kwc; kwc = kwn; kwn = kwnext;
                            t = t + 1;
                        } else {
                            pn = pc; pc = pp; pp = pnext;
                            en = ec; ec = ep; ep = enext;
                            kwn = kwc; kwc = kwp; kwp = kwnext;
                            t = t - 1;
                        }
                    }
                    next_cursor = cursor;
                }
                cursor = (uint32_t)__shfl(next_cursor, 0, 64);
                any = any || __shfl(chain_any ? 1 : 0, 0, 64) != 0;
                lds_order();
                // resume at the next screened candidate of this window, if any
                const uint32_t rel = cursor - wbase;
                const uint64_t rest = rel < 64 ? (cand >> rel) : 0ull;
                if (rest != 0) {
                    cursor += (uint32_t)__builtin_ctzll(rest);
                    // re-screen from there: the window's other flags still hold,
                    // and a fresh screen of the same positions gives them again
                } else if (cursor < wbase + 64) {
                    cursor = wbase + 64;
                }
            }
            if (!any) {
                stop = true;
                break;
            }
        }
        if (stop) break;
    }
    __syncthreads();
    const float limit = half_skin * half_skin;
    bool flag = false;
    for (uint32_t i = lane; i < n; i += 64) {
        const uint32_t v = start + i;
        const float4 old = in[v], q = P[i];
        pos[v] = make_float4(q.x, q.y, q.z, old.w);
        if (track_walls) wall[v] = make_float4(WR[i].x, WR[i].y, WR[i].z, 0.f);
        const float4 r = refpos[v];
        const float dx = q.x - r.x, dy = q.y - r.y, dz = q.z - r.z;
        if (dx * dx + dy * dy + dz * dz > limit) flag = true;
    }
    if (flag) ss[span.z].rebuild = 1u;
    if (__ballot(flag) != 0 && lane == 0) st->nstate[0] = 1u;
}

// Long fibers that do not fit in LDS: one lane per fiber walks the chain in
// global memory with the same evaluate-then-apply projection.
__global__ __launch_bounds__(64) void k_curvature_gs_lane(DevState *st, SpecState *ss, const GridParams *grids,
                                                          const float4 *in, float4 *pos, const float4 *refpos,
                                                          const float4 *vA, const float4 *vB, const uint4 *fspan,
                                                          float4 *wall, uint32_t F, uint32_t sweeps, float stiffness,
                                                          float margin, uint32_t track_walls) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const uint4 span = fspan[f];
    if (!running(ss, span.z) || (__float_as_uint(vB[span.x].z) & kInactive)) return;
    const uint32_t start = span.x, n = span.y;
    const GridParams &g = grids[span.z];
    const float half_skin = 0.5f * g.skin;
    if (in != pos)
        for (uint32_t v = start; v < start + n; ++v) pos[v] = in[v];
    const uint32_t trip = n >= 3 ? n - 2 : 0;
    for (uint32_t sweep = 0; sweep < sweeps && trip > 0; ++sweep) {
        bool any = false;
        for (int dir = 0; dir < 2; ++dir) {
            for (uint32_t order = 0; order < trip; ++order) {
                const uint32_t t = start + (dir == 0 ? order : trip - 1 - order);
                const float4 A = pos[t], B = pos[t + 1], C = pos[t + 2];
                const float4 sa = vA[t], sb = vA[t + 1], sc = vA[t + 2];
                const uint32_t fa = __float_as_uint(vB[t].z), fb = __float_as_uint(vB[t + 1].z),
                               fc = __float_as_uint(vB[t + 2].z);
                float3 mp, mc, mn, rp, rc, rn;
                if (!triplet_projection(make_float3(A.x, A.y, A.z), make_float3(B.x, B.y, B.z),
                                        make_float3(C.x, C.y, C.z), sb.x, fa & kPinned, fb & kPinned, fc & kPinned,
                                        sa.w, sb.w, sc.w, g, stiffness, margin, mp, mc, mn, rp, rc, rn))
                    continue;
                any = true;
                pos[t] = make_float4(mp.x, mp.y, mp.z, A.w);
                pos[t + 1] = make_float4(mc.x, mc.y, mc.z, B.w);
                pos[t + 2] = make_float4(mn.x, mn.y, mn.z, C.w);
                if (track_walls) {
                    if (!(fa & kPinned)) { float4 w = wall[t]; wall[t] = make_float4(w.x + rp.x, w.y + rp.y, w.z + rp.z, 0.f); }
                    if (!(fb & kPinned)) { float4 w = wall[t + 1]; wall[t + 1] = make_float4(w.x + rc.x, w.y + rc.y, w.z + rc.z, 0.f); }
                    if (!(fc & kPinned)) { float4 w = wall[t + 2]; wall[t + 2] = make_float4(w.x + rn.x, w.y + rn.y, w.z + rn.z, 0.f); }
                }
            }
            if (dir == 0 && !any) break;
        }
        if (!any) break;
    }
    const float limit = half_skin * half_skin;
    bool flag = false;
    for (uint32_t v = start; v < start + n; ++v) {
        float4 p = pos[v], r = refpos[v];
        float dx = p.x - r.x, dy = p.y - r.y, dz = p.z - r.z;
        if (dx * dx + dy * dy + dz * dz > limit) flag = true;
    }
    if (flag) {
        ss[span.z].rebuild = 1u;
        st->nstate[0] = 1u;
    }
}

// Rebuild on request: request_neighbor_list_rebuild (neighbor.rs:434-441),
// the rebuild, and finish_neighbor_list_rebuild.
// One workgroup: every specimen's lists are marked stale (or, with `only`,
// those of the flagged specimens).
__global__ void k_request_rebuild(DevState *st, SpecState *ss, const uint32_t *only, uint32_t nspec) {
    bool any = false;
    for (uint32_t s = threadIdx.x; s < nspec; s += blockDim.x) {
        if (only && only[s] == 0u) continue;
        ss[s].rebuild = 1u;
        any = true;
    }
    if (__syncthreads_or(any) && threadIdx.x == 0) st->nstate[0] = 1u;
}

__global__ void k_finish_rebuild(DevState *st, SpecState *ss, uint32_t nspec) {
    const bool pending = st->nstate[0] != 0u;
    __syncthreads();
    if (!pending) return;
    for (uint32_t s = threadIdx.x; s < nspec; s += blockDim.x) ss[s].rebuild = 0u;
    if (threadIdx.x == 0) {
        st->nstate[0] = 0u;
        st->nstate[1] += 1u;
    }
}

// ---------------------------------------------------------------------------
// Formation (kernels.rs:164-221, 538-710): activation of fibers by recipe
// step, manufacturing-layer targets and vertex (needle) targets.

// activate_formation_step (kernels.rs:164-221) for non-adaptive layouts:
// fibers past `maximum_step` become inactive, the others active. The flag is
// kept per fiber and mirrored in the vertex flags the per-vertex kernels read.
__global__ void k_activate(const uint4 *fspan, const uint32_t *fstep, uint32_t *factive, float4 *vB, uint32_t F,
                           uint32_t maximum_step) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const uint32_t on = fstep[f] > maximum_step ? 0u : 1u;
    if (factive[f] == on) return;
    factive[f] = on;
    const uint4 span = fspan[f];
    for (uint32_t i = 0; i < span.y; ++i) {
        float4 b = vB[span.x + i];
        uint32_t flags = __float_as_uint(b.z);
        flags = on ? (flags & ~kInactive) : (flags | kInactive);
        b.z = __uint_as_float(flags);
        vB[span.x + i] = b;
    }
}

// The vertex flags of every fiber set from its activity flag (restore).
__global__ void k_apply_activity(const uint4 *fspan, const uint32_t *factive, float4 *vB, uint32_t F) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const uint32_t on = factive[f];
    const uint4 span = fspan[f];
    for (uint32_t i = 0; i < span.y; ++i) {
        float4 b = vB[span.x + i];
        uint32_t flags = __float_as_uint(b.z);
        flags = on ? (flags & ~kInactive) : (flags | kInactive);
        b.z = __uint_as_float(flags);
        vB[span.x + i] = b;
    }
}

__device__ __forceinline__ float axis_of(const float4 &p, uint32_t axis) {
    return axis == 0 ? p.x : (axis == 1 ? p.y : p.z);
}
__device__ __forceinline__ void add_axis(float4 &p, uint32_t axis, float d) {
    if (axis == 0) p.x += d;
    else if (axis == 1) p.y += d;
    else p.z += d;
}
__device__ __forceinline__ void set_axis(float4 &p, uint32_t axis, float v) {
    if (axis == 0) p.x = v;
    else if (axis == 1) p.y = v;
    else p.z = v;
}

// The mean coordinate of a fiber along `axis`, summed in vertex order.
__device__ __forceinline__ float fiber_center(const float4 *pos, uint4 span, uint32_t axis) {
    float center = 0.0f;
    for (uint32_t i = 0; i < span.y; ++i) center += axis_of(pos[span.x + i], axis);
    return center / (float)span.y;
}

// apply_formation_layer_targets (kernels.rs:538-580): one lane per fiber.
__global__ void k_layer_targets(const SpecState *ss, float4 *pos, const uint4 *fspan, const uint32_t *flayer,
                                const uint32_t *factive, const float *targets, uint32_t layers, uint32_t axis,
                                uint32_t first, uint32_t last, float stiffness, float max_translation, uint32_t F,
                                uint32_t gate) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const uint4 span = fspan[f];
    if (gate && !running(ss, span.z)) return;
    const uint32_t layer = flayer[f];
    if (layer == 0xffffffffu || layer >= layers || layer < first || layer > last) return;
    if (span.y == 0 || factive[f] == 0u) return;
    const float center = fiber_center(pos, span, axis);
    const float translation = rclamp((targets[layer] - center) * stiffness, -max_translation, max_translation);
    for (uint32_t i = 0; i < span.y; ++i) {
        float4 p = pos[span.x + i];
        add_axis(p, axis, translation);
        pos[span.x + i] = p;
    }
}

// initialize_vertex_displacement_targets (kernels.rs:582-603).
__global__ void k_init_vertex_targets(const float4 *pos, const uint32_t *index, float *target, float *command,
                                      uint32_t count, uint32_t axis, float displacement) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float c = axis_of(pos[index[i]], axis);
    target[i] = c + displacement;
    command[i] = c;
}

// apply_vertex_targets (kernels.rs:605-635).
__global__ void k_vertex_targets(const SpecState *ss, float4 *pos, const float4 *vB, const uint32_t *index,
                                 const float *target, float *command, uint32_t count, uint32_t axis, float stiffness,
                                 float max_translation) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t v = index[i];
    const uint32_t flags = __float_as_uint(vB[v].z);
    if (!running(ss, flags >> kSpecShift) || (flags & kInactive)) return;
    const float c = command[i];
    const float translation = rclamp((target[i] - c) * stiffness, -max_translation, max_translation);
    command[i] = c + translation;
    float4 p = pos[v];
    set_axis(p, axis, c + translation);
    pos[v] = p;
}

// measure_layer_target_error and measure_vertex_target_error
// (kernels.rs:637-710): maxima, order-free, over one workgroup.
__global__ __launch_bounds__(256) void k_target_error(const float4 *pos, const uint4 *fspan, const uint32_t *flayer,
                                                      const uint32_t *factive, const float *targets, uint32_t layers,
                                                      uint32_t layer_axis, uint32_t first, uint32_t last, uint32_t F,
                                                      const float4 *vB, const uint32_t *index, const float *vtarget,
                                                      uint32_t vcount, uint32_t vertex_axis, float *error) {
    __shared__ uint32_t scratch[5];
    float le = 0.f, ve = 0.f;
    if (layers) {
        for (uint32_t f = threadIdx.x; f < F; f += blockDim.x) {
            const uint32_t layer = flayer[f];
            const uint4 span = fspan[f];
            if (layer == 0xffffffffu || layer >= layers || layer < first || layer > last) continue;
            if (span.y == 0 || factive[f] == 0u) continue;
            le = fmaxf(le, fabsf(targets[layer] - fiber_center(pos, span, layer_axis)));
        }
    }
    for (uint32_t i = threadIdx.x; i < vcount; i += blockDim.x) {
        const uint32_t v = index[i];
        if (__float_as_uint(vB[v].z) & kInactive) continue;
        ve = fmaxf(ve, fabsf(vtarget[i] - axis_of(pos[v], vertex_axis)));
    }
    const uint32_t lb = block_max_256(__float_as_uint(le), scratch);
    const uint32_t vb = block_max_256(__float_as_uint(ve), scratch);
    if (threadIdx.x == 0) {
        error[0] = __uint_as_float(lb);
        error[1] = __uint_as_float(vb);
    }
}

// ---------------------------------------------------------------------------
// Compaction kinematics (compact_rigid_fiber_centers, compact_affine_vertices,
// compact_moving_walls, kernels.rs:712-809): the cell changes from `old` to
// `nw` and the active material follows it; wall reactions are reset (rigid,
// affine) or set to the projection onto the new walls (moving walls).

// rigid: each active fiber translates so that its vertex-mean centre keeps
// its fraction of the cell, its shape unchanged
__global__ void k_compact_rigid(float4 *pos, const uint4 *fspan, const float4 *vB, float4 *wall, uint32_t F,
                                CellChange box) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const uint4 span = fspan[f];
    if (span.y == 0 || (__float_as_uint(vB[span.x].z) & kInactive)) return;
    float cx = 0.f, cy = 0.f, cz = 0.f;
    for (uint32_t i = 0; i < span.y; ++i) {
        const float4 p = pos[span.x + i];
        cx += p.x;
        cy += p.y;
        cz += p.z;
    }
    const float n = (float)span.y;
    cx /= n;
    cy /= n;
    cz /= n;
    const float c[3] = {cx, cy, cz};
    float d[3];
    for (int k = 0; k < 3; ++k) {
        const float fraction = (c[k] - box.old_lower[k]) / (box.old_upper[k] - box.old_lower[k]);
        d[k] = box.new_lower[k] + fraction * (box.new_upper[k] - box.new_lower[k]) - c[k];
    }
    for (uint32_t i = 0; i < span.y; ++i) {
        float4 p = pos[span.x + i];
        p.x += d[0];
        p.y += d[1];
        p.z += d[2];
        pos[span.x + i] = p;
        if (wall) wall[span.x + i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// affine: every active vertex keeps its fractional coordinates
__global__ void k_compact_affine(float4 *pos, const float4 *vB, float4 *wall, uint32_t V, CellChange box) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V || (__float_as_uint(vB[v].z) & kInactive)) return;
    float4 p = pos[v];
    float q[3] = {p.x, p.y, p.z};
    for (int k = 0; k < 3; ++k) {
        const float fraction = (q[k] - box.old_lower[k]) / (box.old_upper[k] - box.old_lower[k]);
        q[k] = box.new_lower[k] + fraction * (box.new_upper[k] - box.new_lower[k]);
    }
    p.x = q[0];
    p.y = q[1];
    p.z = q[2];
    pos[v] = p;
    if (wall) wall[v] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// moving walls: active vertices are projected inside the new walls of the
// non-periodic axes; the reaction is the projection (negative: lower face)
__global__ void k_compact_walls(float4 *pos, const float4 *vA, const float4 *vB, float4 *wall, uint32_t V,
                                CellChange box) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V || (__float_as_uint(vB[v].z) & kInactive)) return;
    float4 p = pos[v];
    const float extent = vA[v].w;
    float q[3] = {p.x, p.y, p.z}, r[3] = {0.f, 0.f, 0.f};
    for (int k = 0; k < 3; ++k) {
        if (box.periodic[k]) continue;
        const float previous = q[k];
        q[k] = rclamp(previous, box.new_lower[k] + extent, box.new_upper[k] - extent);
        r[k] = previous - q[k];
    }
    p.x = q[0];
    p.y = q[1];
    p.z = q[2];
    pos[v] = p;
    if (wall) wall[v] = make_float4(r[0], r[1], r[2], 0.f);
}

// measure_compaction_metrics (kernels.rs:811-923), in two launches. The
// reference reduces on one thread in index order. Here k_compaction_terms
// computes every item's terms in parallel (zero where the reference skips
// the item, which leaves its nonnegative sums unchanged) into rows [5][S]
// (segments) and [7][V] (vertices); k_compaction_sums then keeps each sum's
// index order on a lane of its own, streaming the rows through LDS tiles
// that the whole workgroup loads one tile ahead, so every f32 sum rounds as
// the reference's does while the memory latency hides under the additions.
// out[13] = {pressure x, y, z, contact, stretch, bending and total energy,
// lower x, upper x, lower y, upper y, lower z, upper z face pressures}.
__global__ void k_compaction_terms(const float4 *pos, const uint4 *sstat, const uint32_t *factive,
                                   const float4 *vA, const float4 *corr0, const float4 *corr1,
                                   const float *ratios, const float4 *wall, uint32_t S, uint32_t V,
                                   float contact_stiffness, float stretch_stiffness, float bending_stiffness,
                                   float *terms) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < S) {
        float t[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
        const uint4 me = sstat[i];
        if (factive[me.y] != 0u) {
            const float4 c0 = corr0[i], c1 = corr1[i];
            t[0] = fabsf(c0.x) + fabsf(c1.x);
            t[1] = fabsf(c0.y) + fabsf(c1.y);
            t[2] = fabsf(c0.z) + fabsf(c1.z);
            const float penetration = c0.w;
            t[3] = 0.5f * contact_stiffness * penetration * penetration;
            const float4 a = pos[me.x], b = pos[me.x + 1];
            const float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
            const float length = sqrtf(dx * dx + dy * dy + dz * dz);
            const float rest = vA[me.x].z;
            if (rest > 1.0e-7f) {
                const float extension = length - rest;
                t[4] = 0.5f * stretch_stiffness * extension * extension / rest;
            }
        }
        for (int k = 0; k < 5; ++k) terms[(size_t)k * S + i] = t[k];
    }
    if (i < V) {
        float t[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        const float excess = fmaxf(ratios[i] - 1.0f, 0.f);
        t[0] = 0.5f * bending_stiffness * excess * excess;
        if (wall) {
            const float4 r = wall[i];
            t[1] = fmaxf(-r.x, 0.f);
            t[2] = fmaxf(r.x, 0.f);
            t[3] = fmaxf(-r.y, 0.f);
            t[4] = fmaxf(r.y, 0.f);
            t[5] = fmaxf(-r.z, 0.f);
            t[6] = fmaxf(r.z, 0.f);
        }
        float *vt = terms + (size_t)5 * S;
        for (int k = 0; k < 7; ++k) vt[(size_t)k * V + i] = t[k];
    }
}

constexpr uint32_t kSumThreads = 256, kSumPerThread = 4, kSumTile = kSumThreads * kSumPerThread;

// Adds rows[k][0..n) in index order on lane k (k < nrows), one LDS tile of
// kSumTile items at a time; the next tile's global loads are in flight while
// the current one is added.
__device__ void ordered_row_sums(const float *rows, uint32_t nrows, uint32_t n, float (*tile)[7][kSumTile + 4],
                                 float *sums) {
    const uint32_t tid = threadIdx.x;
    float acc = 0.f;
    float regs[7][kSumPerThread];
    auto load = [&](uint32_t base) {
        for (uint32_t k = 0; k < nrows; ++k)
            for (uint32_t j = 0; j < kSumPerThread; ++j) {
                const uint32_t idx = base + tid + j * kSumThreads;
                regs[k][j] = idx < n ? rows[(size_t)k * n + idx] : 0.f;
            }
    };
    auto store = [&](uint32_t buf) {
        for (uint32_t k = 0; k < nrows; ++k)
            for (uint32_t j = 0; j < kSumPerThread; ++j) tile[buf][k][tid + j * kSumThreads] = regs[k][j];
    };
    if (n > 0) {
        load(0);
        store(0);
    }
    __syncthreads();
    uint32_t buf = 0;
    for (uint32_t base = 0; base < n; base += kSumTile, buf ^= 1u) {
        const bool more = base + kSumTile < n;
        if (more) load(base + kSumTile);
        if (tid < nrows) {
            const float *row = tile[buf][tid];
            const uint32_t m = min(kSumTile, n - base);
            if (m == kSumTile) {
                for (uint32_t i = 0; i < kSumTile; i += 4u) {
                    const float4 q = *reinterpret_cast<const float4 *>(row + i);
                    acc += q.x;
                    acc += q.y;
                    acc += q.z;
                    acc += q.w;
                }
            } else {
                for (uint32_t i = 0; i < m; ++i) acc += row[i];
            }
        }
        if (more) store(buf ^ 1u);
        __syncthreads();
    }
    if (tid < nrows) sums[tid] = acc;
    __syncthreads();
}

__global__ __launch_bounds__(kSumThreads) void k_compaction_sums(const GridParams *grid, const float *terms,
                                                                 uint32_t S, uint32_t V,
                                                                 float correction_fraction,
                                                                 float contact_stiffness, float *out) {
    __shared__ __align__(16) float tile[2][7][kSumTile + 4];
    __shared__ float sums[12];
    ordered_row_sums(terms, 5u, S, tile, sums);
    ordered_row_sums(terms + (size_t)5 * S, 7u, V, tile, sums + 5);
    if (threadIdx.x != 0) return;
    const float load_x = sums[0], load_y = sums[1], load_z = sums[2], contact = sums[3], stretch = sums[4];
    const float bending = sums[5], lower_x = sums[6], upper_x = sums[7], lower_y = sums[8], upper_y = sums[9],
                lower_z = sums[10], upper_z = sums[11];
    const GridParams g = *grid;
    const float length_x = g.upper[0] - g.lower[0];
    const float length_y = g.upper[1] - g.lower[1];
    const float length_z = g.upper[2] - g.lower[2];
    const float inverse_projection = 1.0f / fmaxf(correction_fraction, 1.0e-7f);
    const float pressure_x = g.periodic[0] ? load_x : lower_x + upper_x;
    const float pressure_y = g.periodic[1] ? load_y : lower_y + upper_y;
    const float pressure_z = g.periodic[2] ? load_z : lower_z + upper_z;
    const float area_x = length_y * length_z, area_y = length_x * length_z, area_z = length_x * length_y;
    out[0] = contact_stiffness * pressure_x * inverse_projection / area_x;
    out[1] = contact_stiffness * pressure_y * inverse_projection / area_y;
    out[2] = contact_stiffness * pressure_z * inverse_projection / area_z;
    out[3] = contact;
    out[4] = stretch;
    out[5] = bending;
    out[6] = contact + stretch + bending;
    out[7] = contact_stiffness * lower_x * inverse_projection / area_x;
    out[8] = contact_stiffness * upper_x * inverse_projection / area_x;
    out[9] = contact_stiffness * lower_y * inverse_projection / area_y;
    out[10] = contact_stiffness * upper_y * inverse_projection / area_y;
    out[11] = contact_stiffness * lower_z * inverse_projection / area_z;
    out[12] = contact_stiffness * upper_z * inverse_projection / area_z;
}

// ---------------------------------------------------------------------------
// Contact capture (capture_segment_contacts, contact.rs:9-230): every active
// segment scans the 27 cells around its midpoint on a whole-segment grid and
// appends each later segment of another fiber whose surface gap is at most
// `max_gap`, with the closest-point coordinates and the crossing angle. The
// append order is unordered; the host sorts the pairs.
// kCaptureLanes lanes share one segment and split each cell's entries; the
// arithmetic of every pair is the one-lane arithmetic, only the append order
// differs, and the host sorts the pairs.
constexpr uint32_t kCaptureLanes = 8;
__global__ void k_capture(const GridParams *grid, const float4 *pos, const uint4 *sstat, const uint32_t *factive,
                          const uint32_t *cell_counts, const uint32_t *cell_offsets, const uint32_t *cell_segs,
                          uint32_t S, float max_gap, uint32_t capacity, uint32_t *counter, uint2 *pairs,
                          float2 *coordinates, float *gaps, float *angles) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t seg = gid / kCaptureLanes, sub = gid % kCaptureLanes;
    if (seg >= S) return;
    const uint4 me = sstat[seg];
    if (factive[me.y] == 0u) return;
    const GridParams g = *grid;
    const float4 P1 = pos[me.x], Q1 = pos[me.x + 1];
    const float p1x = P1.x, p1y = P1.y, p1z = P1.z, q1x = Q1.x, q1y = Q1.y, q1z = Q1.z;
    const float d1x = q1x - p1x, d1y = q1y - p1y, d1z = q1z - p1z;
    const float radius = __uint_as_float(me.z);
    const float eps_len = 1.0e-20f, eps_par = 1.0e-6f;
    uint32_t hx, hy, hz;
    cell_coords(g, p1x, p1y, p1z, q1x, q1y, q1z, hx, hy, hz);
    const uint32_t home = (hz * g.cells[1] + hy) * g.cells[0] + hx;
    for (uint32_t n = 0; n < 27; ++n) {
        uint32_t cell;
        if (!stencil_cell(g, home, n, cell)) continue;
        // the unwrapped neighbour's box shift, then the minimum image
        const int32_t o[3] = {(int32_t)(n % 3) - 1, (int32_t)((n / 3) % 3) - 1, (int32_t)(n / 9) - 1};
        const uint32_t h[3] = {hx, hy, hz};
        float shift[3] = {0.f, 0.f, 0.f};
        for (int k = 0; k < 3; ++k) {
            const int32_t raw = (int32_t)h[k] + o[k];
            if (raw < 0) shift[k] = -(g.upper[k] - g.lower[k]);
            else if (raw >= (int32_t)g.cells[k]) shift[k] = g.upper[k] - g.lower[k];
        }
        const uint32_t start = cell_offsets[cell], count = cell_counts[cell];
        for (uint32_t j = sub; j < count; j += kCaptureLanes) {
            const uint32_t other = cell_segs[start + j];
            if (other <= seg) continue;
            const uint4 ot = sstat[other];
            if (ot.y == me.y) continue;
            const float4 P2 = pos[ot.x], Q2 = pos[ot.x + 1];
            float p2x = P2.x + shift[0], p2y = P2.y + shift[1], p2z = P2.z + shift[2];
            float q2x = Q2.x + shift[0], q2y = Q2.y + shift[1], q2z = Q2.z + shift[2];
            const float lx = g.upper[0] - g.lower[0], ly = g.upper[1] - g.lower[1], lz = g.upper[2] - g.lower[2];
            const float idx = 0.5f * (p2x + q2x - p1x - q1x);
            const float idy = 0.5f * (p2y + q2y - p1y - q1y);
            const float idz = 0.5f * (p2z + q2z - p1z - q1z);
            if (g.periodic[0]) {
                const float sh = roundf(idx / lx) * lx;
                p2x -= sh;
                q2x -= sh;
            }
            if (g.periodic[1]) {
                const float sh = roundf(idy / ly) * ly;
                p2y -= sh;
                q2y -= sh;
            }
            if (g.periodic[2]) {
                const float sh = roundf(idz / lz) * lz;
                p2z -= sh;
                q2z -= sh;
            }
            const float d2x = q2x - p2x, d2y = q2y - p2y, d2z = q2z - p2z;
            const float rx = p1x - p2x, ry = p1y - p2y, rz = p1z - p2z;
            const float a = d1x * d1x + d1y * d1y + d1z * d1z;
            const float e = d2x * d2x + d2y * d2y + d2z * d2z;
            const float f = d2x * rx + d2y * ry + d2z * rz;
            float sp = 0.f, tp = 0.f;
            if (a <= eps_len && e > eps_len) {
                tp = rclamp(f / e, 0.f, 1.f);
            } else if (a > eps_len) {
                const float c = d1x * rx + d1y * ry + d1z * rz;
                if (e <= eps_len) {
                    sp = rclamp(-c / a, 0.f, 1.f);
                } else {
                    const float b = d1x * d2x + d1y * d2y + d1z * d2z;
                    const float den = a * e - b * b;
                    if (fabsf(den) > eps_par * a * e) sp = rclamp((b * f - c * e) / den, 0.f, 1.f);
                    const float proj = (b * sp + f) / e;
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
            const float dx = (p2x + d2x * tp) - (p1x + d1x * sp);
            const float dy = (p2y + d2y * tp) - (p1y + d1y * sp);
            const float dz = (p2z + d2z * tp) - (p1z + d1z * sp);
            const float center = sqrtf(dx * dx + dy * dy + dz * dz);
            const float gap = center - radius - __uint_as_float(ot.z);
            if (!(gap <= max_gap)) continue;
            const uint32_t k = atomicAdd(&counter[0], 1u);
            if (k < capacity) {
                pairs[k] = make_uint2(seg, other);
                coordinates[k] = make_float2(sp, tp);
                gaps[k] = gap;
                const float product = sqrtf(a * e);
                float angle = 0.f;
                if (product > eps_len)
                    angle = acosf_f32(rclamp(fabsf((d1x * d2x + d1y * d2y + d1z * d2z) / product), 0.f, 1.f));
                angles[k] = angle;
            } else {
                atomicOr(&counter[1], 1u);
            }
        }
    }
}

// Sort keys of the captured pairs: (first segment, second segment) in one
// 64-bit key, with the append slot as the value.
__global__ void k_capture_keys(const uint2 *pairs, uint32_t n, uint64_t *keys, uint32_t *slots) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = ((uint64_t)pairs[i].x << 32) | pairs[i].y;
    slots[i] = i;
}

// One captured candidate as the C ABI returns it (tgh_contact_candidate).
struct CandidateRecord {
    uint2 pair;
    float2 coordinates;
    float gap, angle;
};
static_assert(sizeof(CandidateRecord) == 24, "candidate record layout");

// The captured arrays in sorted order, as candidate records.
__global__ void k_capture_gather(const uint32_t *slots, uint32_t n, const uint2 *pairs, const float2 *coordinates,
                                 const float *gaps, const float *angles, CandidateRecord *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = slots[i];
    CandidateRecord r;
    r.pair = pairs[k];
    r.coordinates = coordinates[k];
    r.gap = gaps[k];
    r.angle = angles[k];
    out[i] = r;
}

}  // namespace tgh

// ---------------------------------------------------------------------------
// launchers

namespace tgh {

namespace {
inline uint32_t blocks_for(uint64_t n, uint32_t threads) { return (uint32_t)((n + threads - 1) / threads); }
// grid of a rebuild kernel over n elements, `per_block` a chunk
inline uint32_t rebuild_blocks(uint64_t n, uint32_t per_block) {
    const uint32_t b = blocks_for(n, per_block);
    return b < kRebuildBlocks ? b : kRebuildBlocks;
}
}  // namespace

size_t curvature_lds_bytes(uint32_t max_vertices, uint32_t lanes, uint32_t track_walls) {
    if (lanes == 8u)  // the sequential walk's layout
        return 8u * curvature_seq_stride_floats(max_vertices, track_walls) * sizeof(float);
    return (64u / lanes) * curvature_group_stride_floats(max_vertices, track_walls) * sizeof(float);
}

hipError_t launch_begin_batch(const WorldDev &w, hipStream_t stream) {
    hipLaunchKernelGGL(k_begin_batch, dim3(blocks_for(w.nspec, 256)), dim3(256), 0, stream, w.ss, w.nspec);
    return hipGetLastError();
}

hipError_t launch_rebuild(const WorldDev &w, hipStream_t stream) {
    if (w.S == 0 || w.C == 0 || w.V == 0) return hipErrorInvalidValue;
    const uint32_t clear_n = w.C > w.V ? w.C : w.V;
    hipLaunchKernelGGL(k_rebuild_clear, dim3(rebuild_blocks(clear_n, 256)), dim3(256), 0, stream, w.st, w.ss, w.cell_spec,
                       w.cell_counts, w.cell_cursor, w.C, w.pos, w.refpos, w.vB, w.V, w.cell_reach);
    hipLaunchKernelGGL(k_bin_segments, dim3(rebuild_blocks(w.S, 256)), dim3(256), 0, stream, w.st, w.ss, w.grids, w.pos,
                       w.sstat, w.factive, w.seg_cell, w.cell_counts, w.cell_reach, w.S);
    hipLaunchKernelGGL(k_scan_cells, dim3(w.nspec), dim3(1024), 0, stream, w.st, w.ss, w.grids, w.cell_counts,
                       w.cell_offsets, w.cell_segs, w.scattered, w.slotC);
    hipLaunchKernelGGL(k_scatter_segments, dim3(rebuild_blocks(w.S, 256)), dim3(256), 0, stream, w.st, w.ss, w.sstat,
                       w.cell_spec, w.seg_cell, w.cell_offsets, w.cell_cursor, w.scattered, w.S);
    hipLaunchKernelGGL(k_rank_slots, dim3(rebuild_blocks(w.S, 256)), dim3(256), 0, stream, w.st, w.ss, w.sstat,
                       w.seg_cell, w.cell_counts, w.cell_offsets, w.scattered, w.cell_segs, w.pos, w.slotA, w.slotB,
                       w.slotC, w.cP, w.cQ, w.S);
    {
        const uint32_t blocks = blocks_for(w.S, kBuildGroups);
        const dim3 grid(blocks < kBuildBlocks ? blocks : kBuildBlocks);
#define TGH_BUILD_ARGS                                                                                        \
    w.st, w.ss, w.grids, w.sstat, w.fspan, w.slotA, w.slotB, w.slotC, w.cP, w.cQ, w.seg_cell, w.cell_counts, \
        w.cell_offsets, w.ncount, w.nbr, w.nbr_off, w.ovf, w.cell_reach, w.block, w.ovf_cap, w.S
        if (w.prune_lists)
            hipLaunchKernelGGL(k_build_lists<true>, grid, dim3(64), 0, stream, TGH_BUILD_ARGS);
        else
            hipLaunchKernelGGL(k_build_lists<false>, grid, dim3(64), 0, stream, TGH_BUILD_ARGS);
#undef TGH_BUILD_ARGS
    }
    return hipGetLastError();
}

hipError_t launch_rebuild_now(const WorldDev &w, hipStream_t stream) {
    hipLaunchKernelGGL(k_request_rebuild, dim3(1), dim3(256), 0, stream, w.st, w.ss, nullptr, w.nspec);
    hipError_t e = launch_rebuild(w, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_finish_rebuild, dim3(1), dim3(256), 0, stream, w.st, w.ss, w.nspec);
    return hipGetLastError();
}

hipError_t launch_refresh_lists(const WorldDev &w, hipStream_t stream) {
    hipError_t e = launch_rebuild(w, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_finish_rebuild, dim3(1), dim3(256), 0, stream, w.st, w.ss, w.nspec);
    return hipGetLastError();
}

hipError_t launch_refresh_slots(const WorldDev &w, hipStream_t stream) {
    hipLaunchKernelGGL(k_refresh_slots, dim3(blocks_for(w.S, 256)), dim3(256), 0, stream, w.pos, w.slotC, w.cP, w.cQ,
                       w.S);
    return hipGetLastError();
}

hipError_t launch_request_rebuild(const WorldDev &w, hipStream_t stream) {
    hipLaunchKernelGGL(k_request_rebuild, dim3(1), dim3(256), 0, stream, w.st, w.ss, nullptr, w.nspec);
    return hipGetLastError();
}

hipError_t launch_request_rebuild_specimens(const WorldDev &w, const uint32_t *flags, hipStream_t stream) {
    if (!flags) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_request_rebuild, dim3(1), dim3(256), 0, stream, w.st, w.ss, flags, w.nspec);
    return hipGetLastError();
}

hipError_t launch_contact(const WorldDev &w, const IterParams &p, hipStream_t stream) {
    hipError_t e = launch_refresh_slots(w, stream);
    if (e != hipSuccess) return e;
#define TGH_CONTACT_ARGS                                                                                     \
    w.st, w.ss, w.nspec, w.grids, w.red_slots, w.cP, w.cQ, w.slotA, w.slotB, w.slotC, w.seg_cell, w.cell_counts, w.cell_offsets, \
        w.ncount, w.nbr, w.nbr_off, w.ovf, w.corr0, w.corr1, w.red, p.correction_fraction, p.aggregation, w.block,  \
        w.ovf_cap, w.S
    const dim3 grid(blocks_for(w.S, 256u / w.contact_lanes));
    switch (w.contact_lanes) {
        case 4: hipLaunchKernelGGL(k_contact<4>, grid, dim3(256), 0, stream, TGH_CONTACT_ARGS); break;
        case 8: hipLaunchKernelGGL(k_contact<8>, grid, dim3(256), 0, stream, TGH_CONTACT_ARGS); break;
        case 16: hipLaunchKernelGGL(k_contact<16>, grid, dim3(256), 0, stream, TGH_CONTACT_ARGS); break;
        case 32: hipLaunchKernelGGL(k_contact<32>, grid, dim3(256), 0, stream, TGH_CONTACT_ARGS); break;
        default: return hipErrorInvalidValue;
    }
#undef TGH_CONTACT_ARGS
    return hipGetLastError();
}

hipError_t launch_measure(const WorldDev &w, const IterParams &p, hipStream_t stream) {
    hipLaunchKernelGGL(k_measure, dim3(blocks_for(w.V, 256)), dim3(256), 0, stream, w.ss, w.red_slots, w.pos, w.vA,
                       w.vB, w.red, w.ratio, w.V);
    hipLaunchKernelGGL(k_assess, dim3(1), dim3(256), 0, stream, w.ss, w.nspec, w.red_slots, w.red,
                       p.penetration_tolerance, p.curvature_ratio_tolerance, p.iterations, p.force_full);
    return hipGetLastError();
}

// Jacobi passes alternate between the two position buffers; the last one
// also applies contact. The curvature kernel reads the buffer the last pass
// wrote and writes back to `pos`, so every iteration ends in `pos`.
static uint32_t vertex_passes(const IterParams &p) { return p.constraint_iterations > 0 ? p.constraint_iterations : 1u; }
// two or more passes run as one launch (k_vertex_passes), pos -> alt
static bool fused_vertex_passes(const IterParams &p) {
    return p.constraint_iterations >= 2u && p.constraint_iterations <= kMaxFusedPasses;
}

hipError_t launch_vertex_passes(const WorldDev &w, const IterParams &p, hipStream_t stream) {
    const uint32_t passes = vertex_passes(p);
    if (fused_vertex_passes(p)) {
        hipLaunchKernelGGL(k_vertex_passes, dim3(blocks_for(w.V, kPassOwned)), dim3(256),
                           vertex_passes_lds_bytes(passes), stream, w.ss, w.grids, w.red_slots, w.pos, w.alt, w.vA,
                           w.vB, w.corr0, w.corr1, w.wall, w.red, w.V, p.stretch, p.bend, p.max_step, passes,
                           w.track_walls);
        return hipGetLastError();
    }
    float4 *src = w.pos, *dst = w.alt;
    for (uint32_t i = 0; i < passes; ++i) {
        hipLaunchKernelGGL(k_vertex_pass, dim3(blocks_for(w.V, 256)), dim3(256), 0, stream, w.ss, w.grids,
                           w.red_slots, src, dst, w.vA, w.vB, w.corr0, w.corr1, w.wall, w.red, w.V, p.stretch, p.bend,
                           p.max_step, p.constraint_iterations > 0 ? 1u : 0u, i + 1 == passes ? 1u : 0u,
                           i == 0 ? 1u : 0u, w.track_walls);
        float4 *t = src;
        src = dst;
        dst = t;
    }
    return hipGetLastError();
}

hipError_t launch_curvature(const WorldDev &w, const IterParams &p, hipStream_t stream) {
    const float4 *src = (fused_vertex_passes(p) || vertex_passes(p) % 2 == 1) ? w.alt : w.pos;
    if (w.gs_lanes != 0 && w.max_fiber_vertices <= kMaxLdsFiberVertices) {
        const size_t lds = curvature_lds_bytes(w.max_fiber_vertices, w.gs_lanes, w.track_walls);
        const uint32_t per_block = 64u / w.gs_lanes;
        const dim3 grid(blocks_for(w.F, per_block));
#define TGH_GS_ARGS                                                                                              \
    w.st, w.ss, w.grids, src, w.pos, w.refpos, w.vA, w.vB, w.fspan, w.wall, w.F, p.sweeps, p.curvature_stiffness, \
        p.curvature_margin, w.track_walls, w.max_fiber_vertices
        switch (w.gs_lanes) {
            case 4: hipLaunchKernelGGL(k_curvature_gs_group<4>, grid, dim3(64), lds, stream, TGH_GS_ARGS); break;
            case 8:
                if (w.track_walls)
                    hipLaunchKernelGGL(k_curvature_gs_seq<true>, grid, dim3(64), lds, stream, TGH_GS_ARGS);
                else
                    hipLaunchKernelGGL(k_curvature_gs_seq<false>, grid, dim3(64), lds, stream, TGH_GS_ARGS);
                break;
            case 16: hipLaunchKernelGGL(k_curvature_gs_lean<16>, grid, dim3(64), lds, stream, TGH_GS_ARGS); break;
            case 32: hipLaunchKernelGGL(k_curvature_gs_lean<32>, grid, dim3(64), lds, stream, TGH_GS_ARGS); break;
            case 64: hipLaunchKernelGGL(k_curvature_gs_wave64, grid, dim3(64), lds, stream, TGH_GS_ARGS); break;
            default: return hipErrorInvalidValue;
        }
#undef TGH_GS_ARGS
    } else {
        hipLaunchKernelGGL(k_curvature_gs_lane, dim3(blocks_for(w.F, 64)), dim3(64), 0, stream, w.st, w.ss, w.grids,
                           src, w.pos, w.refpos, w.vA, w.vB, w.fspan, w.wall, w.F, p.sweeps, p.curvature_stiffness,
                           p.curvature_margin, w.track_walls);
    }
    return hipGetLastError();
}

hipError_t launch_activate(const WorldDev &w, uint32_t maximum_step, hipStream_t stream) {
    hipLaunchKernelGGL(k_activate, dim3(blocks_for(w.F, 64)), dim3(64), 0, stream, w.fspan, w.fstep, w.factive, w.vB,
                       w.F, maximum_step);
    return hipGetLastError();
}

hipError_t launch_layer_targets(const WorldDev &w, const FormationDev &f, bool gate_on_running, hipStream_t stream) {
    if (f.layer_count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_layer_targets, dim3(blocks_for(w.F, 64)), dim3(64), 0, stream, w.ss, w.pos, w.fspan,
                       w.flayer, w.factive, f.layer_targets, f.layer_count, f.layer_axis, f.first_layer,
                       f.last_layer, f.layer_stiffness, f.layer_max_translation, w.F, gate_on_running ? 1u : 0u);
    return hipGetLastError();
}

hipError_t launch_compact(const WorldDev &w, const CellChange &box, uint32_t kinematics, hipStream_t stream) {
    float4 *wall = w.track_walls ? w.wall : nullptr;
    if (kinematics == 0u) {
        hipLaunchKernelGGL(k_compact_rigid, dim3(blocks_for(w.F, 64)), dim3(64), 0, stream, w.pos, w.fspan, w.vB,
                           wall, w.F, box);
    } else if (kinematics == 1u) {
        hipLaunchKernelGGL(k_compact_walls, dim3(blocks_for(w.V, 256)), dim3(256), 0, stream, w.pos, w.vA, w.vB,
                           wall, w.V, box);
    } else {
        hipLaunchKernelGGL(k_compact_affine, dim3(blocks_for(w.V, 256)), dim3(256), 0, stream, w.pos, w.vB, wall,
                           w.V, box);
    }
    return hipGetLastError();
}

hipError_t launch_compaction_metrics(const WorldDev &w, const GridParams *grid, float correction_fraction,
                                     float contact_stiffness, float stretch_stiffness, float bending_stiffness,
                                     float *terms, float *out, hipStream_t stream) {
    const uint32_t n = w.S > w.V ? w.S : w.V;
    if (n > 0)
        hipLaunchKernelGGL(k_compaction_terms, dim3(blocks_for(n, 256)), dim3(256), 0, stream, w.pos, w.sstat,
                           w.factive, w.vA, w.corr0, w.corr1, w.ratio, w.track_walls ? w.wall : nullptr, w.S, w.V,
                           contact_stiffness, stretch_stiffness, bending_stiffness, terms);
    hipLaunchKernelGGL(k_compaction_sums, dim3(1), dim3(kSumThreads), 0, stream, grid, terms, w.S, w.V,
                       correction_fraction, contact_stiffness, out);
    return hipGetLastError();
}

hipError_t launch_capture(const WorldDev &w, const GridParams *capture_grid, uint32_t capture_cells, float max_gap,
                          uint32_t capacity, uint32_t *counter, uint2 *pairs, float2 *coordinates, float *gaps,
                          float *angles, hipStream_t stream) {
    if (w.S == 0 || capture_cells == 0 || w.nspec != 1) return hipErrorInvalidValue;
    // the whole-segment cell list on the capture grid, in the shared buffers
    WorldDev c = w;
    c.grids = const_cast<GridParams *>(capture_grid);
    c.C = capture_cells;
    hipLaunchKernelGGL(k_request_rebuild, dim3(1), dim3(256), 0, stream, c.st, c.ss, nullptr, c.nspec);
    const uint32_t clear_n = c.C > c.V ? c.C : c.V;
    hipLaunchKernelGGL(k_rebuild_clear, dim3(rebuild_blocks(clear_n, 256)), dim3(256), 0, stream, c.st, c.ss, c.cell_spec,
                       c.cell_counts, c.cell_cursor, c.C, c.pos, c.refpos, c.vB, c.V, c.cell_reach);
    hipLaunchKernelGGL(k_bin_segments, dim3(rebuild_blocks(c.S, 256)), dim3(256), 0, stream, c.st, c.ss, c.grids, c.pos,
                       c.sstat, c.factive, c.seg_cell, c.cell_counts, c.cell_reach, c.S);
    hipLaunchKernelGGL(k_scan_cells, dim3(1), dim3(1024), 0, stream, c.st, c.ss, c.grids, c.cell_counts,
                       c.cell_offsets, c.cell_segs, c.scattered, c.slotC);
    hipLaunchKernelGGL(k_scatter_segments, dim3(rebuild_blocks(c.S, 256)), dim3(256), 0, stream, c.st, c.ss, c.sstat,
                       c.cell_spec, c.seg_cell, c.cell_offsets, c.cell_cursor, c.scattered, c.S);
    hipLaunchKernelGGL(k_rank_slots, dim3(rebuild_blocks(c.S, 256)), dim3(256), 0, stream, c.st, c.ss, c.sstat,
                       c.seg_cell, c.cell_counts, c.cell_offsets, c.scattered, c.cell_segs, c.pos, c.slotA, c.slotB,
                       c.slotC, c.cP, c.cQ, c.S);
    hipLaunchKernelGGL(k_capture, dim3(blocks_for((size_t)c.S * kCaptureLanes, 256)), dim3(256), 0, stream,
                       capture_grid, c.pos, c.sstat,
                       c.factive, c.cell_counts, c.cell_offsets, c.cell_segs, c.S, max_gap, capacity, counter, pairs,
                       coordinates, gaps, angles);
    // the rebuild request stays pending: the next batch rebuilds the lists
    // on the relaxation grid
    return hipGetLastError();
}

hipError_t launch_capture_sort(uint32_t n, const uint2 *pairs, const float2 *coordinates, const float *gaps,
                               const float *angles, void *records, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    uint64_t *keys = nullptr, *keys_sorted = nullptr;
    uint32_t *slots = nullptr, *slots_sorted = nullptr;
    void *temp = nullptr;
    size_t temp_bytes = 0;
    hipError_t e = hipMalloc((void **)&keys, 2 * (size_t)n * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc((void **)&slots, 2 * (size_t)n * sizeof(uint32_t));
    if (e == hipSuccess) {
        keys_sorted = keys + n;
        slots_sorted = slots + n;
        hipLaunchKernelGGL(k_capture_keys, dim3(blocks_for(n, 256)), dim3(256), 0, stream, pairs, n, keys, slots);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, keys, keys_sorted, slots, slots_sorted, (int)n, 0,
                                               64, stream);
    if (e == hipSuccess) e = hipMalloc(&temp, temp_bytes > 0 ? temp_bytes : 1);
    if (e == hipSuccess)
        e = hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys, keys_sorted, slots, slots_sorted, (int)n, 0,
                                               64, stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_capture_gather, dim3(blocks_for(n, 256)), dim3(256), 0, stream, slots_sorted, n, pairs,
                           coordinates, gaps, angles, static_cast<CandidateRecord *>(records));
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    (void)hipFree(temp);
    (void)hipFree(slots);
    (void)hipFree(keys);
    return e;
}

hipError_t launch_apply_activity(const WorldDev &w, hipStream_t stream) {
    hipLaunchKernelGGL(k_apply_activity, dim3(blocks_for(w.F, 64)), dim3(64), 0, stream, w.fspan, w.factive, w.vB, w.F);
    return hipGetLastError();
}

hipError_t launch_init_vertex_targets(const WorldDev &w, const FormationDev &f, float displacement,
                                      hipStream_t stream) {
    if (f.vertex_count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_init_vertex_targets, dim3(blocks_for(f.vertex_count, 64)), dim3(64), 0, stream, w.pos,
                       f.vertex_index, f.vertex_target, f.vertex_command, f.vertex_count, f.vertex_axis,
                       displacement);
    return hipGetLastError();
}

hipError_t launch_vertex_targets(const WorldDev &w, const FormationDev &f, hipStream_t stream) {
    if (f.vertex_count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_vertex_targets, dim3(blocks_for(f.vertex_count, 64)), dim3(64), 0, stream, w.ss, w.pos,
                       w.vB, f.vertex_index, f.vertex_target, f.vertex_command, f.vertex_count, f.vertex_axis,
                       f.vertex_stiffness, f.vertex_max_translation);
    return hipGetLastError();
}

hipError_t launch_target_error(const WorldDev &w, const FormationDev &f, hipStream_t stream) {
    hipLaunchKernelGGL(k_target_error, dim3(1), dim3(256), 0, stream, w.pos, w.fspan, w.flayer, w.factive,
                       f.layer_targets, f.layer_count, f.layer_axis, f.first_layer, f.last_layer, w.F, w.vB,
                       f.vertex_index, f.vertex_target, f.vertex_count, f.vertex_axis, f.error);
    return hipGetLastError();
}

}  // namespace tgh
