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
//     64 lanes of a wave hold spatially adjacent segments and share their
//     candidate gathers in L1/L2; neighbour lists are stored column-major per
//     slot (nbr[k * S + slot]), so list reads are coalesced across lanes.
//   * Each segment's list is kept sorted while it is built, in LDS, which
//     replaces the separate sort pass (neighbor.rs:366-408).
//   * The Jacobi stretch/bend passes and the contact application read one
//     position buffer and write the other (kernels.rs:1027-1163, 1417-1537):
//     a find/apply pair becomes one launch with identical per-vertex
//     arithmetic, and the last pass also applies the contact corrections.
//   * The curvature Gauss–Seidel keeps a sliding three-vertex window of each
//     fiber in registers (kernels.rs:1165-1415) and stops after a forward
//     sweep that projected nothing, which leaves the result unchanged: the
//     remaining sweeps would revisit the same unchanged triplets.
//   * Max penetration / curvature ratio / vertex step reduce with wave64
//     shuffles and one u32 atomicMax per wave (order-free, so exact), and the
//     convergence assessment runs in the last workgroup of the curvature
//     kernel instead of a separate single-thread launch.
//
// All float arithmetic must be compiled without contraction
// (-ffp-contract=off) and with IEEE division and square root, as the Rust
// reference; cell lists and neighbour pair sets are then bit-exact.

#include <hip/hip_runtime.h>
#include <stdint.h>

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

__global__ void k_begin_batch(DevState *st) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    st->control[2] = 0;
    if (st->control[3] == 0) st->control[0] = 1;
}

// ---------------------------------------------------------------------------
// neighbour-list rebuild, every kernel gated on the device-side request flag
// (world/neighbor_list.rs:19-62)

__global__ void k_rebuild_clear(const DevState *st, uint32_t *cell_counts, uint32_t *cell_cursor, uint32_t C,
                                const float4 *pos, float4 *refpos, uint32_t V) {
    if (st->nstate[0] == 0) return;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < C) {
        cell_counts[i] = 0;
        cell_cursor[i] = 0;
    }
    // The list snapshot (neighbor.rs:410-422) is taken from the positions the
    // lists are built from; nothing moves them before the build finishes.
    if (i < V) refpos[i] = pos[i];
}

__global__ void k_bin_segments(const DevState *st, GridParams g, const float4 *pos, const uint4 *sstat,
                               uint32_t *seg_cell, uint32_t *cell_counts, uint32_t S) {
    if (st->nstate[0] == 0) return;
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    uint32_t v0 = sstat[s].x;
    uint32_t cell = cell_index(g, pos[v0], pos[v0 + 1]);
    seg_cell[s] = cell;
    atomicAdd(&cell_counts[cell], 1u);
}

// Exclusive scan of the cell counts in one 1024-thread workgroup: each thread
// sums a contiguous chunk, the 1024 chunk sums are scanned in LDS, then each
// thread writes its chunk's offsets.
__global__ __launch_bounds__(1024) void k_scan_cells(const DevState *st, const uint32_t *counts,
                                                     uint32_t *offsets, uint32_t C) {
    if (st->nstate[0] == 0) return;
    __shared__ uint32_t sums[1024];
    uint32_t t = threadIdx.x;
    uint32_t chunk = (C + 1023) / 1024;
    uint32_t begin = t * chunk;
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
    uint32_t run = sums[t] - local;
    for (uint32_t i = begin; i < end; ++i) {
        offsets[i] = run;
        run += counts[i];
    }
}

__global__ void k_scatter_segments(const DevState *st, const uint32_t *seg_cell, const uint32_t *cell_offsets,
                                   uint32_t *cell_cursor, uint32_t *scattered, uint32_t S) {
    if (st->nstate[0] == 0) return;
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    uint32_t cell = seg_cell[s];
    uint32_t slot = cell_offsets[cell] + atomicAdd(&cell_cursor[cell], 1u);
    scattered[slot] = s;
}

// One wave per cell: every lane ranks its scattered entries against the whole
// cell, which puts the cell in ascending segment order (the order
// rank_cell_segments produces, cell_list.rs:344-378), and copies the slot
// geometry the list build reads (gather_cell_slot_geometry, neighbor.rs:99-153).
__global__ __launch_bounds__(256) void k_rank_cells(const DevState *st, const uint32_t *cell_counts,
                                                    const uint32_t *cell_offsets, const uint32_t *scattered,
                                                    uint32_t *cell_segs, const float4 *pos, const uint4 *sstat,
                                                    float4 *slotA, float4 *slotB, uint4 *slotC, uint32_t C) {
    if (st->nstate[0] == 0) return;
    uint32_t cell = blockIdx.x * 4 + (threadIdx.x >> 6);
    uint32_t lane = threadIdx.x & 63;
    if (cell >= C) return;
    uint32_t n = cell_counts[cell], off = cell_offsets[cell];
    for (uint32_t i = lane; i < n; i += 64) {
        uint32_t mine = scattered[off + i];
        uint32_t rank = 0;
        for (uint32_t k = 0; k < n; ++k) rank += scattered[off + k] < mine ? 1u : 0u;
        uint32_t slot = off + rank;
        cell_segs[slot] = mine;
        uint4 ss = sstat[mine];
        float4 p = pos[ss.x], q = pos[ss.x + 1];
        slotA[slot] = make_float4(p.x, p.y, p.z, __uint_as_float(ss.z));
        slotB[slot] = make_float4(q.x, q.y, q.z, p.w);
        slotC[slot] = make_uint4(mine, ss.y, __float_as_uint(q.w), 0u);
    }
}

// One lane per cell-list slot builds that slot's segment list from the 27
// stencil cells (build_segment_neighbor_lists, neighbor.rs:155-364), keeping
// it sorted by insertion. `Lds` keeps the lane's list in LDS (block size up
// to kLdsListCapacity); otherwise it is sorted in place in global memory.
template <bool Lds>
__global__ __launch_bounds__(kBuildThreads) void k_build_lists(const DevState *st, GridParams g,
                                                               const float4 *slotA, const float4 *slotB,
                                                               const uint4 *slotC, const uint32_t *seg_cell,
                                                               const uint32_t *cell_counts,
                                                               const uint32_t *cell_offsets, uint32_t *ncount,
                                                               uint32_t *nbr, float skin, uint32_t block,
                                                               uint32_t S) {
    if (st->nstate[0] == 0) return;
    __shared__ uint32_t lds[Lds ? kLdsListCapacity * kBuildThreads : 1];
    uint32_t slot = blockIdx.x * kBuildThreads + threadIdx.x;
    if (slot >= S) return;
    const uint32_t tid = threadIdx.x;
    float4 A = slotA[slot], B = slotB[slot];
    uint4 M = slotC[slot];
    const uint32_t seg = M.x, owner = M.y;
    const float arc_p = B.w, arc_q = __uint_as_float(M.z);
    const float p1x = A.x, p1y = A.y, p1z = A.z, radius = A.w;
    const float q1x = B.x, q1y = B.y, q1z = B.z;
    const float d1x = q1x - p1x, d1y = q1y - p1y, d1z = q1z - p1z;
    const float half_length = 0.5f * sqrtf(d1x * d1x + d1y * d1y + d1z * d1z);
    const uint32_t home = seg_cell[seg];
    uint32_t count = 0;
    for (uint32_t n = 0; n < 27; ++n) {
        uint32_t cell;
        if (!stencil_cell(g, home, n, cell)) continue;
        uint32_t cs = cell_offsets[cell], cn = cell_counts[cell];
        for (uint32_t k = 0; k < cn; ++k) {
            uint32_t o = cs + k;
            uint4 OM = slotC[o];
            uint32_t other = OM.x;
            float4 OA = slotA[o], OB = slotB[o];
            float gap = fmaxf(OB.w - arc_q, arc_p - __uint_as_float(OM.z));
            bool near = OM.y == owner && gap < radius + OA.w;
            if (other == seg || near) continue;
            float p2x = OA.x, p2y = OA.y, p2z = OA.z, q2x = OB.x, q2y = OB.y, q2z = OB.z;
            if (g.periodic[0]) {
                float sh = roundf(0.5f * (p2x + q2x - p1x - q1x) / g.length[0]) * g.length[0];
                p2x -= sh;
                q2x -= sh;
            }
            if (g.periodic[1]) {
                float sh = roundf(0.5f * (p2y + q2y - p1y - q1y) / g.length[1]) * g.length[1];
                p2y -= sh;
                q2y -= sh;
            }
            if (g.periodic[2]) {
                float sh = roundf(0.5f * (p2z + q2z - p1z - q1z) / g.length[2]) * g.length[2];
                p2z -= sh;
                q2z -= sh;
            }
            float d2x = q2x - p2x, d2y = q2y - p2y, d2z = q2z - p2z;
            float inter = (radius + OA.w + skin) * (1.0f + 1.0e-4f);
            float mdx = 0.5f * (p2x + q2x - p1x - q1x);
            float mdy = 0.5f * (p2y + q2y - p1y - q1y);
            float mdz = 0.5f * (p2z + q2z - p1z - q1z);
            float reach = half_length + 0.5f * sqrtf(d2x * d2x + d2y * d2y + d2z * d2z) + inter;
            if (!(mdx * mdx + mdy * mdy + mdz * mdz <= reach * reach)) continue;
            if (!(axis_distance(p1x, p1y, p1z, d1x, d1y, d1z, p2x, p2y, p2z, d2x, d2y, d2z) <= inter)) continue;
            if (count < block) {
                uint32_t at = count;
                if (Lds) {
                    while (at > 0 && lds[(at - 1) * kBuildThreads + tid] > other) {
                        lds[at * kBuildThreads + tid] = lds[(at - 1) * kBuildThreads + tid];
                        --at;
                    }
                    lds[at * kBuildThreads + tid] = other;
                } else {
                    while (at > 0 && nbr[(size_t)(at - 1) * S + slot] > other) {
                        nbr[(size_t)at * S + slot] = nbr[(size_t)(at - 1) * S + slot];
                        --at;
                    }
                    nbr[(size_t)at * S + slot] = other;
                }
            }
            ++count;
        }
    }
    ncount[slot] = count;
    if (Lds) {
        uint32_t keep = count < block ? count : block;
        for (uint32_t k = 0; k < keep; ++k) nbr[(size_t)k * S + slot] = lds[k * kBuildThreads + tid];
    }
}

// ---------------------------------------------------------------------------
// contact projection (find_segment_corrections, contact.rs:232-902, round
// sections), one lane per slot. Thread 0 of the grid first closes a pending
// rebuild (finish_neighbor_list_rebuild, neighbor.rs:424-432).

__global__ __launch_bounds__(256) void k_contact(DevState *st, GridParams g, const float4 *pos,
                                                 const uint4 *sstat, const uint32_t *cell_segs,
                                                 const uint32_t *seg_cell, const uint32_t *cell_counts,
                                                 const uint32_t *cell_offsets, const uint32_t *ncount,
                                                 const uint32_t *nbr, float4 *corr0, float4 *corr1,
                                                 float correction_fraction, uint32_t aggregation, uint32_t block,
                                                 uint32_t S) {
    uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot == 0 && st->nstate[0] != 0) {
        st->nstate[0] = 0;
        st->nstate[1] += 1;
    }
    if (st->control[0] == 0) return;
    float max_pen = 0.f;
    if (slot < S) {
        const uint32_t seg = cell_segs[slot];
        const uint4 me = sstat[seg];
        const float4 P1 = pos[me.x], Q1 = pos[me.x + 1];
        const float p1x = P1.x, p1y = P1.y, p1z = P1.z, q1x = Q1.x, q1y = Q1.y, q1z = Q1.z;
        const float d1x = q1x - p1x, d1y = q1y - p1y, d1z = q1z - p1z;
        const uint32_t owner = me.y;
        const float radius = __uint_as_float(me.z);
        const float half_length = 0.5f * sqrtf(d1x * d1x + d1y * d1y + d1z * d1z);
        const float eps_len = 1.0e-20f, eps_par = 1.0e-6f;
        float c0x = 0.f, c0y = 0.f, c0z = 0.f, c1x = 0.f, c1y = 0.f, c1z = 0.f, cweight = 0.f;
        const uint32_t listed_count = ncount[slot];
        const bool listed = listed_count <= block;
        const uint32_t ranges = listed ? 1u : 27u;
        const uint32_t home = seg_cell[seg];
        for (uint32_t range = 0; range < ranges; ++range) {
            uint32_t count = 0, start = 0;
            if (listed) {
                count = listed_count;
            } else {
                uint32_t cell;
                if (stencil_cell(g, home, range, cell)) {
                    count = cell_counts[cell];
                    start = cell_offsets[cell];
                }
            }
            for (uint32_t k = 0; k < count; ++k) {
                uint32_t other = listed ? nbr[(size_t)k * S + slot] : cell_segs[start + k];
                uint4 ot = sstat[other];
                float4 P2 = pos[ot.x], Q2 = pos[ot.x + 1];
                float other_radius = __uint_as_float(ot.z);
                float gap = fmaxf(P2.w - Q1.w, P1.w - Q2.w);
                bool near = ot.y == owner && gap < radius + other_radius;
                if (other == seg || near) continue;
                float p2x = P2.x, p2y = P2.y, p2z = P2.z, q2x = Q2.x, q2y = Q2.y, q2z = Q2.z;
                float idx = 0.5f * (p2x + q2x - p1x - q1x);
                float idy = 0.5f * (p2y + q2y - p1y - q1y);
                float idz = 0.5f * (p2z + q2z - p1z - q1z);
                if (g.periodic[0]) {
                    float sh = roundf(idx / g.length[0]) * g.length[0];
                    p2x -= sh;
                    q2x -= sh;
                }
                if (g.periodic[1]) {
                    float sh = roundf(idy / g.length[1]) * g.length[1];
                    p2y -= sh;
                    q2y -= sh;
                }
                if (g.periodic[2]) {
                    float sh = roundf(idz / g.length[2]) * g.length[2];
                    p2z -= sh;
                    q2z -= sh;
                }
                float d2x = q2x - p2x, d2y = q2y - p2y, d2z = q2z - p2z;
                float mdx = 0.5f * (p2x + q2x - p1x - q1x);
                float mdy = 0.5f * (p2y + q2y - p1y - q1y);
                float mdz = 0.5f * (p2z + q2z - p1z - q1z);
                float reach = (half_length + 0.5f * sqrtf(d2x * d2x + d2y * d2y + d2z * d2z) + radius + other_radius) *
                              (1.0f + 1.0e-4f);
                if (!(mdx * mdx + mdy * mdy + mdz * mdz <= reach * reach)) continue;
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
                        bool parallel = fabsf(den) <= eps_par * a * e;
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
                            float gx = rx + d1x * s - d2x * t, gy = ry + d1y * s - d2y * t, gz = rz + d1z * s - d2z * t;
                            float ex = rx + d1x * es - d2x * et, ey = ry + d1y * es - d2y * et,
                                  ez = rz + d1z * es - d2z * et;
                            if (ex * ex + ey * ey + ez * ez < gx * gx + gy * gy + gz * gz) {
                                s = es;
                                t = et;
                            }
                        }
                    }
                }
                float dx = (p2x + d2x * t) - (p1x + d1x * s);
                float dy = (p2y + d2y * t) - (p1y + d1y * s);
                float dz = (p2z + d2z * t) - (p1z + d1z * s);
                float distance = sqrtf(dx * dx + dy * dy + dz * dz);
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
                    float dom = (fabsf(nx) >= fabsf(ny) && fabsf(nx) >= fabsf(nz)) ? nx
                                : (fabsf(ny) >= fabsf(nz) ? ny : nz);
                    if (dom < 0.f) { nx = -nx; ny = -ny; nz = -nz; }
                    if (seg > other) { nx = -nx; ny = -ny; nz = -nz; }
                }
                float w0 = 1.0f - s, w1 = s, ow0 = 1.0f - t;
                float den = w0 * w0 + w1 * w1 + ow0 * ow0 + t * t + 0.0f;
                float mag = -correction_fraction * pen / den;
                float aw = aggregation == 1u ? pen : 1.0f;
                if (aggregation == 2u) {
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
        if (aggregation != 2u && cweight > 1.0e-20f) {
            c0x /= cweight; c0y /= cweight; c0z /= cweight;
            c1x /= cweight; c1y /= cweight; c1z /= cweight;
        }
        corr0[seg] = make_float4(c0x, c0y, c0z, max_pen);
        corr1[seg] = make_float4(c1x, c1y, c1z, 0.f);
    }
    uint32_t m = wave_max_u32(__float_as_uint(fmaxf(max_pen, 0.f)));
    if ((threadIdx.x & 63) == 0 && m != 0u) atomicMax(&st->red_pen, m);
}

// ---------------------------------------------------------------------------
// curvature ratio (measure_curvature_ratio, kernels.rs:925-980) with the
// metric reductions (kernels.rs:95-127) and, in the last workgroup to finish,
// the convergence assessment (assess_reduced_metrics, kernels.rs:129-162).

__global__ __launch_bounds__(256) void k_measure_assess(DevState *st, const float4 *pos, const float4 *vA,
                                                        const float4 *vB, uint32_t V, float penetration_tolerance,
                                                        float curvature_ratio_tolerance, uint32_t iterations,
                                                        uint32_t force_full) {
    if (st->control[0] == 0) return;
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    float ratio = 0.f;
    if (v < V) {
        float kmax = vA[v].x;
        uint32_t flags = __float_as_uint(vB[v].z);
        if (kmax > 0.f && (flags & kHasPrev) && (flags & kHasNext)) {
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
    uint32_t m = wave_max_u32(__float_as_uint(fmaxf(ratio, 0.f)));
    if ((threadIdx.x & 63) == 0 && m != 0u) atomicMax(&st->red_curv, m);
    __shared__ uint32_t last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        last = atomicAdd(&st->done_blocks, 1u) == gridDim.x - 1 ? 1u : 0u;
    }
    __syncthreads();
    if (last == 0u || threadIdx.x != 0) return;
    __threadfence();
    uint32_t pen_bits = atomicAdd(&st->red_pen, 0u);
    uint32_t step_bits = atomicAdd(&st->red_step, 0u);
    uint32_t curv_bits = atomicAdd(&st->red_curv, 0u);
    float pen = __uint_as_float(pen_bits), curv = __uint_as_float(curv_bits);
    st->metrics[0] = pen;
    st->metrics[1] = __uint_as_float(step_bits);
    st->metrics[2] = curv;
    bool stop = (force_full == 0u && pen <= penetration_tolerance && curv <= 1.0f + curvature_ratio_tolerance) ||
                st->control[2] >= iterations;
    if (stop) {
        st->control[0] = 0;
    } else {
        st->control[1] += 1;
        st->control[2] += 1;
        st->red_step = 0u;  // recomputed by this iteration's contact application
    }
    st->red_pen = 0u;
    st->red_curv = 0u;
    st->done_blocks = 0u;
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

__global__ __launch_bounds__(256) void k_vertex_pass(DevState *st, GridParams g, const float4 *in, float4 *out,
                                                     const float4 *vA, const float4 *vB, const float4 *corr0,
                                                     const float4 *corr1, float4 *wall, uint32_t V,
                                                     float stretch, float bend, float max_step, uint32_t do_internal,
                                                     uint32_t do_contact, uint32_t first_pass, uint32_t track_walls) {
    if (st->control[0] == 0) return;
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    float step = 0.f;
    if (v < V) {
        float4 me = in[v];
        float4 sa = vA[v], sb = vB[v];
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
    if (do_contact) {
        uint32_t m = wave_max_u32(__float_as_uint(fmaxf(step, 0.f)));
        if ((threadIdx.x & 63) == 0 && m != 0u) atomicMax(&st->red_step, m);
    }
}

// ---------------------------------------------------------------------------
// Per-fiber curvature Gauss–Seidel (project_fiber_curvature_in_place and
// project_curvature_triplet_in_place, kernels.rs:1165-1415), one lane per
// fiber, then the neighbour-list displacement trigger for the fiber's
// vertices (flag_neighbor_list_displacement, neighbor.rs:443-465).

// Projects one triplet in material order (previous, center, next) held in
// registers, accumulating wall reactions into wp/wc/wn in the reference's
// order. Returns true when the projection branch ran.
__device__ __forceinline__ bool project_triplet(float3 &pp, float3 &pc, float3 &pn, float kmax_center,
                                                bool pin_p, bool pin_c, bool pin_n, float ext_p, float ext_c,
                                                float ext_n, const GridParams &g, float stiffness, float margin,
                                                float3 &wp, float3 &wc, float3 &wn) {
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
    if (sine > 1.0e-3f) violation = (acosf(cosine) - acosf(min_cos)) * sine;
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
    float *P = &pp.x, *Cc = &pc.x, *N = &pn.x, *WP = &wp.x, *WC = &wc.x, *WN = &wn.x;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float mp = P[k] - fwp * scale * gp[k];
        float mc = Cc[k] + fwc * scale * gc[k];
        float mn = N[k] + fwn * scale * gn[k];
        float vp = mp, vc = mc, vn = mn;
        if (!g.periodic[k]) {
            vp = rclamp(mp, g.lower[k] + ext_p, g.upper[k] - ext_p);
            vc = rclamp(mc, g.lower[k] + ext_c, g.upper[k] - ext_c);
            vn = rclamp(mn, g.lower[k] + ext_n, g.upper[k] - ext_n);
            if (fwp > 0.f) WP[k] += mp - vp;
            if (fwc > 0.f) WC[k] += mc - vc;
            if (fwn > 0.f) WN[k] += mn - vn;
        }
        if (fwp > 0.f) P[k] = vp;
        if (fwc > 0.f) Cc[k] = vc;
        if (fwn > 0.f) N[k] = vn;
    }
    return true;
}

__device__ __forceinline__ float3 load3(const float4 *p, int32_t i) {
    float4 v = p[i];
    return make_float3(v.x, v.y, v.z);
}

__global__ __launch_bounds__(64) void k_curvature_gs(DevState *st, GridParams g, const float4 *in, float4 *pos,
                                                     const float4 *refpos, const float4 *vA, const float4 *vB,
                                                     const uint2 *fspan, float4 *wall, uint32_t F, uint32_t sweeps,
                                                     float stiffness, float margin, float half_skin,
                                                     uint32_t track_walls) {
    if (st->control[0] == 0) return;
    uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const uint2 span = fspan[f];
    const uint32_t start = span.x, count = span.y;
    if (in != pos) {
        for (uint32_t v = start; v < start + count; ++v) pos[v] = in[v];
    }
    const float3 zero3 = make_float3(0.f, 0.f, 0.f);
    if (count >= 3) {
        for (uint32_t sweep = 0; sweep < sweeps; ++sweep) {
            bool any = false;
            for (int dir = 0; dir < 2; ++dir) {
                // Window over three consecutive vertices in visiting order:
                // a (oldest), b, c (just loaded). Forward, (a, b, c) is
                // (previous, center, next); backward it is (next, center,
                // previous) in material order.
                const int32_t stepv = dir == 0 ? 1 : -1;
                const int32_t first = dir == 0 ? (int32_t)start : (int32_t)(start + count - 1);
                int32_t ia = first, ib = first + stepv;
                float3 a = load3(pos, ia), b = load3(pos, ib);
                float3 wa = track_walls ? load3(wall, ia) : zero3, wb = track_walls ? load3(wall, ib) : zero3;
                bool moved_a = false, moved_b = false;
                for (uint32_t l = 2; l < count; ++l) {
                    const int32_t ic = first + stepv * (int32_t)l;
                    float3 c = load3(pos, ic);
                    float3 wc = track_walls ? load3(wall, ic) : zero3;
                    const int32_t ip = dir == 0 ? ia : ic, inx = dir == 0 ? ic : ia;
                    const float4 sp = vA[ip], sc = vA[ib], sn = vA[inx];
                    const uint32_t fp = __float_as_uint(vB[ip].z), fc = __float_as_uint(vB[ib].z),
                                   fn = __float_as_uint(vB[inx].z);
                    bool ran = dir == 0
                                   ? project_triplet(a, b, c, sc.x, fp & kPinned, fc & kPinned, fn & kPinned, sp.w,
                                                     sc.w, sn.w, g, stiffness, margin, wa, wb, wc)
                                   : project_triplet(c, b, a, sc.x, fp & kPinned, fc & kPinned, fn & kPinned, sp.w,
                                                     sc.w, sn.w, g, stiffness, margin, wc, wb, wa);
                    if (ran) any = true;
                    moved_a = moved_a || ran;
                    if (moved_a) {  // vertex a leaves the window
                        pos[ia] = make_float4(a.x, a.y, a.z, pos[ia].w);
                        if (track_walls) wall[ia] = make_float4(wa.x, wa.y, wa.z, 0.f);
                    }
                    a = b; wa = wb; moved_a = moved_b || ran; ia = ib;
                    b = c; wb = wc; moved_b = ran; ib = ic;
                }
                if (moved_a) {
                    pos[ia] = make_float4(a.x, a.y, a.z, pos[ia].w);
                    if (track_walls) wall[ia] = make_float4(wa.x, wa.y, wa.z, 0.f);
                }
                if (moved_b) {
                    pos[ib] = make_float4(b.x, b.y, b.z, pos[ib].w);
                    if (track_walls) wall[ib] = make_float4(wb.x, wb.y, wb.z, 0.f);
                }
                if (dir == 0 && !any) break;
            }
            if (!any) break;
        }
    }
    const float limit = half_skin * half_skin;
    bool flag = false;
    for (uint32_t v = start; v < start + count; ++v) {
        float4 p = pos[v], r = refpos[v];
        float dx = p.x - r.x, dy = p.y - r.y, dz = p.z - r.z;
        if (dx * dx + dy * dy + dz * dz > limit) flag = true;
    }
    if (flag) st->nstate[0] = 1u;
}

}  // namespace tgh

// ---------------------------------------------------------------------------
// launchers

namespace tgh {

namespace {
inline uint32_t blocks_for(uint64_t n, uint32_t threads) { return (uint32_t)((n + threads - 1) / threads); }
}  // namespace

hipError_t launch_begin_batch(const WorldDev &w, hipStream_t stream) {
    hipLaunchKernelGGL(k_begin_batch, dim3(1), dim3(64), 0, stream, w.st);
    return hipGetLastError();
}

hipError_t launch_rebuild(const WorldDev &w, hipStream_t stream) {
    if (w.S == 0 || w.C == 0 || w.V == 0) return hipErrorInvalidValue;
    const uint32_t clear_n = w.C > w.V ? w.C : w.V;
    hipLaunchKernelGGL(k_rebuild_clear, dim3(blocks_for(clear_n, 256)), dim3(256), 0, stream, w.st, w.cell_counts,
                       w.cell_cursor, w.C, w.pos, w.refpos, w.V);
    hipLaunchKernelGGL(k_bin_segments, dim3(blocks_for(w.S, 256)), dim3(256), 0, stream, w.st, w.g, w.pos, w.sstat,
                       w.seg_cell, w.cell_counts, w.S);
    hipLaunchKernelGGL(k_scan_cells, dim3(1), dim3(1024), 0, stream, w.st, w.cell_counts, w.cell_offsets, w.C);
    hipLaunchKernelGGL(k_scatter_segments, dim3(blocks_for(w.S, 256)), dim3(256), 0, stream, w.st, w.seg_cell,
                       w.cell_offsets, w.cell_cursor, w.scattered, w.S);
    hipLaunchKernelGGL(k_rank_cells, dim3(blocks_for(w.C, 4)), dim3(256), 0, stream, w.st, w.cell_counts,
                       w.cell_offsets, w.scattered, w.cell_segs, w.pos, w.sstat, w.slotA, w.slotB, w.slotC, w.C);
    if (w.block <= kLdsListCapacity) {
        hipLaunchKernelGGL(k_build_lists<true>, dim3(blocks_for(w.S, kBuildThreads)), dim3(kBuildThreads), 0, stream,
                           w.st, w.g, w.slotA, w.slotB, w.slotC, w.seg_cell, w.cell_counts, w.cell_offsets, w.ncount,
                           w.nbr, w.skin, w.block, w.S);
    } else {
        hipLaunchKernelGGL(k_build_lists<false>, dim3(blocks_for(w.S, kBuildThreads)), dim3(kBuildThreads), 0,
                           stream, w.st, w.g, w.slotA, w.slotB, w.slotC, w.seg_cell, w.cell_counts, w.cell_offsets,
                           w.ncount, w.nbr, w.skin, w.block, w.S);
    }
    return hipGetLastError();
}

// Rebuilds the cell list and neighbour lists from the current positions now,
// whatever the displacement trigger says: request_neighbor_list_rebuild
// (neighbor.rs:434-441), the rebuild, and finish_neighbor_list_rebuild.
__global__ void k_request_rebuild(DevState *st) {
    if (threadIdx.x == 0 && blockIdx.x == 0) st->nstate[0] = 1u;
}

__global__ void k_finish_rebuild(DevState *st) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || st->nstate[0] == 0) return;
    st->nstate[0] = 0u;
    st->nstate[1] += 1u;
}

hipError_t launch_rebuild_now(const WorldDev &w, hipStream_t stream) {
    hipLaunchKernelGGL(k_request_rebuild, dim3(1), dim3(64), 0, stream, w.st);
    hipError_t e = launch_rebuild(w, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_finish_rebuild, dim3(1), dim3(64), 0, stream, w.st);
    return hipGetLastError();
}

hipError_t launch_contact(const WorldDev &w, const IterParams &p, hipStream_t stream) {
    hipLaunchKernelGGL(k_contact, dim3(blocks_for(w.S, 256)), dim3(256), 0, stream, w.st, w.g, w.pos, w.sstat,
                       w.cell_segs, w.seg_cell, w.cell_counts, w.cell_offsets, w.ncount, w.nbr, w.corr0, w.corr1,
                       p.correction_fraction, p.aggregation, w.block, w.S);
    return hipGetLastError();
}

hipError_t launch_measure(const WorldDev &w, const IterParams &p, hipStream_t stream) {
    hipLaunchKernelGGL(k_measure_assess, dim3(blocks_for(w.V, 256)), dim3(256), 0, stream, w.st, w.pos, w.vA, w.vB,
                       w.V, p.penetration_tolerance, p.curvature_ratio_tolerance, p.iterations, p.force_full);
    return hipGetLastError();
}

// Jacobi passes alternate between the two position buffers; the last one
// also applies contact. The curvature kernel reads the buffer the last pass
// wrote and writes back to `pos`, so every iteration ends in `pos`.
static uint32_t vertex_passes(const IterParams &p) { return p.constraint_iterations > 0 ? p.constraint_iterations : 1u; }

hipError_t launch_vertex_passes(const WorldDev &w, const IterParams &p, hipStream_t stream) {
    const uint32_t passes = vertex_passes(p);
    float4 *src = w.pos, *dst = w.alt;
    for (uint32_t i = 0; i < passes; ++i) {
        hipLaunchKernelGGL(k_vertex_pass, dim3(blocks_for(w.V, 256)), dim3(256), 0, stream, w.st, w.g, src, dst,
                           w.vA, w.vB, w.corr0, w.corr1, w.wall, w.V, p.stretch, p.bend, p.max_step,
                           p.constraint_iterations > 0 ? 1u : 0u, i + 1 == passes ? 1u : 0u, i == 0 ? 1u : 0u,
                           w.track_walls);
        float4 *t = src;
        src = dst;
        dst = t;
    }
    return hipGetLastError();
}

hipError_t launch_curvature(const WorldDev &w, const IterParams &p, hipStream_t stream) {
    const float4 *src = (vertex_passes(p) % 2 == 1) ? w.alt : w.pos;
    hipLaunchKernelGGL(k_curvature_gs, dim3(blocks_for(w.F, 64)), dim3(64), 0, stream, w.st, w.g, src, w.pos,
                       w.refpos, w.vA, w.vB, w.fspan, w.wall, w.F, p.sweeps, p.curvature_stiffness,
                       p.curvature_margin, 0.5f * w.skin, w.track_walls);
    return hipGetLastError();
}

}  // namespace tgh
