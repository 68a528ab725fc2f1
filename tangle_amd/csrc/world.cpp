// Host side of the HIP relaxation world: packs a solver-neutral assembly into
// the device layout, sizes the uniform grid and neighbour lists, owns the
// device buffers and drives batches of relaxation iterations as captured
// hipGraphs. Exposes the C ABI declared in include/tangle_hip.h.
//
// Reference behaviour followed here:
//   packing        crates/tangle_relax/src/device/packed.rs:150-320, 527-580
//   sizing         crates/tangle_relax/src/device/world.rs:322-417, 496-509, 868-917
//   batch/status   crates/tangle_relax/src/device/world.rs:1020-1505, 2337-2369
//   defaults       crates/tangle_relax/src/device/plugin.rs:145-175,
//                  crates/tangle_contact/src/lib.rs:32-40

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <string>
#include <vector>

#include "../../include/tangle_hip.h"
#include "errors.h"
#include "relax_device.h"
#include "world_internal.h"

using tgh::DevState;
using tgh::SpecState;
using tgh::GridParams;
using tgh::IterParams;
using tgh::WorldDev;
using tgh::Packed;
using tgh::Sizing;

namespace {

thread_local std::string g_error;

}  // namespace

int tgh::fail(int code, const std::string &message) {
    g_error = message;
    return code;
}

namespace {

inline int set_error(int code, const std::string &message) { return tgh::fail(code, message); }

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess)                                                                      \
            return set_error(TGH_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e));  \
    } while (0)

constexpr size_t kReferenceMaxListBytes = size_t(120) << 20;  // world.rs:36

inline float fclamp_rust(float x, float lo, float hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}
inline uint32_t to_u32_saturating(float x) {
    if (std::isnan(x) || x <= 0.0f) return 0;
    if (x >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)x;
}

int pack_assembly(const tgh_assembly *a, Packed &p) {
    if (!a) return set_error(TGH_ERR_INVALID, "assembly is null");
    if (a->fiber_count == 0 || a->vertex_count == 0)
        return set_error(TGH_ERR_INVALID, "assembly holds no fibers");
    if (!a->fiber_vertex_counts || !a->placed || !a->intrinsic || !a->fiber_radius)
        return set_error(TGH_ERR_INVALID, "assembly is missing an array");
    for (int k = 0; k < 3; ++k)
        if (!(a->cell_lengths[k] > 0.0) || !std::isfinite(a->cell_lengths[k]))
            return set_error(TGH_ERR_INVALID, "cell lengths must be positive and finite");
    p.F = a->fiber_count;
    p.V = a->vertex_count;
    uint64_t sum = 0;
    p.fstart.resize(p.F);
    p.fcount.resize(p.F);
    for (uint32_t f = 0; f < p.F; ++f) {
        uint32_t n = a->fiber_vertex_counts[f];
        if (n < 2) return set_error(TGH_ERR_INVALID, "fiber " + std::to_string(f) + " has fewer than two vertices");
        p.fstart[f] = (uint32_t)sum;
        p.fcount[f] = n;
        sum += n;
    }
    if (sum != p.V) return set_error(TGH_ERR_INVALID, "fiber vertex counts do not add up to vertex_count");
    p.S = p.V - p.F;
    p.pos.resize(3 * (size_t)p.V);
    p.intr.resize(3 * (size_t)p.V);
    for (size_t i = 0; i < 3 * (size_t)p.V; ++i) {
        if (!std::isfinite(a->placed[i]) || !std::isfinite(a->intrinsic[i]))
            return set_error(TGH_ERR_INVALID, "assembly positions must be finite");
        p.pos[i] = (float)a->placed[i];
        p.intr[i] = (float)a->intrinsic[i];
    }
    p.radius.resize(p.F);
    p.rest.resize(p.S);
    p.kmax.assign(p.V, 0.0f);
    p.pinned.assign(p.V, 0);
    for (uint32_t f = 0; f < p.F; ++f) {
        double r = a->fiber_radius[f];
        if (!(r > 0.0) || !std::isfinite(r))
            return set_error(TGH_ERR_INVALID, "fiber " + std::to_string(f) + " radius must be positive");
        p.radius[f] = (float)r;
        const uint32_t st = p.fstart[f], n = p.fcount[f];
        for (uint32_t l = 0; l + 1 < n; ++l) {
            const double *x0 = &a->intrinsic[3 * (size_t)(st + l)];
            const double *x1 = &a->intrinsic[3 * (size_t)(st + l + 1)];
            double dx = x1[0] - x0[0], dy = x1[1] - x0[1], dz = x1[2] - x0[2];
            p.rest[st - f + l] = (float)std::sqrt(dx * dx + dy * dy + dz * dz);
        }
        double limit = a->fiber_min_bend_radius ? a->fiber_min_bend_radius[f] : 0.0;
        if (limit > 0.0 && std::isfinite(limit)) {
            float k = (float)(1.0 / limit);
            for (uint32_t l = 1; l + 1 < n; ++l) p.kmax[st + l] = k;
        }
        if (a->pin_fiber_ends) {
            p.pinned[st] = 1;
            p.pinned[st + n - 1] = 1;
        }
    }
    if (a->vertex_pinned)
        for (uint32_t v = 0; v < p.V; ++v) p.pinned[v] |= a->vertex_pinned[v] ? 1 : 0;
    p.fstep.assign(p.F, 0u);
    p.flayer.assign(p.F, 0xffffffffu);
    for (uint32_t f = 0; f < p.F; ++f) {
        if (a->fiber_formation_step) p.fstep[f] = a->fiber_formation_step[f];
        if (a->fiber_formation_layer) p.flayer[f] = a->fiber_formation_layer[f];
    }
    // rest arc length along the intrinsic centreline, f64 sums of f32 data
    p.arc.assign(p.V, 0.0f);
    for (uint32_t f = 0; f < p.F; ++f) {
        double total = 0.0;
        for (uint32_t v = p.fstart[f] + 1; v < p.fstart[f] + p.fcount[f]; ++v) {
            double sq = 0.0;
            for (int k = 0; k < 3; ++k) {
                double d = (double)p.intr[3 * (size_t)v + k] - (double)p.intr[3 * (size_t)(v - 1) + k];
                sq += d * d;
            }
            total += std::sqrt(sq);
            p.arc[v] = (float)total;
        }
    }
    for (int k = 0; k < 3; ++k) {
        p.lower[k] = (float)a->cell_origin[k];
        p.upper[k] = (float)(a->cell_origin[k] + a->cell_lengths[k]);
        p.periodic[k] = a->cell_periodic[k] ? 1u : 0u;
    }
    return TGH_OK;
}

// PackedAssembly (packed.rs:47-141) of a uniform, round layout: the f32
// arrays are taken as they are; the layout must be the one
// from_assembly_with_options writes for such an assembly (packed.rs:150-320):
// fibers contiguous, segment f-th fiber's l-th = vertices (first + l, first +
// l + 1) numbered first - f + l, one radius per fiber, every segment active.
int pack_packed(const tgh_packed *a, Packed &p) {
    if (!a) return set_error(TGH_ERR_INVALID, "packed assembly is null");
    if (a->fiber_count == 0 || a->vertex_count == 0 || a->segment_count == 0)
        return set_error(TGH_ERR_INVALID, "packed assembly holds no fibers");
    if (!a->positions || !a->intrinsic_positions || !a->segment_vertices || !a->segment_fibers ||
        !a->segment_radii || !a->segment_rest_lengths || !a->fiber_vertex_spans || !a->vertex_max_curvature ||
        !a->vertex_pinned)
        return set_error(TGH_ERR_INVALID, "packed assembly is missing an array");
    if (a->adaptive) return set_error(TGH_ERR_INVALID, "adaptive (dyadic tree) layouts are not supported");
    if (a->has_ovals) return set_error(TGH_ERR_INVALID, "oval sections are not supported");
    p.F = a->fiber_count;
    p.V = a->vertex_count;
    if ((uint64_t)p.V < (uint64_t)p.F * 2 || a->segment_count != p.V - p.F)
        return set_error(TGH_ERR_INVALID, "segment_count must equal vertex_count - fiber_count");
    p.S = a->segment_count;
    p.fstart.resize(p.F);
    p.fcount.resize(p.F);
    p.radius.resize(p.F);
    uint64_t next = 0;
    for (uint32_t f = 0; f < p.F; ++f) {
        uint32_t st = a->fiber_vertex_spans[2 * (size_t)f], n = a->fiber_vertex_spans[2 * (size_t)f + 1];
        if (st != next || n < 2)
            return set_error(TGH_ERR_INVALID, "fiber " + std::to_string(f) +
                                                  " vertex span must follow the previous one and hold two vertices");
        p.fstart[f] = st;
        p.fcount[f] = n;
        next += n;
    }
    if (next != p.V) return set_error(TGH_ERR_INVALID, "fiber vertex spans do not cover vertex_count");
    for (uint32_t f = 0; f < p.F; ++f) {
        const uint32_t st = p.fstart[f], n = p.fcount[f];
        for (uint32_t l = 0; l + 1 < n; ++l) {
            const size_t s = (size_t)st - f + l;
            if (a->segment_vertices[2 * s] != st + l || a->segment_vertices[2 * s + 1] != st + l + 1 ||
                a->segment_fibers[s] != f)
                return set_error(TGH_ERR_INVALID, "segment " + std::to_string(s) +
                                                      " is not the uniform layout's segment of fiber " +
                                                      std::to_string(f));
            if (a->segment_active && a->segment_active[s] == 0)
                return set_error(TGH_ERR_INVALID, "inactive segments (adaptive layouts) are not supported");
            float r = a->segment_radii[s];
            if (!(r > 0.0f) || !std::isfinite(r))
                return set_error(TGH_ERR_INVALID, "segment " + std::to_string(s) + " radius must be positive");
            if (l == 0)
                p.radius[f] = r;
            else if (r != p.radius[f])
                return set_error(TGH_ERR_INVALID, "fiber " + std::to_string(f) + " segments differ in radius");
        }
    }
    p.pos.assign(a->positions, a->positions + 3 * (size_t)p.V);
    p.intr.assign(a->intrinsic_positions, a->intrinsic_positions + 3 * (size_t)p.V);
    for (size_t i = 0; i < 3 * (size_t)p.V; ++i)
        if (!std::isfinite(p.pos[i]) || !std::isfinite(p.intr[i]))
            return set_error(TGH_ERR_INVALID, "packed positions must be finite");
    p.rest.assign(a->segment_rest_lengths, a->segment_rest_lengths + p.S);
    for (float r : p.rest)
        if (!(r >= 0.0f) || !std::isfinite(r)) return set_error(TGH_ERR_INVALID, "rest lengths must be finite");
    p.kmax.assign(a->vertex_max_curvature, a->vertex_max_curvature + p.V);
    for (float k : p.kmax)
        if (!(k >= 0.0f) || !std::isfinite(k))
            return set_error(TGH_ERR_INVALID, "vertex_max_curvature must be finite and non-negative");
    p.pinned.resize(p.V);
    for (uint32_t v = 0; v < p.V; ++v) p.pinned[v] = a->vertex_pinned[v] ? 1 : 0;
    p.fstep.assign(p.F, 0u);
    p.flayer.assign(p.F, 0xffffffffu);
    for (uint32_t f = 0; f < p.F; ++f) {
        if (a->fiber_formation_steps) p.fstep[f] = a->fiber_formation_steps[f];
        if (a->fiber_formation_layers) p.flayer[f] = a->fiber_formation_layers[f];
    }
    // vertex_arc_lengths (packed.rs:561-580): f64 sums over the f32 intrinsic chain
    p.arc.assign(p.V, 0.0f);
    for (uint32_t f = 0; f < p.F; ++f) {
        double total = 0.0;
        for (uint32_t v = p.fstart[f] + 1; v < p.fstart[f] + p.fcount[f]; ++v) {
            double sq = 0.0;
            for (int k = 0; k < 3; ++k) {
                double d = (double)p.intr[3 * (size_t)v + k] - (double)p.intr[3 * (size_t)(v - 1) + k];
                sq += d * d;
            }
            total += std::sqrt(sq);
            p.arc[v] = (float)total;
        }
    }
    for (int k = 0; k < 3; ++k) {
        if (!std::isfinite(a->cell_lower[k]) || !std::isfinite(a->cell_upper[k]) ||
            !(a->cell_upper[k] > a->cell_lower[k]))
            return set_error(TGH_ERR_INVALID, "cell bounds must be finite with upper > lower");
        p.lower[k] = a->cell_lower[k];
        p.upper[k] = a->cell_upper[k];
        p.periodic[k] = a->cell_periodic[k] ? 1u : 0u;
    }
    return TGH_OK;
}

int validate_config(const tgh_config *c) {
    if (!c) return set_error(TGH_ERR_INVALID, "config is null");
    if (!(c->cell_size_scale >= 1.0f) || !std::isfinite(c->cell_size_scale))
        return set_error(TGH_ERR_INVALID, "cell_size_scale must be finite and at least 1");
    if (!std::isfinite(c->neighbor_skin_scale) || c->neighbor_skin_scale < 0.0f)
        return set_error(TGH_ERR_INVALID, "neighbor_skin_scale must be finite and non-negative");
    if (c->neighbor_capacity == 0) return set_error(TGH_ERR_INVALID, "neighbor_capacity must be at least 1");
    if (!(c->max_step > 0.0f) || !std::isfinite(c->max_step))
        return set_error(TGH_ERR_INVALID, "max_step must be positive and finite");
    if (!std::isfinite(c->penetration_tolerance) || c->penetration_tolerance < 0.0f)
        return set_error(TGH_ERR_INVALID, "penetration_tolerance must be nonnegative and finite");
    if (c->contact_aggregation > 2) return set_error(TGH_ERR_INVALID, "contact_aggregation must be 0, 1 or 2");
    return TGH_OK;
}

float segment_length_now(const Packed &p, uint32_t f, uint32_t l) {
    const float *a = &p.pos[3 * (size_t)(p.fstart[f] + l)];
    const float *b = a + 3;
    float dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
    return std::sqrt(dx * dx + dy * dy + dz * dz);
}

int size_grid(const Packed &p, const tgh_config *c, Sizing &z) {
    float max_rest = 0.f, max_init = 0.f, max_r = 0.f, max_leaf = 0.f;
    double ext_sum = 0.0;  // half length plus radius, summed over segments
    uint64_t segments = 0;
    for (uint32_t f = 0; f < p.F; ++f) {
        max_r = std::fmax(max_r, p.radius[f]);
        for (uint32_t l = 0; l + 1 < p.fcount[f]; ++l) {
            float rest = p.rest[p.fstart[f] - f + l];
            float now = segment_length_now(p, f, l);
            max_rest = std::fmax(max_rest, rest);
            max_init = std::fmax(max_init, now);
            max_leaf = std::fmax(max_leaf, std::fmax(rest, now));
            ext_sum += 0.5 * now + p.radius[f];
            ++segments;
        }
    }
    const float max_len = std::fmax(max_rest, max_init);
    z.skin = c->neighbor_skin_scale * max_r;
    const float margin = 2.0f * max_r + std::fmax(2.0f * c->max_step, z.skin);
    const float block_len = std::fmin(max_len, std::fmax(max_leaf, 2.0f * margin));
    // A segment longer than one list block would be binned as several proxy
    // pieces; with a uniform (non-adaptive) layout every segment is at most
    // the longest one, so this never triggers, but guard it all the same.
    for (uint32_t f = 0; f < p.F; ++f)
        for (uint32_t l = 0; l + 1 < p.fcount[f]; ++l) {
            float len = std::fmax(p.rest[p.fstart[f] - f + l], segment_length_now(p, f, l));
            if (to_u32_saturating(std::ceil(len / block_len)) > 1)
                return set_error(TGH_ERR_INVALID, "segments split into proxy pieces are not supported");
        }
    const float piece = block_len < max_len ? 1.05f * block_len : block_len;
    z.cell_size = (piece + margin) * c->cell_size_scale;
    z.segment_cell_size = (max_len + margin) * c->cell_size_scale;
    z.max_radius = max_r;
    z.list_reach = (float)(2.0 * ext_sum / (double)std::max<uint64_t>(1, segments)) + z.skin;
    uint64_t C = 1;
    for (int k = 0; k < 3; ++k) {
        float extent = p.upper[k] - p.lower[k];
        z.cells[k] = to_u32_saturating(std::fmax(std::floor(extent / z.cell_size), 1.0f));
        C *= z.cells[k];
    }
    if (C > (uint64_t(1) << 28)) return set_error(TGH_ERR_INVALID, "cell grid exceeds 2^28 cells");
    z.C = (uint32_t)C;
    const size_t max_bytes = c->maximum_neighbor_list_bytes ? c->maximum_neighbor_list_bytes : kReferenceMaxListBytes;
    z.slots = std::max<size_t>(1, std::min<size_t>((size_t)c->neighbor_capacity * p.S, max_bytes / 4));
    z.block = (uint32_t)std::min<size_t>(c->neighbor_capacity, z.slots / std::max<uint32_t>(1, p.S));
    return TGH_OK;
}

IterParams iter_params(const tgh_config *c, uint32_t iterations) {
    IterParams q;
    q.penetration_tolerance = c->penetration_tolerance;
    q.curvature_ratio_tolerance = c->curvature_ratio_tolerance;
    q.correction_fraction = c->correction_fraction;
    q.stretch = c->stretch_stiffness;
    q.bend = c->bend_stiffness;
    q.max_step = c->max_step;
    q.curvature_stiffness = c->curvature_limit_stiffness;
    q.curvature_margin = c->curvature_limit_safety_margin;
    q.aggregation = c->contact_aggregation;
    q.constraint_iterations = c->constraint_iterations;
    q.sweeps = c->curvature_cleanup_sweeps;
    q.force_full = c->force_full_iterations ? 1u : 0u;
    q.iterations = iterations;
    return q;
}

}  // namespace

namespace {

int upload_static(tgh_world *w) {
    WorldDev &d = w->d;
    const uint32_t V = d.V, S = d.S, F = d.F;
    std::vector<float4> pos4(V), vA(V), vB(V);
    std::vector<uint4> sstat(S);
    std::vector<uint4> fspan(F);
    for (uint32_t sp = 0; sp < d.nspec; ++sp) {
        const Packed &p = w->specimens[sp];
        const uint32_t vb = w->vbase[sp], fb = w->fbase[sp], sb = w->sbase[sp];
        for (uint32_t f = 0; f < p.F; ++f) {
            const uint32_t st = p.fstart[f], n = p.fcount[f];
            const uint32_t gf = fb + f;  // world fiber index
            fspan[gf] = make_uint4(vb + st, n, sp, 0u);
            for (uint32_t l = 0; l < n; ++l) {
                const uint32_t v = st + l, gv = vb + v;
                pos4[gv] = make_float4(p.pos[3 * (size_t)v], p.pos[3 * (size_t)v + 1], p.pos[3 * (size_t)v + 2],
                                       p.arc[v]);
                uint32_t flags = 0;
                float rest_prev = 0.f, rest_next = 0.f, chord_prev = 0.f, chord_next = 0.f;
                if (l > 0) {
                    flags |= tgh::kHasPrev;
                    rest_prev = p.rest[v - f - 1];
                }
                if (l + 1 < n) {
                    flags |= tgh::kHasNext;
                    rest_next = p.rest[v - f];
                }
                // rest chords to the second neighbours, f32 as in kernels.rs:1098-1101
                auto chord = [&](uint32_t other) {
                    const float *I = p.intr.data();
                    float rx = I[3 * (size_t)other] - I[3 * (size_t)v];
                    float ry = I[3 * (size_t)other + 1] - I[3 * (size_t)v + 1];
                    float rz = I[3 * (size_t)other + 2] - I[3 * (size_t)v + 2];
                    return std::sqrt(rx * rx + ry * ry + rz * rz);
                };
                if (l >= 2) {
                    flags |= tgh::kHasPrev2;
                    chord_prev = chord(v - 2);
                }
                if (l + 2 < n) {
                    flags |= tgh::kHasNext2;
                    chord_next = chord(v + 2);
                }
                if (p.pinned[v]) flags |= tgh::kPinned;
                flags |= sp << tgh::kSpecShift;
                float flag_bits, fiber_bits;
                std::memcpy(&flag_bits, &flags, 4);
                std::memcpy(&fiber_bits, &gf, 4);
                vA[gv] = make_float4(p.kmax[v], rest_prev, rest_next, p.radius[f]);
                vB[gv] = make_float4(chord_prev, chord_next, flag_bits, fiber_bits);
            }
            for (uint32_t l = 0; l + 1 < n; ++l) {
                const uint32_t seg = sb + (st - f + l);
                uint32_t rbits;
                std::memcpy(&rbits, &p.radius[f], 4);
                sstat[seg] = make_uint4(vb + st + l, gf, rbits, sp);
            }
        }
    }
    HIP_TRY(hipMemcpy(d.pos, pos4.data(), V * sizeof(float4), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.refpos, pos4.data(), V * sizeof(float4), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.alt, pos4.data(), V * sizeof(float4), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.vA, vA.data(), V * sizeof(float4), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.vB, vB.data(), V * sizeof(float4), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.sstat, sstat.data(), S * sizeof(uint4), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.fspan, fspan.data(), F * sizeof(uint4), hipMemcpyHostToDevice));
    {
        std::vector<uint32_t> fstep, flayer, factive(F, 1u);  // every fiber active at upload (packed.rs:289, 303)
        for (const Packed &p : w->specimens) {
            fstep.insert(fstep.end(), p.fstep.begin(), p.fstep.end());
            flayer.insert(flayer.end(), p.flayer.begin(), p.flayer.end());
        }
        HIP_TRY(hipMemcpy(d.fstep, fstep.data(), F * sizeof(uint32_t), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d.flayer, flayer.data(), F * sizeof(uint32_t), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d.factive, factive.data(), F * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMemcpy((void *)d.grids, w->grids.data(), d.nspec * sizeof(GridParams), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(d.wall, 0, V * sizeof(float4)));
    HIP_TRY(hipMemset(d.corr0, 0, S * sizeof(float4)));
    HIP_TRY(hipMemset(d.corr1, 0, S * sizeof(float4)));
    HIP_TRY(hipMemset(d.ratio, 0, V * sizeof(float)));
    HIP_TRY(hipMemset(d.ncount, 0, S * sizeof(uint32_t)));
    HIP_TRY(hipMemset(d.nbr_off, 0, S * sizeof(uint32_t)));
    HIP_TRY(hipMemset(d.red, 0, (size_t)d.nspec * 3 * d.red_slots * sizeof(uint32_t)));
    HIP_TRY(hipMemset(d.cell_reach, 0, d.C * sizeof(uint32_t)));
    HIP_TRY(hipMemset(d.cell_counts, 0, d.C * sizeof(uint32_t)));
    HIP_TRY(hipMemset(d.cell_offsets, 0, d.C * sizeof(uint32_t)));
    {
        std::vector<uint32_t> cell_spec(d.C);
        for (uint32_t sp = 0; sp < d.nspec; ++sp)
            std::fill(cell_spec.begin() + w->cbase[sp], cell_spec.begin() + w->cbase[sp + 1], sp);
        HIP_TRY(hipMemcpy(d.cell_spec, cell_spec.data(), d.C * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    // no slot holds a segment until the first rebuild
    HIP_TRY(hipMemset(d.cell_segs, 0xff, S * sizeof(uint32_t)));
    HIP_TRY(hipMemset(d.slotC, 0xff, S * sizeof(uint4)));
    std::vector<SpecState> spec(d.nspec);
    for (SpecState &sp : spec) {
        std::memset(&sp, 0, sizeof(sp));
        sp.metrics[0] = std::numeric_limits<float>::infinity();
        sp.metrics[2] = std::numeric_limits<float>::infinity();
        sp.rebuild = 1;  // lists start stale so the first contact pass builds them (world.rs:510-511)
    }
    HIP_TRY(hipMemcpy(d.ss, spec.data(), d.nspec * sizeof(SpecState), hipMemcpyHostToDevice));
    DevState init;
    std::memset(&init, 0, sizeof(init));
    init.nstate[0] = 1;  // lists start stale so the first contact pass builds them (world.rs:510-511)
    HIP_TRY(hipMemcpy(d.st, &init, sizeof(init), hipMemcpyHostToDevice));
    return TGH_OK;
}

// Creates a world from specimens already in the f32 layout (pack_assembly or
// pack_packed); every input check has run before the first HIP call.
int create_world_packed(std::vector<Packed> &&packed, const tgh_config *c, tgh_world **out) {
    int rc = validate_config(c);
    if (rc) return rc;
    const uint32_t count = (uint32_t)packed.size();
    if (count == 0) return set_error(TGH_ERR_INVALID, "no assembly");
    if (packed.size() > (size_t(1) << (32 - tgh::kSpecShift))) return set_error(TGH_ERR_INVALID, "too many specimens");
    tgh_world *w = new tgh_world();
    auto fail = [&](int code) {
        delete w;
        return code;
    };
    w->specimens = std::move(packed);
    w->sizes.resize(count);
    w->vbase.assign(count + 1, 0);
    w->fbase.assign(count + 1, 0);
    w->sbase.assign(count + 1, 0);
    w->cbase.assign(count + 1, 0);
    uint64_t V = 0, F = 0, S = 0, C = 0;
    for (uint32_t sp = 0; sp < count; ++sp) {
        rc = size_grid(w->specimens[sp], c, w->sizes[sp]);
        if (rc) {
            if (count > 1) set_error(rc, "specimen " + std::to_string(sp) + ": " + tgh_last_error());
            return fail(rc);
        }
        V += w->specimens[sp].V;
        F += w->specimens[sp].F;
        S += w->specimens[sp].S;
        C += w->sizes[sp].C;
        if (V > 0xffffffffull || C > (uint64_t(1) << 28))
            return fail(set_error(TGH_ERR_SIZE, "the specimens exceed 2^32 vertices or 2^28 cells"));
        w->vbase[sp + 1] = (uint32_t)V;
        w->fbase[sp + 1] = (uint32_t)F;
        w->sbase[sp + 1] = (uint32_t)S;
        w->cbase[sp + 1] = (uint32_t)C;
    }
    const char *env = std::getenv("TANGLE_HIP_GRAPHS");
    if (env && env[0] == '0') w->use_graphs = false;
    hipError_t e = hipGetDevice(&w->device);
    if (e != hipSuccess) return fail(set_error(TGH_ERR_DEVICE, std::string("no HIP device: ") + hipGetErrorString(e)));
    e = hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(set_error(TGH_ERR_DEVICE, hipGetErrorString(e)));
    e = hipHostMalloc((void **)&w->host_spec, count * sizeof(SpecState), hipHostMallocDefault);
    if (e != hipSuccess) return fail(set_error(TGH_ERR_DEVICE, hipGetErrorString(e)));
    WorldDev &d = w->d;
    d.F = (uint32_t)F;
    d.V = (uint32_t)V;
    d.S = (uint32_t)S;
    d.C = (uint32_t)C;
    d.nspec = count;
    d.red_slots = count == 1 ? tgh::kRedSlots : tgh::kRedSlotsBatch;
    // one list width for all specimens: the narrowest any specimen's budget allows
    d.block = w->sizes[0].block;
    for (const Sizing &z : w->sizes) d.block = std::min(d.block, z.block);
    d.ovf_cap = std::max<uint32_t>(tgh::kOverflowListCapacity, d.block);
    d.max_fiber_vertices = 0;
    d.track_walls = 0;
    w->grids.resize(count);
    for (uint32_t sp = 0; sp < count; ++sp) {
        const Packed &p = w->specimens[sp];
        d.max_fiber_vertices = std::max(d.max_fiber_vertices, *std::max_element(p.fcount.begin(), p.fcount.end()));
        if (!(p.periodic[0] && p.periodic[1] && p.periodic[2])) d.track_walls = 1u;
        GridParams &g = w->grids[sp];
        std::memset(&g, 0, sizeof(g));
        for (int k = 0; k < 3; ++k) {
            g.lower[k] = p.lower[k];
            g.upper[k] = p.upper[k];
            g.cells[k] = w->sizes[sp].cells[k];
            g.periodic[k] = p.periodic[k];
            g.width[k] = (p.upper[k] - p.lower[k]) / (float)w->sizes[sp].cells[k];
            g.length[k] = p.upper[k] - p.lower[k];
        }
        g.cell_base = w->cbase[sp];
        g.skin = w->sizes[sp].skin;
        g.seg_base = w->sbase[sp];
        g.seg_count = p.S;
    }
    // lanes per fiber in the curvature sweep: 8 (eight fibers per wave, the
    // fastest measured on the bimodal felt) unless a workgroup's LDS would
    // exceed 64 KiB, then wider groups
    d.gs_lanes = 0u;
    for (uint32_t lanes : {8u, 16u, 32u, 64u})
        if (d.max_fiber_vertices <= tgh::kMaxLdsFiberVertices &&
            tgh::curvature_lds_bytes(d.max_fiber_vertices, lanes, d.track_walls) <= 64u * 1024u) {
            d.gs_lanes = lanes;
            break;
        }
    // lanes per segment in the contact pass, one candidate each: 8 lanes
    // issuing three rounds of gathers at once cover the ~20-candidate lists
    // of the bimodal felt and the slab (faster there than 16, which leaves
    // more lanes idle); the autotuner picks 4 for shorter lists
    d.contact_lanes = 8u;
    // the list build leaves out stencil cells beyond reach when cells are
    // wide against a typical pair's reach (the bimodal felt's are 1.9 times
    // it, their width set by its longest segments): where they are about as
    // wide, as in the batched specimens, few cells drop out and the bound's
    // registers cost more than it saves
    d.prune_lists = 1u;
    for (uint32_t sp = 0; sp < count; ++sp)
        if (!(w->sizes[sp].cell_size >= 1.25f * w->sizes[sp].list_reach)) d.prune_lists = 0u;
    e = hipSuccess;
    GridParams *grids = nullptr;
    if (e == hipSuccess) e = w->alloc(&d.st, 1);
    if (e == hipSuccess) e = w->alloc(&d.ss, count);
    if (e == hipSuccess) e = w->alloc(&grids, count);
    if (e == hipSuccess) e = w->alloc(&d.red, (size_t)count * 3 * d.red_slots);
    if (e == hipSuccess) e = w->alloc(&d.ovf, std::max<size_t>(1, (size_t)d.ovf_cap * S));
    if (e == hipSuccess) e = w->alloc(&d.pos, V);
    if (e == hipSuccess) e = w->alloc(&d.alt, V);
    if (e == hipSuccess) e = w->alloc(&d.refpos, V);
    if (e == hipSuccess) e = w->alloc(&d.wall, V);
    if (e == hipSuccess) e = w->alloc(&d.vA, V);
    if (e == hipSuccess) e = w->alloc(&d.vB, V);
    if (e == hipSuccess) e = w->alloc(&d.sstat, S);
    if (e == hipSuccess) e = w->alloc(&d.fspan, F);
    if (e == hipSuccess) e = w->alloc(&d.fstep, F);
    if (e == hipSuccess) e = w->alloc(&d.flayer, F);
    if (e == hipSuccess) e = w->alloc(&d.factive, F);
    if (e == hipSuccess) e = w->alloc(&w->formation.error, 2);
    if (e == hipSuccess) e = w->alloc(&d.cell_counts, C);
    if (e == hipSuccess) e = w->alloc(&d.cell_offsets, C);
    if (e == hipSuccess) e = w->alloc(&d.cell_cursor, C);
    if (e == hipSuccess) e = w->alloc(&d.cell_spec, std::max<uint64_t>(1, C));
    if (e == hipSuccess) e = w->alloc(&d.cell_reach, std::max<uint64_t>(1, C));
    if (e == hipSuccess) e = w->alloc(&d.cell_segs, S);
    if (e == hipSuccess) e = w->alloc(&d.seg_cell, S);
    if (e == hipSuccess) e = w->alloc(&d.scattered, S);
    if (e == hipSuccess) e = w->alloc(&d.slotA, S);
    if (e == hipSuccess) e = w->alloc(&d.slotB, S);
    if (e == hipSuccess) e = w->alloc(&d.slotC, S);
    if (e == hipSuccess) e = w->alloc(&d.cP, S);
    if (e == hipSuccess) e = w->alloc(&d.cQ, S);
    if (e == hipSuccess) e = w->alloc(&d.ncount, S);
    if (e == hipSuccess) e = w->alloc(&d.nbr, std::max<size_t>(1, (size_t)d.block * S));
    if (e == hipSuccess) e = w->alloc(&d.nbr_off, S);
    if (e == hipSuccess) e = w->alloc(&d.corr0, S);
    if (e == hipSuccess) e = w->alloc(&d.corr1, S);
    if (e == hipSuccess) e = w->alloc(&d.ratio, V);
    if (e != hipSuccess) return fail(set_error(TGH_ERR_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e)));
    d.grids = grids;
    rc = upload_static(w);
    if (rc) return fail(rc);
    *out = w;
    return TGH_OK;
}

int create_world(const tgh_assembly *a, uint32_t count, const tgh_config *c, tgh_world **out) {
    int rc = validate_config(c);
    if (rc) return rc;
    if (!a || count == 0) return set_error(TGH_ERR_INVALID, "no assembly");
    if (count > (1u << (32 - tgh::kSpecShift))) return set_error(TGH_ERR_INVALID, "too many specimens");
    std::vector<Packed> packed(count);
    for (uint32_t sp = 0; sp < count; ++sp) {
        rc = pack_assembly(&a[sp], packed[sp]);
        if (rc) {
            if (count > 1) set_error(rc, "specimen " + std::to_string(sp) + ": " + tgh_last_error());
            return rc;
        }
    }
    return create_world_packed(std::move(packed), c, out);
}

std::string graph_key(const IterParams &q, bool timing) {
    std::string k(sizeof(q) + 1, '\0');
    std::memcpy(&k[0], &q, sizeof(q));
    k[sizeof(q)] = timing ? 1 : 0;
    return k;
}

hipError_t ensure_events(std::vector<hipEvent_t> &events, size_t count) {
    while (events.size() < count) {
        hipEvent_t e = nullptr;
        hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) return r;
        events.push_back(e);
    }
    return hipSuccess;
}

// Enqueues begin_relaxation_batch and iterations+1 passes (world.rs:1020-1505).
// With `events`, event i*G+g marks the start of group g of pass i.
int enqueue_batch(tgh_world *w, const IterParams &q, std::vector<hipEvent_t> *events) {
    constexpr int G = tgh::kGroupCount;
    const WorldDev &d = w->d;
    hipStream_t s = w->stream;
    hipError_t e = tgh::launch_begin_batch(d, s);
    // begin_relaxation_batch is followed by the displacement check
    // (world.rs:1033-1035); between batches nothing moves the positions
    // except tgh_world_upload_positions, which requests a rebuild itself.
    if (events && e == hipSuccess) e = ensure_events(*events, (size_t)(q.iterations + 1) * G + 1);
    auto mark = [&](size_t i) {
        if (events && e == hipSuccess) e = hipEventRecord((*events)[i], s);
    };
    for (uint32_t step = 0; e == hipSuccess && step <= q.iterations; ++step) {
        const size_t base = (size_t)step * G;
        mark(base + tgh::kGroupRebuild);
        if (e == hipSuccess) e = tgh::launch_rebuild(d, s);
        mark(base + tgh::kGroupContact);
        if (e == hipSuccess) e = tgh::launch_contact(d, q, s);
        mark(base + tgh::kGroupMeasure);
        if (e == hipSuccess) e = tgh::launch_measure(d, q, s);
        mark(base + tgh::kGroupVertex);
        if (e == hipSuccess) e = tgh::launch_vertex_passes(d, q, s);
        mark(base + tgh::kGroupCurvature);
        if (e == hipSuccess) e = tgh::launch_curvature(d, q, s);
        // manufacturing targets end each iteration at the commanded position;
        // the extra pass only measures (world.rs:1420-1482)
        if (step < q.iterations) {
            if (e == hipSuccess) e = tgh::launch_layer_targets(d, w->formation, true, s);
            if (e == hipSuccess) e = tgh::launch_vertex_targets(d, w->formation, s);
        }
    }
    mark((size_t)(q.iterations + 1) * G);
    if (e != hipSuccess) return set_error(TGH_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
    return TGH_OK;
}

int accumulate_timing(tgh_world *w, const IterParams &q, const std::vector<hipEvent_t> &events) {
    constexpr int G = tgh::kGroupCount;
    const size_t marks = (size_t)(q.iterations + 1) * G + 1;
    for (size_t i = 0; i + 1 < marks; ++i) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, events[i], events[i + 1]));
        w->group_ms[i % G] += ms;
        w->group_launches[i % G] += 1;
    }
    return TGH_OK;
}

// BatchStatus of one specimen (read_status, world.rs:2337-2369).
void specimen_status(const SpecState &s, const tgh_config &c, tgh_batch_status &st) {
    st.total_iterations = s.control[1];
    st.batch_iterations = s.control[2];
    st.max_penetration = s.metrics[0];
    st.max_displacement = s.metrics[1];
    st.max_curvature_ratio = s.metrics[2];
    const bool overflow = s.control[3] != 0;
    st.cell_list_overflow = overflow ? 1u : 0u;
    st.converged = (!c.force_full_iterations && s.metrics[0] <= c.penetration_tolerance &&
                    s.metrics[2] <= 1.0f + c.curvature_ratio_tolerance && !overflow)
                       ? 1u
                       : 0u;
}

int run_batch(tgh_world *w, const tgh_config *c, uint32_t iterations, tgh_batch_status *st) {
    int rc = validate_config(c);
    if (rc) return rc;
    if (iterations == 0) return set_error(TGH_ERR_INVALID, "iterations must be positive");
    if (c->neighbor_capacity != 0 && c->neighbor_skin_scale != c->neighbor_skin_scale)
        return set_error(TGH_ERR_INVALID, "neighbor_skin_scale is NaN");
    const IterParams q = iter_params(c, iterations);
    std::vector<hipEvent_t> *events = nullptr;
    // Event records captured into a graph cannot be queried for elapsed time
    // on this runtime, so a timed batch is launched directly.
    if (w->use_graphs && !w->timing) {
        const std::string key = graph_key(q, false);
        auto it = w->graphs.find(key);
        if (it == w->graphs.end()) {
            HIP_TRY(hipStreamBeginCapture(w->stream, hipStreamCaptureModeThreadLocal));
            int erc = enqueue_batch(w, q, nullptr);
            hipGraph_t graph = nullptr;
            hipError_t e = hipStreamEndCapture(w->stream, &graph);
            if (erc) return erc;
            if (e != hipSuccess) return set_error(TGH_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
            hipGraphExec_t exec = nullptr;
            e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGraphDestroy(graph);
                return set_error(TGH_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
            }
            w->graph_templates.push_back(graph);
            it = w->graphs.emplace(key, exec).first;
        }
        HIP_TRY(hipGraphLaunch(it->second, w->stream));
    } else {
        if (w->timing) events = &w->direct_events;
        rc = enqueue_batch(w, q, events);
        if (rc) return rc;
    }
    HIP_TRY(hipMemcpyAsync(w->host_spec, w->d.ss, w->d.nspec * sizeof(SpecState), hipMemcpyDeviceToHost, w->stream));
    HIP_TRY(hipStreamSynchronize(w->stream));
    if (events) {
        rc = accumulate_timing(w, q, *events);
        if (rc) return rc;
    }
    w->last_config = *c;
    w->ran = true;
    // one specimen: its status; several: the worst of each field, converged
    // only when every specimen converged
    for (uint32_t sp = 0; sp < w->d.nspec; ++sp) {
        tgh_batch_status one;
        specimen_status(w->host_spec[sp], *c, one);
        if (sp == 0) {
            *st = one;
            continue;
        }
        st->total_iterations = std::max(st->total_iterations, one.total_iterations);
        st->batch_iterations = std::max(st->batch_iterations, one.batch_iterations);
        st->max_penetration = std::fmax(st->max_penetration, one.max_penetration);
        st->max_displacement = std::fmax(st->max_displacement, one.max_displacement);
        st->max_curvature_ratio = std::fmax(st->max_curvature_ratio, one.max_curvature_ratio);
        st->converged = st->converged && one.converged;
        st->cell_list_overflow = st->cell_list_overflow || one.cell_list_overflow;
    }
    return TGH_OK;
}

}  // namespace

extern "C" {

void tgh_config_default(tgh_config *c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->penetration_tolerance = 1.0e-4f;
    c->correction_fraction = 0.8f;
    c->stretch_stiffness = 0.35f;
    c->bend_stiffness = 0.03f;
    c->curvature_limit_stiffness = 1.0f;
    c->curvature_limit_safety_margin = 1.0e-3f;
    c->curvature_ratio_tolerance = 1.0e-5f;
    c->max_step = 0.006f;
    c->cell_size_scale = 1.0f;
    c->neighbor_skin_scale = 2.0f;
    c->contact_aggregation = 0;
    c->constraint_iterations = 2;
    c->curvature_cleanup_sweeps = 4;
    c->force_full_iterations = 0;
    c->neighbor_capacity = 48;
    c->maximum_neighbor_list_bytes = 0;
}

int tgh_world_create(const tgh_assembly *assembly, const tgh_config *config, tgh_world **out) {
    if (!out) return set_error(TGH_ERR_INVALID, "output pointer is null");
    *out = nullptr;
    return create_world(assembly, 1, config, out);
}

int tgh_world_create_packed(const tgh_packed *packed, const tgh_config *config, tgh_world **out) {
    if (!out) return set_error(TGH_ERR_INVALID, "output pointer is null");
    *out = nullptr;
    std::vector<Packed> specimens(1);
    int rc = pack_packed(packed, specimens[0]);
    if (rc) return rc;
    return create_world_packed(std::move(specimens), config, out);
}

int tgh_world_create_batch(const tgh_assembly *specimens, uint32_t count, const tgh_config *config, tgh_world **out) {
    if (!out) return set_error(TGH_ERR_INVALID, "output pointer is null");
    *out = nullptr;
    return create_world(specimens, count, config, out);
}

int tgh_world_specimen_status(const tgh_world *world, tgh_batch_status *out, size_t count) {
    if (!world || !out) return set_error(TGH_ERR_INVALID, "null argument");
    if (count < world->d.nspec) return set_error(TGH_ERR_SIZE, "status buffer holds fewer than specimen_count entries");
    if (!world->ran) return set_error(TGH_ERR_INVALID, "no batch has run yet");
    for (uint32_t sp = 0; sp < world->d.nspec; ++sp) specimen_status(world->host_spec[sp], world->last_config, out[sp]);
    return TGH_OK;
}

int tgh_world_run_batch(tgh_world *world, const tgh_config *config, uint32_t iterations, tgh_batch_status *status) {
    if (!world || !status) return set_error(TGH_ERR_INVALID, "world or status is null");
    return run_batch(world, config, iterations, status);
}

int tgh_world_download_positions(const tgh_world *world, float *out, size_t count) {
    if (!world || !out) return set_error(TGH_ERR_INVALID, "null argument");
    const size_t V = world->d.V;
    if (count < 3 * V) return set_error(TGH_ERR_SIZE, "position buffer holds fewer than 3*vertex_count floats");
    std::vector<float4> tmp(V);
    HIP_TRY(hipStreamSynchronize(world->stream));
    HIP_TRY(hipMemcpy(tmp.data(), world->d.pos, V * sizeof(float4), hipMemcpyDeviceToHost));
    for (size_t v = 0; v < V; ++v) {
        out[3 * v] = tmp[v].x;
        out[3 * v + 1] = tmp[v].y;
        out[3 * v + 2] = tmp[v].z;
    }
    return TGH_OK;
}

int tgh_world_upload_positions(tgh_world *world, const float *positions, size_t count) {
    if (!world || !positions) return set_error(TGH_ERR_INVALID, "null argument");
    const size_t V = world->d.V;
    if (count != 3 * V) return set_error(TGH_ERR_SIZE, "expected 3*vertex_count floats");
    std::vector<float4> tmp(V);
    HIP_TRY(hipStreamSynchronize(world->stream));
    HIP_TRY(hipMemcpy(tmp.data(), world->d.pos, V * sizeof(float4), hipMemcpyDeviceToHost));
    for (size_t v = 0; v < V; ++v) {
        tmp[v].x = positions[3 * v];
        tmp[v].y = positions[3 * v + 1];
        tmp[v].z = positions[3 * v + 2];
    }
    HIP_TRY(hipMemcpy(world->d.pos, tmp.data(), V * sizeof(float4), hipMemcpyHostToDevice));
    world->history_valid = false;  // the motion history described the replaced positions
    world->proposal_pending = false;
    HIP_TRY(tgh::launch_request_rebuild(world->d, world->stream));
    HIP_TRY(hipStreamSynchronize(world->stream));
    return TGH_OK;
}

int tgh_world_info_get(const tgh_world *world, tgh_world_info *info) {
    if (!world || !info) return set_error(TGH_ERR_INVALID, "null argument");
    std::memset(info, 0, sizeof(*info));
    info->fiber_count = world->d.F;
    info->segment_count = world->d.S;
    info->vertex_count = world->d.V;
    for (int k = 0; k < 3; ++k) info->cells[k] = world->grids[0].cells[k];
    info->cell_count = world->d.C;
    info->cell_size = world->sizes[0].cell_size;
    info->neighbor_skin = world->grids[0].skin;
    info->specimen_count = world->d.nspec;
    info->neighbor_block = world->d.block;
    DevState s;
    HIP_TRY(hipStreamSynchronize(world->stream));
    HIP_TRY(hipMemcpy(&s, world->d.st, sizeof(s), hipMemcpyDeviceToHost));
    info->neighbor_list_rebuilds = s.nstate[1];
    info->device_bytes = world->device_bytes;
    return TGH_OK;
}

int tgh_world_cell_list(const tgh_world *world, uint32_t *counts, uint32_t *offsets, uint32_t *segments,
                        size_t cell_capacity, size_t segment_capacity) {
    if (!world || !counts || !offsets || !segments) return set_error(TGH_ERR_INVALID, "null argument");
    if (cell_capacity < world->d.C || segment_capacity < world->d.S)
        return set_error(TGH_ERR_SIZE, "cell-list buffers too small");
    HIP_TRY(hipStreamSynchronize(world->stream));
    HIP_TRY(hipMemcpy(counts, world->d.cell_counts, world->d.C * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(offsets, world->d.cell_offsets, world->d.C * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(segments, world->d.cell_segs, (size_t)world->d.S * 4, hipMemcpyDeviceToHost));
    return TGH_OK;
}

int tgh_world_neighbor_lists(const tgh_world *world, uint32_t *counts, uint32_t *lists, size_t list_capacity) {
    if (!world || !counts || !lists) return set_error(TGH_ERR_INVALID, "null argument");
    const size_t S = world->d.S, B = world->d.block;
    if (list_capacity < S * B) return set_error(TGH_ERR_SIZE, "list buffer smaller than segment_count*neighbor_block");
    std::vector<uint32_t> slot_seg(S), slot_count(S), slot_off(S), cols(std::max<size_t>(1, S * B));
    HIP_TRY(hipStreamSynchronize(world->stream));
    HIP_TRY(hipMemcpy(slot_off.data(), world->d.nbr_off, S * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(slot_seg.data(), world->d.cell_segs, S * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(slot_count.data(), world->d.ncount, S * 4, hipMemcpyDeviceToHost));
    if (S * B) HIP_TRY(hipMemcpy(cols.data(), world->d.nbr, S * B * 4, hipMemcpyDeviceToHost));
    std::memset(lists, 0, S * B * 4);
    std::memset(counts, 0, S * 4);  // segments outside the cell list (inactive) have no list
    for (size_t slot = 0; slot < S; ++slot) {
        const uint32_t seg = slot_seg[slot];
        if (seg >= S) continue;  // no segment in this slot
        counts[seg] = slot_count[slot];
        // rows hold cell-list slots; the boundary returns segments. An
        // overflowed list's row is never written (unspecified), so only
        // listed rows are mapped.
        if (slot_count[slot] > B) continue;
        if ((size_t)slot_off[slot] + slot_count[slot] > S * B) continue;  // not built
        for (size_t k = 0; k < slot_count[slot]; ++k) {
            const uint32_t other = cols[slot_off[slot] + k];
            lists[(size_t)seg * B + k] = other < S ? slot_seg[other] : 0xffffffffu;
        }
    }
    return TGH_OK;
}

int tgh_world_set_timing(tgh_world *world, uint32_t enable) {
    if (!world) return set_error(TGH_ERR_INVALID, "world is null");
    world->timing = enable != 0;
    return TGH_OK;
}

int tgh_world_timing(tgh_world *world, double *group_ms, uint64_t *group_launches, size_t groups, uint32_t reset) {
    if (!world || !group_ms || !group_launches) return set_error(TGH_ERR_INVALID, "null argument");
    if (groups < (size_t)tgh::kGroupCount) return set_error(TGH_ERR_SIZE, "timing buffers hold fewer than 5 groups");
    for (int g = 0; g < tgh::kGroupCount; ++g) {
        group_ms[g] = world->group_ms[g];
        group_launches[g] = world->group_launches[g];
        if (reset) {
            world->group_ms[g] = 0.0;
            world->group_launches[g] = 0;
        }
    }
    return TGH_OK;
}

// Captured batch graphs hold the formation launches of their time.
static void drop_graphs(tgh_world *w) {
    (void)hipStreamSynchronize(w->stream);
    for (auto &kv : w->graphs) (void)hipGraphExecDestroy(kv.second);
    for (hipGraph_t g : w->graph_templates) (void)hipGraphDestroy(g);
    w->graphs.clear();
    w->graph_templates.clear();
}

int tgh_world_activate_formation_step(tgh_world *world, uint32_t maximum_step, uint64_t *active_segments,
                                      uint64_t *active_vertices) {
    if (!world) return set_error(TGH_ERR_INVALID, "world is null");
    HIP_TRY(tgh::launch_activate(world->d, maximum_step, world->stream));
    // newly active fibers are missing from the lists (world.rs:2176-2178)
    HIP_TRY(tgh::launch_request_rebuild(world->d, world->stream));
    HIP_TRY(hipStreamSynchronize(world->stream));
    uint64_t segs = 0, verts = 0;
    for (const Packed &p : world->specimens)
        for (uint32_t f = 0; f < p.F; ++f)
            if (p.fstep[f] <= maximum_step) {
                segs += p.fcount[f] - 1;
                verts += p.fcount[f];
            }
    if (active_segments) *active_segments = segs;
    if (active_vertices) *active_vertices = verts;
    return TGH_OK;
}

// Holds `targets` as the world's layer target; with apply_now the first
// correction is applied at once (apply_layer_target_range, world.rs:1752-1803),
// without it the target is only reinstated (restore, world.rs:721-733).
static int install_layer_targets(tgh_world *world, uint32_t axis, const float *targets, size_t count,
                                 uint32_t first_layer, uint32_t last_layer, float stiffness, float max_translation,
                                 bool apply_now) {
    if (!world || !targets) return set_error(TGH_ERR_INVALID, "null argument");
    if (axis > 2 || count == 0 || count > 0xffffffffu) return set_error(TGH_ERR_INVALID, "axis must be 0-2 and targets non-empty");
    if (!(stiffness > 0.f && stiffness <= 1.f) || !(max_translation > 0.f))
        return set_error(TGH_ERR_INVALID, "stiffness must be in (0, 1] and max_translation positive");
    if (first_layer > last_layer) return set_error(TGH_ERR_INVALID, "first_layer after last_layer");
    for (size_t i = 0; i < count; ++i)
        if (!std::isfinite(targets[i])) return set_error(TGH_ERR_INVALID, "layer targets must be finite");
    tgh::FormationDev &f = world->formation;
    if (count > world->formation_layer_capacity) {
        HIP_TRY(world->alloc(&f.layer_targets, count));
        world->formation_layer_capacity = (uint32_t)count;
    }
    HIP_TRY(hipStreamSynchronize(world->stream));
    HIP_TRY(hipMemcpy(f.layer_targets, targets, count * sizeof(float), hipMemcpyHostToDevice));
    f.layer_count = (uint32_t)count;
    f.layer_axis = axis;
    f.first_layer = first_layer;
    f.last_layer = last_layer;
    f.layer_stiffness = stiffness;
    f.layer_max_translation = max_translation;
    if (apply_now) {
        HIP_TRY(tgh::launch_layer_targets(world->d, f, false, world->stream));
        HIP_TRY(hipStreamSynchronize(world->stream));
    }
    drop_graphs(world);
    return TGH_OK;
}

// Holds vertex targets; `commands` null: each command starts at the
// vertex's coordinate and the target lies `displacement` from it
// (initialize_vertex_displacement_targets); otherwise targets and commands
// are reinstated as given (restore).
static int install_vertex_targets(tgh_world *world, uint32_t axis, const uint32_t *vertices, size_t count,
                                  float displacement, const float *targets, const float *commands, float stiffness,
                                  float max_translation) {
    if (!world || !vertices) return set_error(TGH_ERR_INVALID, "null argument");
    if (axis > 2 || count == 0 || count > 0xffffffffu) return set_error(TGH_ERR_INVALID, "axis must be 0-2 and vertices non-empty");
    for (size_t i = 0; i < count; ++i)
        if (vertices[i] >= world->d.V) return set_error(TGH_ERR_INVALID, "vertex index out of range");
    if (!commands && (!std::isfinite(displacement) || displacement == 0.f))
        return set_error(TGH_ERR_INVALID, "displacement must be finite and nonzero");
    if (!(stiffness > 0.f && stiffness <= 1.f) || !(max_translation > 0.f))
        return set_error(TGH_ERR_INVALID, "stiffness must be in (0, 1] and max_translation positive");
    tgh::FormationDev &f = world->formation;
    if (count > world->formation_vertex_capacity) {
        HIP_TRY(world->alloc(&f.vertex_index, count));
        HIP_TRY(world->alloc(&f.vertex_target, count));
        HIP_TRY(world->alloc(&f.vertex_command, count));
        world->formation_vertex_capacity = (uint32_t)count;
    }
    HIP_TRY(hipStreamSynchronize(world->stream));
    HIP_TRY(hipMemcpy(f.vertex_index, vertices, count * sizeof(uint32_t), hipMemcpyHostToDevice));
    f.vertex_count = (uint32_t)count;
    f.vertex_axis = axis;
    f.vertex_stiffness = stiffness;
    f.vertex_max_translation = max_translation;
    if (commands) {
        HIP_TRY(hipMemcpy(f.vertex_target, targets, count * sizeof(float), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(f.vertex_command, commands, count * sizeof(float), hipMemcpyHostToDevice));
    } else {
        HIP_TRY(tgh::launch_init_vertex_targets(world->d, f, displacement, world->stream));
        HIP_TRY(hipStreamSynchronize(world->stream));
    }
    drop_graphs(world);
    return TGH_OK;
}

int tgh_world_set_layer_targets(tgh_world *world, uint32_t axis, const float *targets, size_t count,
                                uint32_t first_layer, uint32_t last_layer, float stiffness, float max_translation) {
    return install_layer_targets(world, axis, targets, count, first_layer, last_layer, stiffness, max_translation,
                                 true);
}

int tgh_world_clear_layer_targets(tgh_world *world) {
    if (!world) return set_error(TGH_ERR_INVALID, "world is null");
    world->formation.layer_count = 0;
    drop_graphs(world);
    return TGH_OK;
}

int tgh_world_set_vertex_targets(tgh_world *world, uint32_t axis, const uint32_t *vertices, size_t count,
                                 float displacement, float stiffness, float max_translation) {
    return install_vertex_targets(world, axis, vertices, count, displacement, nullptr, nullptr, stiffness,
                                  max_translation);
}

int tgh_world_clear_vertex_targets(tgh_world *world) {
    if (!world) return set_error(TGH_ERR_INVALID, "world is null");
    world->formation.vertex_count = 0;
    drop_graphs(world);
    return TGH_OK;
}

int tgh_world_formation_target_error(tgh_world *world, float *layer_error, float *vertex_error,
                                     uint32_t *has_layer, uint32_t *has_vertex) {
    if (!world || !layer_error || !vertex_error || !has_layer || !has_vertex)
        return set_error(TGH_ERR_INVALID, "null argument");
    const tgh::FormationDev &f = world->formation;
    *has_layer = f.layer_count ? 1u : 0u;
    *has_vertex = f.vertex_count ? 1u : 0u;
    *layer_error = 0.f;
    *vertex_error = 0.f;
    if (!f.layer_count && !f.vertex_count) return TGH_OK;
    float err[2];
    HIP_TRY(tgh::launch_target_error(world->d, f, world->stream));
    HIP_TRY(hipMemcpyAsync(err, f.error, sizeof(err), hipMemcpyDeviceToHost, world->stream));
    HIP_TRY(hipStreamSynchronize(world->stream));
    *layer_error = err[0];
    *vertex_error = err[1];
    return TGH_OK;
}

namespace {

// Checks a new cell for a one-specimen world (compact_cell's assertions,
// world.rs:1937-1948) and moves the grid to it: same cell size, the cell
// counts of the new extent, every list stale.
int set_cell(tgh_world *w, const float *lower, const float *upper) {
    if (w->d.nspec != 1) return set_error(TGH_ERR_INVALID, "cell changes need a one-specimen world");
    GridParams g = w->grids[0];
    const tgh::Sizing &z = w->sizes[0];
    uint64_t C = 1;
    uint32_t cells[3];
    for (int k = 0; k < 3; ++k) {
        if (!std::isfinite(lower[k]) || !std::isfinite(upper[k]))
            return set_error(TGH_ERR_INVALID, "cell bounds must be finite");
        if (!(upper[k] - lower[k] > 2.0f * z.max_radius))
            return set_error(TGH_ERR_INVALID, "the cell must be wider than two fiber radii on every axis");
        if (!(lower[k] >= g.lower[k]) || !(upper[k] <= g.upper[k]))
            return set_error(TGH_ERR_INVALID, "the new cell must lie inside the current one");
        cells[k] = to_u32_saturating(std::fmax(std::floor((upper[k] - lower[k]) / z.cell_size), 1.0f));
        C *= cells[k];
    }
    if (C > w->d.C) return set_error(TGH_ERR_SIZE, "the new cell needs more grid cells than were allocated");
    for (int k = 0; k < 3; ++k) {
        g.lower[k] = lower[k];
        g.upper[k] = upper[k];
        g.cells[k] = cells[k];
        g.width[k] = (upper[k] - lower[k]) / (float)cells[k];
        g.length[k] = upper[k] - lower[k];
    }
    HIP_TRY(hipMemcpyAsync((void *)w->d.grids, &g, sizeof(g), hipMemcpyHostToDevice, w->stream));
    HIP_TRY(hipStreamSynchronize(w->stream));
    w->grids[0] = g;
    tgh::Packed &p = w->specimens[0];
    for (int k = 0; k < 3; ++k) {
        p.lower[k] = lower[k];
        p.upper[k] = upper[k];
        w->sizes[0].cells[k] = cells[k];
    }
    HIP_TRY(tgh::launch_request_rebuild(w->d, w->stream));
    HIP_TRY(hipStreamSynchronize(w->stream));
    return TGH_OK;
}

}  // namespace

int tgh_world_checkpoint_sizes(tgh_world *world, tgh_checkpoint *out) {
    if (!world || !out) return set_error(TGH_ERR_INVALID, "null argument");
    if (world->d.nspec != 1) return set_error(TGH_ERR_INVALID, "checkpoints hold one-specimen worlds");
    out->fiber_count = world->d.F;
    out->vertex_count = world->d.V;
    out->layer_target_count = world->formation.layer_count;
    out->vertex_target_count = world->formation.vertex_count;
    return TGH_OK;
}

int tgh_world_checkpoint(tgh_world *world, tgh_checkpoint *out) {
    if (!world || !out) return set_error(TGH_ERR_INVALID, "null argument");
    if (world->d.nspec != 1) return set_error(TGH_ERR_INVALID, "checkpoints hold one-specimen worlds");
    const tgh::FormationDev &f = world->formation;
    if (out->fiber_count != world->d.F || out->vertex_count != world->d.V ||
        out->layer_target_count != f.layer_count || out->vertex_target_count != f.vertex_count)
        return set_error(TGH_ERR_SIZE, "checkpoint counts differ from tgh_world_checkpoint_sizes");
    if (!out->positions || !out->fiber_active || !out->wall_reactions ||
        (f.layer_count && !out->layer_targets) ||
        (f.vertex_count && (!out->vertex_indices || !out->vertex_targets || !out->vertex_commands)))
        return set_error(TGH_ERR_INVALID, "checkpoint array is null");
    int rc = tgh_world_download_positions(world, out->positions, 3 * (size_t)world->d.V);
    if (rc) return rc;
    const size_t V = world->d.V;
    HIP_TRY(hipMemcpy(out->fiber_active, world->d.factive, world->d.F * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (world->d.track_walls) {
        std::vector<float4> w(V);
        HIP_TRY(hipMemcpy(w.data(), world->d.wall, V * sizeof(float4), hipMemcpyDeviceToHost));
        for (size_t v = 0; v < V; ++v) {
            out->wall_reactions[3 * v] = w[v].x;
            out->wall_reactions[3 * v + 1] = w[v].y;
            out->wall_reactions[3 * v + 2] = w[v].z;
        }
    } else {
        std::fill(out->wall_reactions, out->wall_reactions + 3 * V, 0.f);
    }
    tgh::SpecState spec;
    HIP_TRY(hipMemcpy(&spec, world->d.ss, sizeof(spec), hipMemcpyDeviceToHost));
    out->total_iterations = spec.control[1];
    for (int k = 0; k < 3; ++k) {
        out->cell_lower[k] = world->grids[0].lower[k];
        out->cell_upper[k] = world->grids[0].upper[k];
    }
    out->layer_axis = f.layer_axis;
    out->layer_first = f.first_layer;
    out->layer_last = f.last_layer;
    out->layer_stiffness = f.layer_stiffness;
    out->layer_max_translation = f.layer_max_translation;
    if (f.layer_count)
        HIP_TRY(hipMemcpy(out->layer_targets, f.layer_targets, f.layer_count * sizeof(float), hipMemcpyDeviceToHost));
    out->vertex_axis = f.vertex_axis;
    out->vertex_stiffness = f.vertex_stiffness;
    out->vertex_max_translation = f.vertex_max_translation;
    if (f.vertex_count) {
        HIP_TRY(hipMemcpy(out->vertex_indices, f.vertex_index, f.vertex_count * sizeof(uint32_t),
                          hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(out->vertex_targets, f.vertex_target, f.vertex_count * sizeof(float),
                          hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(out->vertex_commands, f.vertex_command, f.vertex_count * sizeof(float),
                          hipMemcpyDeviceToHost));
    }
    return TGH_OK;
}

int tgh_world_restore(const tgh_assembly *assembly, const tgh_config *config, const tgh_checkpoint *ck,
                      tgh_world **out) {
    if (!out || !ck) return set_error(TGH_ERR_INVALID, "null argument");
    *out = nullptr;
    if (!ck->positions || !ck->fiber_active || !ck->wall_reactions)
        return set_error(TGH_ERR_INVALID, "checkpoint array is null");
    if (ck->total_iterations > 0xffffffffull)
        return set_error(TGH_ERR_INVALID, "checkpoint iteration count exceeds u32");
    tgh_world *w = nullptr;
    int rc = create_world(assembly, 1, config, &w);
    if (rc) return rc;
    auto fail = [&](int code) {
        delete w;
        return code;
    };
    if (ck->fiber_count != w->d.F || ck->vertex_count != w->d.V)
        return fail(set_error(TGH_ERR_SIZE, "the checkpoint was taken from another assembly"));
    {
        // a cell compacted since creation: the grid follows it, the positions
        // come from the checkpoint
        const bool keep = std::isnan(ck->cell_lower[0]);  // NaN: the assembly's cell
        bool same = true;
        for (int k = 0; k < 3; ++k)
            same = same && ck->cell_lower[k] == w->grids[0].lower[k] && ck->cell_upper[k] == w->grids[0].upper[k];
        if (!keep && !same) {
            rc = set_cell(w, ck->cell_lower, ck->cell_upper);
            if (rc) return fail(rc);
        }
    }
    rc = tgh_world_upload_positions(w, ck->positions, 3 * (size_t)w->d.V);
    if (rc) return fail(rc);
    for (uint32_t fi = 0; fi < w->d.F; ++fi)
        if (ck->fiber_active[fi] > 1u) return fail(set_error(TGH_ERR_INVALID, "fiber_active entries must be 0 or 1"));
    if (hipMemcpy(w->d.factive, ck->fiber_active, w->d.F * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess ||
        tgh::launch_apply_activity(w->d, w->stream) != hipSuccess || hipStreamSynchronize(w->stream) != hipSuccess)
        return fail(set_error(TGH_ERR_DEVICE, "restoring fiber activity failed"));
    if (w->d.track_walls) {
        const size_t V = w->d.V;
        std::vector<float4> wr(V);
        for (size_t v = 0; v < V; ++v)
            wr[v] = make_float4(ck->wall_reactions[3 * v], ck->wall_reactions[3 * v + 1],
                                ck->wall_reactions[3 * v + 2], 0.f);
        if (hipMemcpy(w->d.wall, wr.data(), V * sizeof(float4), hipMemcpyHostToDevice) != hipSuccess)
            return fail(set_error(TGH_ERR_DEVICE, "restoring wall reactions failed"));
    }
    tgh::SpecState spec;
    if (hipMemcpy(&spec, w->d.ss, sizeof(spec), hipMemcpyDeviceToHost) != hipSuccess)
        return fail(set_error(TGH_ERR_DEVICE, "reading the specimen state failed"));
    spec.control[0] = 0;
    spec.control[1] = (uint32_t)ck->total_iterations;
    spec.control[2] = 0;
    spec.control[3] = 0;
    if (hipMemcpy(w->d.ss, &spec, sizeof(spec), hipMemcpyHostToDevice) != hipSuccess)
        return fail(set_error(TGH_ERR_DEVICE, "writing the specimen state failed"));
    if (ck->layer_target_count) {
        rc = install_layer_targets(w, ck->layer_axis, ck->layer_targets, ck->layer_target_count, ck->layer_first,
                                   ck->layer_last, ck->layer_stiffness, ck->layer_max_translation, false);
        if (rc) return fail(rc);
    }
    if (ck->vertex_target_count) {
        if (!ck->vertex_targets || !ck->vertex_commands) return fail(set_error(TGH_ERR_INVALID, "checkpoint array is null"));
        rc = install_vertex_targets(w, ck->vertex_axis, ck->vertex_indices, ck->vertex_target_count, 0.f,
                                    ck->vertex_targets, ck->vertex_commands, ck->vertex_stiffness,
                                    ck->vertex_max_translation);
        if (rc) return fail(rc);
    }
    *out = w;
    return TGH_OK;
}

int tgh_world_cell_bounds(const tgh_world *world, float lower[3], float upper[3]) {
    if (!world || !lower || !upper) return set_error(TGH_ERR_INVALID, "null argument");
    if (world->d.nspec != 1) return set_error(TGH_ERR_INVALID, "cell bounds of a one-specimen world");
    for (int k = 0; k < 3; ++k) {
        lower[k] = world->grids[0].lower[k];
        upper[k] = world->grids[0].upper[k];
    }
    return TGH_OK;
}

// compact_cell (world.rs:1931-2053): the kinematics move the material from
// the current cell to the new one, then the grid follows.
int tgh_world_compact_cell(tgh_world *world, const float new_lower[3], const float new_upper[3],
                           uint32_t kinematics) {
    if (!world || !new_lower || !new_upper) return set_error(TGH_ERR_INVALID, "null argument");
    if (kinematics > TGH_COMPACT_AFFINE_VERTICES) return set_error(TGH_ERR_INVALID, "unknown compaction kinematics");
    if (world->d.nspec != 1) return set_error(TGH_ERR_INVALID, "cell changes need a one-specimen world");
    const GridParams old = world->grids[0];
    tgh::CellChange box{};
    for (int k = 0; k < 3; ++k) {
        box.old_lower[k] = old.lower[k];
        box.old_upper[k] = old.upper[k];
        box.new_lower[k] = new_lower[k];
        box.new_upper[k] = new_upper[k];
        box.periodic[k] = old.periodic[k];
    }
    // validate before moving anything: set_cell's checks, on a copy
    {
        const tgh::Sizing &z = world->sizes[0];
        uint64_t C = 1;
        for (int k = 0; k < 3; ++k) {
            if (!std::isfinite(new_lower[k]) || !std::isfinite(new_upper[k]))
                return set_error(TGH_ERR_INVALID, "cell bounds must be finite");
            if (!(new_upper[k] - new_lower[k] > 2.0f * z.max_radius))
                return set_error(TGH_ERR_INVALID, "the cell must be wider than two fiber radii on every axis");
            if (!(new_lower[k] >= old.lower[k]) || !(new_upper[k] <= old.upper[k]))
                return set_error(TGH_ERR_INVALID, "the new cell must lie inside the current one");
            C *= to_u32_saturating(std::fmax(std::floor((new_upper[k] - new_lower[k]) / z.cell_size), 1.0f));
        }
        if (C > world->d.C) return set_error(TGH_ERR_SIZE, "the new cell needs more grid cells than were allocated");
    }
    world->history_valid = false;
    HIP_TRY(tgh::launch_compact(world->d, box, kinematics, world->stream));
    return set_cell(world, new_lower, new_upper);
}

int tgh_world_compaction_metrics(tgh_world *world, float correction_fraction, float contact_stiffness,
                                 float stretch_stiffness, float bending_stiffness, tgh_compaction_metrics *out) {
    if (!world || !out) return set_error(TGH_ERR_INVALID, "null argument");
    if (world->d.nspec != 1) return set_error(TGH_ERR_INVALID, "compaction metrics of a one-specimen world");
    if (!(correction_fraction > 0.0f)) return set_error(TGH_ERR_INVALID, "correction_fraction must be positive");
    if (!(contact_stiffness >= 0.0f) || !(stretch_stiffness >= 0.0f) || !(bending_stiffness >= 0.0f))
        return set_error(TGH_ERR_INVALID, "stiffnesses must be nonnegative");
    if (!world->metrics_scratch) {
        // allocated on first use and kept: every call on this world needs the same size
        const size_t scratch = 13 + 5 * (size_t)world->d.S + 7 * (size_t)world->d.V;
        HIP_TRY(world->alloc(&world->metrics_scratch, scratch));
    }
    float *dev = world->metrics_scratch;
    float values[13];
    hipError_t e = tgh::launch_compaction_metrics(world->d, world->d.grids, correction_fraction, contact_stiffness,
                                                  stretch_stiffness, bending_stiffness, dev + 13, dev,
                                                  world->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(values, dev, sizeof(values), hipMemcpyDeviceToHost, world->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(world->stream);
    if (e != hipSuccess) return set_error(TGH_ERR_DEVICE, std::string("compaction metrics: ") + hipGetErrorString(e));
    for (int k = 0; k < 3; ++k) {
        out->pressure[k] = values[k];
        out->face_pressure[k][0] = values[7 + 2 * k];
        out->face_pressure[k][1] = values[8 + 2 * k];
    }
    out->contact_energy = values[3];
    out->stretch_energy = values[4];
    out->bending_energy = values[5];
    out->total_energy = values[6];
    return TGH_OK;
}

namespace {

// f32::total_cmp's key: the bits, with the magnitude bits of negatives flipped
inline int32_t total_order_key(float x) {
    int32_t i;
    std::memcpy(&i, &x, 4);
    return i ^ (int32_t)((uint32_t)(i >> 31) >> 1);
}

}  // namespace

// capture_contacts (world.rs:2193-2334): a whole-segment cell list on a grid
// of cells at least max(segment cell, list cell) + gap wide, one append per
// accepted pair, then the host sort.
int tgh_world_capture_contacts(tgh_world *world, float maximum_surface_gap, size_t capacity,
                               tgh_contact_candidate *out, size_t *count, uint32_t *overflow) {
    if (!world || !count || !overflow || (capacity && !out)) return set_error(TGH_ERR_INVALID, "null argument");
    if (world->d.nspec != 1) return set_error(TGH_ERR_INVALID, "contact capture of a one-specimen world");
    if (!std::isfinite(maximum_surface_gap) || !(maximum_surface_gap >= 0.0f))
        return set_error(TGH_ERR_INVALID, "maximum_surface_gap must be finite and nonnegative");
    if (capacity == 0) return set_error(TGH_ERR_INVALID, "capacity must be positive");
    if (capacity > 0xffffffffull) return set_error(TGH_ERR_INVALID, "capacity exceeds u32");
    *count = 0;
    *overflow = 0;
    const tgh::Sizing &z = world->sizes[0];
    GridParams g = world->grids[0];
    const float size = std::fmax(z.segment_cell_size, z.cell_size) + maximum_surface_gap;
    uint64_t C = 1;
    for (int k = 0; k < 3; ++k) {
        g.cells[k] = to_u32_saturating(std::fmax(std::floor((g.upper[k] - g.lower[k]) / size), 1.0f));
        g.width[k] = (g.upper[k] - g.lower[k]) / (float)g.cells[k];
        C *= g.cells[k];
    }
    g.cell_base = 0;
    if (C > world->d.C) return set_error(TGH_ERR_SIZE, "the capture grid needs more cells than were allocated");
    if (world->d.S == 0) return TGH_OK;
    struct Buffers {
        GridParams *grid = nullptr;
        uint32_t *counter = nullptr;
        uint2 *pairs = nullptr;
        float2 *coordinates = nullptr;
        float *gaps = nullptr, *angles = nullptr;
        ~Buffers() {
            (void)hipFree(grid);
            (void)hipFree(counter);
            (void)hipFree(pairs);
            (void)hipFree(coordinates);
            (void)hipFree(gaps);
            (void)hipFree(angles);
        }
    } b;
    HIP_TRY(hipMalloc((void **)&b.grid, sizeof(GridParams)));
    HIP_TRY(hipMalloc((void **)&b.counter, 2 * sizeof(uint32_t)));
    HIP_TRY(hipMalloc((void **)&b.pairs, capacity * sizeof(uint2)));
    HIP_TRY(hipMalloc((void **)&b.coordinates, capacity * sizeof(float2)));
    HIP_TRY(hipMalloc((void **)&b.gaps, capacity * sizeof(float)));
    HIP_TRY(hipMalloc((void **)&b.angles, capacity * sizeof(float)));
    HIP_TRY(hipMemcpyAsync(b.grid, &g, sizeof(g), hipMemcpyHostToDevice, world->stream));
    HIP_TRY(hipMemsetAsync(b.counter, 0, 2 * sizeof(uint32_t), world->stream));
    HIP_TRY(tgh::launch_capture(world->d, b.grid, (uint32_t)C, maximum_surface_gap, (uint32_t)capacity, b.counter,
                                b.pairs, b.coordinates, b.gaps, b.angles, world->stream));
    uint32_t counter[2];
    HIP_TRY(hipMemcpyAsync(counter, b.counter, sizeof(counter), hipMemcpyDeviceToHost, world->stream));
    HIP_TRY(hipStreamSynchronize(world->stream));
    const size_t n = std::min<size_t>(counter[0], capacity);
    if (n) {
        // sorted by (first, second) on the device, then copied out as records
        void *records = nullptr;
        HIP_TRY(hipMalloc(&records, n * sizeof(tgh_contact_candidate)));
        hipError_t e = tgh::launch_capture_sort((uint32_t)n, b.pairs, b.coordinates, b.gaps, b.angles, records,
                                                world->stream);
        if (e == hipSuccess) e = hipMemcpy(out, records, n * sizeof(tgh_contact_candidate), hipMemcpyDeviceToHost);
        (void)hipFree(records);
        if (e != hipSuccess) return set_error(TGH_ERR_DEVICE, std::string("capture sort: ") + hipGetErrorString(e));
        // the reference breaks ties of (first, second) by the coordinates'
        // total order; a pair is found once per capture, but order any run of
        // equal pairs that way all the same
        auto less = [](const tgh_contact_candidate &l, const tgh_contact_candidate &r) {
            const int32_t a = total_order_key(l.first_coordinate), b2 = total_order_key(r.first_coordinate);
            if (a != b2) return a < b2;
            return total_order_key(l.second_coordinate) < total_order_key(r.second_coordinate);
        };
        for (size_t i = 0; i < n;) {
            size_t j = i + 1;
            while (j < n && out[j].first_segment == out[i].first_segment &&
                   out[j].second_segment == out[i].second_segment)
                ++j;
            if (j - i > 1) std::sort(out + i, out + j, less);
            i = j;
        }
    }
    *count = n;
    *overflow = counter[1] != 0u ? 1u : 0u;
    return TGH_OK;
}

int tgh_world_set_fiber_maximum_curvatures(tgh_world *world, const float *maximum_curvatures, size_t count) {
    if (!world || (count && !maximum_curvatures)) return set_error(TGH_ERR_INVALID, "null argument");
    if (count != world->d.F) return set_error(TGH_ERR_SIZE, "one curvature limit per fiber");
    for (size_t f = 0; f < count; ++f)
        if (!std::isfinite(maximum_curvatures[f]) || maximum_curvatures[f] < 0.0f)
            return set_error(TGH_ERR_INVALID, "curvature limits must be finite and nonnegative");
    const size_t V = world->d.V;
    std::vector<float4> vA(V);
    std::vector<uint4> spans(world->d.F);
    HIP_TRY(hipMemcpyAsync(vA.data(), world->d.vA, V * sizeof(float4), hipMemcpyDeviceToHost, world->stream));
    HIP_TRY(hipMemcpyAsync(spans.data(), world->d.fspan, spans.size() * sizeof(uint4), hipMemcpyDeviceToHost,
                           world->stream));
    HIP_TRY(hipStreamSynchronize(world->stream));
    // the limit acts on internal vertices only, as at creation
    for (size_t f = 0; f < count; ++f)
        for (uint32_t l = 1; l + 1 < spans[f].y; ++l) vA[spans[f].x + l].x = maximum_curvatures[f];
    HIP_TRY(hipMemcpyAsync(world->d.vA, vA.data(), V * sizeof(float4), hipMemcpyHostToDevice, world->stream));
    HIP_TRY(hipStreamSynchronize(world->stream));
    for (uint32_t sp = 0, f = 0; sp < world->specimens.size(); ++sp) {
        tgh::Packed &p = world->specimens[sp];
        for (uint32_t lf = 0; lf < p.F; ++lf, ++f)
            for (uint32_t l = 1; l + 1 < p.fcount[lf]; ++l) p.kmax[p.fstart[lf] + l] = maximum_curvatures[f];
    }
    return TGH_OK;
}

int tgh_device_info_get(tgh_device_info *out) {
    if (!out) return set_error(TGH_ERR_INVALID, "null argument");
    *out = tgh_device_info{};
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    hipDeviceProp_t prop{};
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return set_error(TGH_ERR_DEVICE, hipGetErrorString(e));
    std::snprintf(out->name, sizeof out->name, "%s", prop.name);
    std::snprintf(out->arch, sizeof out->arch, "%s", prop.gcnArchName);
    out->compute_units = (uint32_t)prop.multiProcessorCount;
    return TGH_OK;
}

int tgh_world_get_tuning(const tgh_world *world, tgh_tuning *out) {
    if (!world || !out) return set_error(TGH_ERR_INVALID, "null argument");
    *out = tgh_tuning{};
    out->curvature_lanes = world->d.gs_lanes;
    out->contact_lanes = world->d.contact_lanes;
    return TGH_OK;
}

int tgh_world_set_tuning(tgh_world *world, const tgh_tuning *tuning) {
    if (!world || !tuning) return set_error(TGH_ERR_INVALID, "null argument");
    const uint32_t lanes = tuning->curvature_lanes;
    if (lanes != 0 && lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64)
        return set_error(TGH_ERR_INVALID, "curvature_lanes must be 0, 4, 8, 16, 32 or 64");
    const uint32_t cl = tuning->contact_lanes;
    if (cl != 4 && cl != 8 && cl != 16 && cl != 32)
        return set_error(TGH_ERR_INVALID, "contact_lanes must be 4, 8, 16 or 32");
    if (lanes != 0) {
        const size_t lds = tgh::curvature_lds_bytes(world->d.max_fiber_vertices, lanes, world->d.track_walls);
        if (world->d.max_fiber_vertices > tgh::kMaxLdsFiberVertices || lds > 64u * 1024u)
            return set_error(TGH_ERR_SIZE, "the longest fiber does not fit in LDS at this lane count");
    }
    HIP_TRY(hipStreamSynchronize(world->stream));
    for (auto &kv : world->graphs) (void)hipGraphExecDestroy(kv.second);
    for (hipGraph_t g : world->graph_templates) (void)hipGraphDestroy(g);
    world->graphs.clear();
    world->graph_templates.clear();
    world->d.gs_lanes = lanes;
    world->d.contact_lanes = cl;
    return TGH_OK;
}

int tgh_world_rebuild_lists(tgh_world *world) {
    if (!world) return set_error(TGH_ERR_INVALID, "world is null");
    hipError_t e = tgh::launch_rebuild_now(world->d, world->stream);
    if (e != hipSuccess) return set_error(TGH_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
    HIP_TRY(hipStreamSynchronize(world->stream));
    return TGH_OK;
}

void tgh_world_destroy(tgh_world *world) { delete world; }

const char *tgh_last_error(void) { return g_error.c_str(); }

}  // extern "C"
