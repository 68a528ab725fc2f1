// Device-side data layout shared by the relaxation kernels and the host world.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tgh {

// Vertex flag bits (vB.z): neighbours along the fiber and pinning.
constexpr uint32_t kHasPrev = 1u;   // previous segment exists
constexpr uint32_t kHasNext = 2u;   // next segment exists
constexpr uint32_t kHasPrev2 = 4u;  // second previous vertex exists (rest-chord bend)
constexpr uint32_t kHasNext2 = 8u;  // second next vertex exists
constexpr uint32_t kPinned = 16u;
constexpr uint32_t kInactive = 32u;  // the vertex's fiber is not active (formation)

constexpr uint32_t kBuildThreads = 128;
// Entries kept per overflowed list (candidates in the reference's scan order).
constexpr uint32_t kOverflowListCapacity = 256;
// Longest fiber the per-wave curvature kernel stages in LDS (9 floats each).
constexpr uint32_t kMaxLdsFiberVertices = 1536;
// Distributed atomic-max slots per metric (penetration, step, curvature) of
// a single-specimen world; a world of several specimens uses kRedSlotsBatch
// per specimen.
constexpr uint32_t kRedSlots = 256;
constexpr uint32_t kRedSlotsBatch = 8;
constexpr uint32_t kMetricPen = 0, kMetricStep = 1, kMetricCurv = 2;

// Vertex flag bits above kSpecShift hold the vertex's specimen index.
constexpr uint32_t kSpecShift = 8;
constexpr uint32_t kFlagMask = (1u << kSpecShift) - 1u;

// One specimen's uniform grid: widths are (upper - lower) / cells computed
// once in f32, exactly as the reference's proxy_cell computes them per call.
// The specimen's cells occupy [cell_base, cell_base + cells product) of the
// world's cell list and its segments [seg_base, seg_base + seg_count) of the
// world's segments, which are also the cell-list slots its active segments
// fill from seg_base on; `skin` is its Verlet skin.
// A change of the cell: old and new bounds and the periodic axes.
struct CellChange {
    float old_lower[3], old_upper[3], new_lower[3], new_upper[3];
    uint32_t periodic[3];
};

struct GridParams {
    float lower[3];
    float upper[3];
    float width[3];
    float length[3];
    uint32_t cells[3];
    uint32_t periodic[3];
    uint32_t cell_base;
    float skin;
    uint32_t seg_base, seg_count;
};

// Per-specimen control block. `control` follows the reference's buffer
// (world.rs:466-467): {running, lifetime iterations, batch iterations,
// terminal failure}; `metrics` = {max penetration, max vertex step, max
// curvature ratio} measured at the specimen's last assessment; `rebuild` =
// the specimen's lists are stale (its own neighbor_state[0], neighbor.rs:1-9:
// each specimen rebuilds when its own vertices have moved, as it would alone).
struct SpecState {
    uint32_t control[4];
    float metrics[3];
    uint32_t rebuild;
};

// World-wide state: nstate = {some specimen's rebuild is pending, passes that
// rebuilt lists} (neighbor.rs:1-9). A cell-list slot that holds no segment
// (past a specimen's active segments) has slotC.x = cell_segs = kNoSegment.
struct DevState {
    uint32_t nstate[2];
    uint32_t pad[6];
};
constexpr uint32_t kNoSegment = 0xffffffffu;

// Installed formation targets: layer targets (per layer coordinate along
// `layer_axis`, fibers of layers first..last held) and vertex targets
// (per vertex: index, target coordinate, command coordinate along
// `vertex_axis`). Counts of 0 mean none.
struct FormationDev {
    uint32_t layer_count, layer_axis, first_layer, last_layer;
    float layer_stiffness, layer_max_translation;
    float *layer_targets;
    uint32_t vertex_count, vertex_axis;
    float vertex_stiffness, vertex_max_translation;
    uint32_t *vertex_index;
    float *vertex_target, *vertex_command;
    float *error;  // [2]: layer, vertex
};

// Device pointers and sizes of one world (one or more specimens relaxed
// together), passed to the launchers. Specimens are stored one after
// another: fibers, vertices, segments and cells of specimen s follow those
// of specimen s - 1.
struct WorldDev {
    uint32_t F, V, S, C;
    uint32_t nspec;
    uint32_t red_slots;  // atomic-max slots per specimen and metric
    const GridParams *grids;  // [nspec], device
    SpecState *ss;            // [nspec], device
    uint32_t block;  // neighbour-list slots per segment
    uint32_t ovf_cap;
    uint32_t track_walls;
    uint32_t max_fiber_vertices;
    uint32_t gs_lanes;       // lanes per fiber in the curvature sweep (0, 4, 8, 16, 32 or 64)
    uint32_t contact_lanes;  // lanes per segment in the contact pass (4, 8, 16 or 32)
    uint32_t prune_lists;    // the list build bounds stencil cells by distance (k_build_lists<true>)
    DevState *st;
    uint32_t *red;  // [nspec][3][red_slots] u32 bit patterns of nonnegative floats
    float4 *pos, *alt, *refpos, *wall, *vA, *vB;
    uint4 *sstat;   // per segment: first vertex, fiber, radius bits, specimen
    uint4 *fspan;   // per fiber: first vertex, vertex count, specimen, 0
    uint32_t *fstep;    // per fiber: formation step
    uint32_t *flayer;   // per fiber: formation layer (0xffffffff: none)
    uint32_t *factive;  // per fiber: 1 active, 0 not
    uint32_t *cell_counts, *cell_offsets, *cell_cursor, *cell_segs, *seg_cell, *scattered;
    uint32_t *cell_spec;  // per cell: its specimen
    // per cell: the largest half length plus radius of its segments at the
    // last rebuild's binning (f32 bits), for the list build's cell bound
    uint32_t *cell_reach;
    // slot records in cell-list order, refreshed from `pos` every iteration:
    // slotA/slotB = the segment's two vertex records (x, y, z, rest arc
    // length), slotC = (segment, fiber, radius bits, specimen)
    float4 *slotA, *slotB;
    uint4 *slotC;
    // the contact pass's compact slot records, refreshed every iteration:
    // cP = (first vertex x, y, z, radius), cQ = (second vertex x, y, z,
    // segment bits)
    float4 *cP, *cQ;
    // lists: per slot its count and, for a listed one, the first entry of
    // its row in `nbr` (the rows of each 8 consecutive slots are packed in
    // their 8 x block entries); an overflowed list's row in `ovf`
    uint32_t *ncount, *nbr, *nbr_off, *ovf;
    float4 *corr0, *corr1;
    float *ratio;  // per vertex: curvature ratio at the last measure (0 until measured)
};

// Per-batch scalar settings of the relaxation loop (RelaxationConfig).
struct IterParams {
    float penetration_tolerance;
    float curvature_ratio_tolerance;
    float correction_fraction;
    float stretch;
    float bend;
    float max_step;
    float curvature_stiffness;
    float curvature_margin;
    uint32_t aggregation;
    uint32_t constraint_iterations;
    uint32_t sweeps;
    uint32_t force_full;
    uint32_t iterations;  // batch size, compared with control[2]
};

// Dynamic LDS of one curvature-sweep workgroup (64 / lanes fibers).
// The slot records' positions from `pos` (before any reader of the records
// that does not follow a rebuild).
hipError_t launch_refresh_slots(const WorldDev &w, hipStream_t stream);
size_t curvature_lds_bytes(uint32_t max_vertices, uint32_t lanes, uint32_t track_walls);

// Enqueue work on `stream`; every launcher checks its grid against the
// world's sizes and returns a HIP error code (hipSuccess when enqueued).
hipError_t launch_begin_batch(const WorldDev &w, hipStream_t stream);
hipError_t launch_rebuild(const WorldDev &w, hipStream_t stream);
hipError_t launch_rebuild_now(const WorldDev &w, hipStream_t stream);
// Finishes a pending rebuild (if any) now, from the current positions: the
// same lists, snapshot and rebuild count the next pass would have produced.
hipError_t launch_refresh_lists(const WorldDev &w, hipStream_t stream);
// Marks the lists stale so the next pass rebuilds them: every specimen's,
// or those of the specimens whose entry of `flags` ([nspec], device) is set.
hipError_t launch_request_rebuild(const WorldDev &w, hipStream_t stream);
hipError_t launch_request_rebuild_specimens(const WorldDev &w, const uint32_t *flags, hipStream_t stream);
// One relaxation iteration is these five groups in order (world.rs:1038-1501).
hipError_t launch_contact(const WorldDev &w, const IterParams &p, hipStream_t stream);
hipError_t launch_measure(const WorldDev &w, const IterParams &p, hipStream_t stream);
hipError_t launch_vertex_passes(const WorldDev &w, const IterParams &p, hipStream_t stream);
hipError_t launch_curvature(const WorldDev &w, const IterParams &p, hipStream_t stream);
// Formation: activation of fibers through a step; layer and vertex targets.
hipError_t launch_activate(const WorldDev &w, uint32_t maximum_step, hipStream_t stream);
// mirrors w.factive (as uploaded) into every vertex's inactive flag
hipError_t launch_apply_activity(const WorldDev &w, hipStream_t stream);
// compaction: kinematics 0 rigid fiber centres, 1 moving walls, 2 affine
hipError_t launch_compact(const WorldDev &w, const CellChange &box, uint32_t kinematics, hipStream_t stream);
// the 13 compaction metrics (measure_compaction_metrics order) into out[13];
// `terms` is scratch of 5 S + 7 V floats
hipError_t launch_compaction_metrics(const WorldDev &w, const GridParams *grid, float correction_fraction,
                                     float contact_stiffness, float stretch_stiffness, float bending_stiffness,
                                     float *terms, float *out, hipStream_t stream);
// the first n captured pairs sorted by (first, second) segment on the device
// (radix sort of 64-bit keys), gathered as 24-byte candidate records (the
// tgh_contact_candidate layout) into `records`; synchronises
hipError_t launch_capture_sort(uint32_t n, const uint2 *pairs, const float2 *coordinates, const float *gaps,
                               const float *angles, void *records, hipStream_t stream);
hipError_t launch_capture(const WorldDev &w, const GridParams *capture_grid, uint32_t capture_cells, float max_gap,
                          uint32_t capacity, uint32_t *counter, uint2 *pairs, float2 *coordinates, float *gaps,
                          float *angles, hipStream_t stream);
hipError_t launch_layer_targets(const WorldDev &w, const FormationDev &f, bool gate_on_running, hipStream_t stream);
hipError_t launch_init_vertex_targets(const WorldDev &w, const FormationDev &f, float displacement,
                                      hipStream_t stream);
hipError_t launch_vertex_targets(const WorldDev &w, const FormationDev &f, hipStream_t stream);
hipError_t launch_target_error(const WorldDev &w, const FormationDev &f, hipStream_t stream);

// Kernel groups of one iteration, in launch order; used for event timing.
enum KernelGroup { kGroupRebuild = 0, kGroupContact, kGroupMeasure, kGroupVertex, kGroupCurvature, kGroupCount };

}  // namespace tgh
