/*
 * tangle_hip.h — C ABI of the MI355X relaxation backend for TANGLE.
 *
 * One world holds one fiber assembly resident on the GPU and relaxes it with
 * position-based contact, stretch/bend and admissible-curvature projections.
 * The entry points mirror TANGLE's device-world operator API (Rust):
 *
 *   tgh_world_create            DeviceWorld::new / DeviceFiberWorld::upload
 *                               crates/tangle_relax/src/device/plugin.rs:422-424,
 *                               crates/tangle_relax/src/device/world.rs:308-690
 *                               (packing: crates/tangle_relax/src/device/packed.rs:150-320)
 *   tgh_world_run_batch         DeviceWorld::run_batch + read_status
 *                               plugin.rs:407-418, world.rs:1020-1505, 2337-2369
 *   tgh_world_download_positions DeviceWorld::download_positions   world.rs:2384-2390
 *   tgh_world_upload_positions  (no direct counterpart: the reference re-uploads
 *                               through DeviceFiberWorld::restore, world.rs:692-823;
 *                               here it replaces positions in place and requests a
 *                               neighbour-list rebuild, request_neighbor_list_rebuild
 *                               crates/tangle_contact/src/device/neighbor.rs:434-441)
 *   tgh_world_info              cell_size()/cell_count()/neighbor_list_rebuilds()
 *                               plugin.rs:324-893, world/neighbor_list.rs:166-176
 *   tgh_world_cell_list,        read-back of cell_counts/cell_offsets/cell_segments
 *   tgh_world_neighbor_lists    and neighbor_counts/neighbor_segments, as the
 *                               reference's own tests read them
 *                               (crates/tangle_relax/src/device/world_tests.rs:1342-1362)
 *   tgh_generate_population     generate_biased_fiber_population
 *                               crates/tangle_generate/src/population.rs:246-357
 *
 * Conventions. Every function returns 0 on success and a negative code on
 * failure; tgh_last_error() then names the problem (thread-local string). The
 * caller owns every input and output buffer; the world owns device memory.
 * A world is not thread-safe: use one world from one thread at a time.
 * All lengths are SI metres, positions are interleaved xyz.
 */
#ifndef TANGLE_HIP_H
#define TANGLE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TGH_OK 0
#define TGH_ERR_INVALID (-1)   /* invalid argument or configuration (Python: ValueError) */
#define TGH_ERR_DEVICE (-2)    /* HIP runtime failure (Python: RuntimeError)             */
#define TGH_ERR_SIZE (-3)      /* caller buffer too small                                 */

/* A solver-neutral assembly: the FiberAssembly fields the relaxation reads
 * (crates/tangle_core/src/assembly.rs; packed by packed.rs:150-320).
 * Fibers are stored one after another: fiber f owns vertices
 * [sum(counts[0..f]), sum(counts[0..f+1])). Round sections only. */
typedef struct tgh_assembly {
    uint32_t fiber_count;
    uint32_t vertex_count;
    const uint32_t *fiber_vertex_counts;  /* [fiber_count], each >= 2            */
    const double *placed;                 /* [3*vertex_count] current positions  */
    const double *intrinsic;              /* [3*vertex_count] rest shape          */
    const double *fiber_radius;           /* [fiber_count] > 0                    */
    const double *fiber_min_bend_radius;  /* [fiber_count]; <= 0 or NaN: no limit, NULL: none */
    const uint8_t *vertex_pinned;         /* [vertex_count] or NULL               */
    double cell_origin[3];
    double cell_lengths[3];
    uint32_t cell_periodic[3];
    uint32_t pin_fiber_ends;              /* pin first/last vertex of every fiber */
    /* formation (crates/tangle_relax/src/device/packed.rs:94, 202-210):
     * the recipe step that inserts each fiber (NULL: all 0) and its
     * manufacturing layer (NULL, or 0xffffffff for a fiber: none) */
    const uint32_t *fiber_formation_step;
    const uint32_t *fiber_formation_layer;
} tgh_assembly;

/* RelaxationConfig (plugin.rs:77-175) restricted to the flexible, round,
 * non-adaptive relaxation loop; CellListConfig (tangle_contact/src/lib.rs:14-40). */
typedef struct tgh_config {
    float penetration_tolerance;
    float correction_fraction;
    float stretch_stiffness;
    float bend_stiffness;
    float curvature_limit_stiffness;
    float curvature_limit_safety_margin;
    float curvature_ratio_tolerance;
    float max_step;
    float cell_size_scale;
    float neighbor_skin_scale;
    uint32_t contact_aggregation;        /* 0 average, 1 penetration-weighted, 2 deepest */
    uint32_t constraint_iterations;
    uint32_t curvature_cleanup_sweeps;
    uint32_t force_full_iterations;
    uint32_t neighbor_capacity;
    uint32_t maximum_neighbor_list_bytes; /* 0: the reference's 120 MiB (world.rs:36) */
} tgh_config;

/* BatchStatus (world.rs:89-114). */
typedef struct tgh_batch_status {
    uint64_t total_iterations;
    uint64_t batch_iterations;
    float max_penetration;
    float max_displacement;
    float max_curvature_ratio;
    uint32_t converged;
    uint32_t cell_list_overflow;
} tgh_batch_status;

typedef struct tgh_world_info {
    uint32_t fiber_count;
    uint32_t segment_count;
    uint32_t vertex_count;
    uint32_t cells[3];
    uint32_t cell_count;
    float cell_size;
    float neighbor_skin;
    uint32_t neighbor_block;             /* list slots per segment actually used */
    uint32_t neighbor_list_rebuilds;     /* neighbor_state[1] */
    uint64_t device_bytes;               /* device memory held by the world */
    uint32_t specimen_count;             /* specimens relaxed together (1 unless batched) */
    uint32_t reserved;
} tgh_world_info;

typedef struct tgh_world tgh_world;

/* Fills `config` with RelaxationConfig::default() (plugin.rs:145-175). */
void tgh_config_default(tgh_config *config);

int tgh_world_create(const tgh_assembly *assembly, const tgh_config *config, tgh_world **out);

/* PackedAssembly (crates/tangle_relax/src/device/packed.rs:47-141), the f32
 * structure of arrays DeviceWorld::new takes (plugin.rs:422-424), for the
 * uniform round layout this backend relaxes. Field meanings are the
 * reference's: fibers follow one another (fiber_vertex_spans: first vertex
 * and count, each span starting where the previous ended); segment first - f
 * + l of fiber f joins vertices first + l and first + l + 1 (the layout
 * from_assembly_with_options writes, packed.rs:150-320); one radius per
 * fiber. Arrays the uniform layout does not need (segment_children,
 * refinement levels and epochs, vertex_active, directors, lane offsets,
 * junction tethers) are not taken; `adaptive` and `has_ovals` must be 0 and
 * segment_active, when given, all 1. */
typedef struct tgh_packed {
    uint32_t fiber_count;
    uint32_t vertex_count;
    uint32_t segment_count;                /* vertex_count - fiber_count */
    const float *positions;                /* [3 * vertex_count] */
    const float *intrinsic_positions;      /* [3 * vertex_count] */
    const uint32_t *segment_vertices;      /* [2 * segment_count] */
    const uint32_t *segment_fibers;        /* [segment_count] */
    const float *segment_radii;            /* [segment_count] */
    const float *segment_rest_lengths;     /* [segment_count] */
    const uint32_t *segment_active;        /* [segment_count] or NULL (all active) */
    const uint32_t *fiber_vertex_spans;    /* [2 * fiber_count] */
    const uint32_t *fiber_formation_layers; /* [fiber_count] or NULL; 0xffffffff: none */
    const uint32_t *fiber_formation_steps;  /* [fiber_count] or NULL: all 0 */
    const float *vertex_max_curvature;     /* [vertex_count]; 0: unrestricted */
    const uint32_t *vertex_pinned;         /* [vertex_count] */
    float cell_lower[3];
    float cell_upper[3];
    uint32_t cell_periodic[3];
    uint32_t adaptive;                     /* must be 0 */
    uint32_t has_ovals;                    /* must be 0 */
} tgh_packed;

/* DeviceWorld::new(&RelaxationConfig, PackedAssembly) (plugin.rs:422-424,
 * world.rs:308-690): creates a world from a packed assembly. A layout that is
 * not the uniform round one above is rejected with TGH_ERR_INVALID before any
 * device call. */
int tgh_world_create_packed(const tgh_packed *packed, const tgh_config *config, tgh_world **out);
/* Batched specimens (no counterpart in the reference, which relaxes one
 * world per DeviceWorld): `count` independent assemblies, each with its own
 * cell, grid, skin and convergence, relaxed together by the same launches.
 * Vertices, fibers and segments are numbered specimen after specimen, and
 * tgh_world_download_positions returns them in that order. Each specimen
 * stops on its own gate; tgh_world_run_batch reports the worst of each
 * field (converged only when all converged) and tgh_world_specimen_status
 * each specimen's status of the last batch. The neighbour-list rebuild is
 * shared: any specimen's displacement trigger rebuilds every list. */
int tgh_world_create_batch(const tgh_assembly *specimens, uint32_t count, const tgh_config *config,
                           tgh_world **out);
int tgh_world_specimen_status(const tgh_world *world, tgh_batch_status *out, size_t count);
int tgh_world_run_batch(tgh_world *world, const tgh_config *config, uint32_t iterations,
                        tgh_batch_status *status);
int tgh_world_download_positions(const tgh_world *world, float *out, size_t count);
int tgh_world_upload_positions(tgh_world *world, const float *positions, size_t count);
int tgh_world_info_get(const tgh_world *world, tgh_world_info *info);
/* Rebuilds the cell list and neighbour lists from the current positions now
 * (request_neighbor_list_rebuild followed by rebuild_neighbor_lists_if_requested,
 * crates/tangle_contact/src/device/neighbor.rs:434-441,
 * crates/tangle_relax/src/device/world/neighbor_list.rs:19-62); used to compare
 * lists built from identical positions. */
int tgh_world_rebuild_lists(tgh_world *world);
/* counts/offsets: [cell_count]; segments: [segment_count] in cell order. */
int tgh_world_cell_list(const tgh_world *world, uint32_t *counts, uint32_t *offsets,
                        uint32_t *segments, size_t cell_capacity, size_t segment_capacity);
/* counts: [segment_count] true neighbour counts (may exceed the block);
 * lists: [segment_count * neighbor_block], row s holds segment s's list in
 * ascending order when counts[s] <= neighbor_block (unspecified otherwise). */
int tgh_world_neighbor_lists(const tgh_world *world, uint32_t *counts, uint32_t *lists,
                             size_t list_capacity);
/* Kernel-group timing with HIP events on the world's stream (no counterpart
 * in the reference, which has no kernel profiler). When enabled, later
 * batches are launched directly (not as a graph) and every group of every
 * iteration is bracketed by events: 0 list rebuild, 1 contact, 2 curvature measure and
 * convergence assessment, 3 stretch/bend/contact-apply passes, 4 curvature
 * Gauss-Seidel. tgh_world_timing returns the accumulated milliseconds and
 * launch counts per group (groups >= 5) and clears them when `reset`. */
int tgh_world_set_timing(tgh_world *world, uint32_t enable);
int tgh_world_timing(tgh_world *world, double *group_ms, uint64_t *group_launches, size_t groups,
                     uint32_t reset);
/* ---- formation targets (kernels.rs:164-221, 538-710; world.rs:1424-1482,
 * 1723-1927, 2139-2180) -------------------------------------------------
 * tgh_world_activate_formation_step: every fiber whose formation step is at
 * most `maximum_step` becomes active, the others inactive (activate_formation
 * _step); inactive fibers are left out of the cell list, the neighbour lists,
 * contact and every projection. Returns the active segment and vertex counts.
 * tgh_world_set_layer_targets: fibers of layers first..last are translated
 * along `axis` toward targets[layer] (their vertex-mean coordinate) by
 * clamp((target - mean) * stiffness, +-max_translation), once now and after
 * every iteration's projections (apply_layer_target_range, world.rs:1752-1803).
 * tgh_world_set_vertex_targets: each listed vertex is commanded from its
 * current coordinate toward coordinate + displacement along `axis`, by at
 * most max_translation per iteration, and set to the command
 * (apply_vertex_displacement_targets, world.rs:1813-1858).
 * tgh_world_formation_target_error: the largest |target - mean| over held
 * layers and |target - coordinate| over held vertices (formation_target_error,
 * world.rs:1867-1927); has_layer / has_vertex tell which targets exist. */
int tgh_world_activate_formation_step(tgh_world *world, uint32_t maximum_step, uint64_t *active_segments,
                                      uint64_t *active_vertices);
int tgh_world_set_layer_targets(tgh_world *world, uint32_t axis, const float *targets, size_t count,
                                uint32_t first_layer, uint32_t last_layer, float stiffness, float max_translation);
int tgh_world_clear_layer_targets(tgh_world *world);
int tgh_world_set_vertex_targets(tgh_world *world, uint32_t axis, const uint32_t *vertices, size_t count,
                                 float displacement, float stiffness, float max_translation);
int tgh_world_clear_vertex_targets(tgh_world *world);
int tgh_world_formation_target_error(tgh_world *world, float *layer_error, float *vertex_error,
                                     uint32_t *has_layer, uint32_t *has_vertex);

/* ---- checkpoint and restore (DeviceWorldCheckpoint, world.rs:281-304, 692-823) ----
 * The persistent state of a one-specimen world: positions, fiber activity,
 * wall reactions, the lifetime iteration count and the held formation
 * targets. Broad-phase and correction scratch is left out; the first batch
 * after a restore rebuilds it, as the reference's restore does.
 * tgh_world_checkpoint_sizes fills the counts (fiber_count, vertex_count,
 * layer_target_count, vertex_target_count) and leaves the pointers alone;
 * tgh_world_checkpoint then fills a checkpoint whose counts are those and
 * whose arrays hold them. tgh_world_restore creates a world from the
 * assembly the checkpointed world was created from and puts it in the
 * checkpoint's state (DeviceFiberWorld::restore), including a cell compacted
 * since creation (it must lie inside the assembly's cell). */
typedef struct tgh_checkpoint {
    uint64_t total_iterations;
    uint32_t fiber_count;
    uint32_t vertex_count;
    float *positions;             /* [3 * vertex_count] */
    uint32_t *fiber_active;       /* [fiber_count]: 1 active, 0 not yet inserted */
    float *wall_reactions;        /* [3 * vertex_count]: zero on periodic axes */
    uint32_t layer_target_count;  /* 0: no layer target held */
    uint32_t layer_axis, layer_first, layer_last;
    float layer_stiffness, layer_max_translation;
    float *layer_targets;         /* [layer_target_count] */
    uint32_t vertex_target_count; /* 0: no vertex target held */
    uint32_t vertex_axis;
    float vertex_stiffness, vertex_max_translation;
    uint32_t *vertex_indices;     /* [vertex_target_count] */
    float *vertex_targets;        /* [vertex_target_count] target coordinates */
    float *vertex_commands;       /* [vertex_target_count] commanded coordinates */
    float cell_lower[3];          /* the world's cell at the checkpoint (it may */
    float cell_upper[3];          /* have been compacted since creation) */
} tgh_checkpoint;

int tgh_world_checkpoint_sizes(tgh_world *world, tgh_checkpoint *out);
int tgh_world_checkpoint(tgh_world *world, tgh_checkpoint *out);
int tgh_world_restore(const tgh_assembly *assembly, const tgh_config *config, const tgh_checkpoint *checkpoint,
                      tgh_world **out);

/* ---- compaction and contact capture (world.rs:1931-2334, kernels.rs:712-923,
 * crates/tangle_contact/src/device/contact.rs:9-230); one-specimen worlds ----
 * tgh_world_cell_bounds: the resident cell (cell_bounds).
 * tgh_world_compact_cell: the cell becomes [new_lower, new_upper], inside the
 * current one on every axis and wider than two fiber radii, and the active
 * material follows it by `kinematics` (compact_cell):
 *   TGH_COMPACT_RIGID_FIBER_CENTERS: each fiber translates so that its
 *     vertex-mean centre keeps its fraction of the cell; reactions reset;
 *   TGH_COMPACT_MOVING_WALLS: vertices are projected inside the new walls of
 *     the bounded axes; the wall reaction is the projection;
 *   TGH_COMPACT_AFFINE_VERTICES: every vertex keeps its fractional
 *     coordinates; reactions reset.
 * The grid keeps its cell size (fewer cells) and the lists are rebuilt by the
 * next batch.
 * tgh_world_compaction_metrics: directional pressures and penalty energies of
 * the current state from the last iteration's contact corrections, curvature
 * ratios and wall reactions (compaction_metrics, measure_compaction_metrics);
 * the reductions run in index order, as the reference's single thread does.
 * tgh_world_capture_contacts: every pair of active segments of different
 * fibers whose capsule surfaces are at most `maximum_surface_gap` apart, with
 * the closest-point coordinates and the crossing angle, sorted by (first,
 * second, first coordinate, second coordinate) (capture_contacts). At most
 * `capacity` are returned; *overflow is set when more were found. The
 * capture uses the shared cell buffers, so the next batch rebuilds its lists.
 * tgh_world_set_fiber_maximum_curvatures: per-fiber curvature limits
 * (inverse length; 0 disables) replace the assembly's bend radii
 * (set_fiber_maximum_curvatures). */
#define TGH_COMPACT_RIGID_FIBER_CENTERS 0u
#define TGH_COMPACT_MOVING_WALLS 1u
#define TGH_COMPACT_AFFINE_VERTICES 2u
typedef struct tgh_compaction_metrics {
    float pressure[3];        /* per axis: contact stiffness x load / (projection x face area) */
    float contact_energy;     /* sum of k_c p^2 / 2 over segment penetrations */
    float stretch_energy;     /* sum of k_s (l - l0)^2 / (2 l0) */
    float bending_energy;     /* sum of k_b max(ratio - 1, 0)^2 / 2 */
    float total_energy;
    float face_pressure[3][2]; /* per axis: [lower face, upper face] from wall reactions */
} tgh_compaction_metrics;
typedef struct tgh_contact_candidate {
    uint32_t first_segment, second_segment; /* first < second, different fibers */
    float first_coordinate, second_coordinate; /* closest points, in [0, 1] */
    float surface_gap;                      /* centre distance minus both radii */
    float crossing_angle;                   /* acos |cos| of the axes, radians */
} tgh_contact_candidate;
int tgh_world_cell_bounds(const tgh_world *world, float lower[3], float upper[3]);
int tgh_world_compact_cell(tgh_world *world, const float new_lower[3], const float new_upper[3],
                           uint32_t kinematics);
int tgh_world_compaction_metrics(tgh_world *world, float correction_fraction, float contact_stiffness,
                                 float stretch_stiffness, float bending_stiffness, tgh_compaction_metrics *out);
int tgh_world_capture_contacts(tgh_world *world, float maximum_surface_gap, size_t capacity,
                               tgh_contact_candidate *out, size_t *count, uint32_t *overflow);
int tgh_world_set_fiber_maximum_curvatures(tgh_world *world, const float *maximum_curvatures, size_t count);

/* Launch shape of the relaxation kernels (no counterpart in the reference,
 * whose launch shapes are fixed). Any setting gives the same positions bit
 * for bit; tgh_world_set_tuning validates the settings and drops captured
 * graphs so the next batch is captured with the new shape.
 *   curvature_lanes: lanes that sweep one fiber in the curvature Gauss-Seidel
 *     pass: 4, 8, 16, 32 or 64 (fibers staged in LDS, at most 64 KiB per
 *     workgroup; at 8 the sweep walks every triplet with the triplet
 *     carried in registers, the other widths screen ahead), or 0 for one
 *     lane per fiber walking it in global memory.
 *   contact_lanes: lanes that evaluate one segment's neighbour list in the
 *     contact pass, one candidate each: 4, 8, 16 or 32. */
typedef struct tgh_tuning {
    uint32_t curvature_lanes;
    uint32_t contact_lanes;
    uint32_t reserved[6];
} tgh_tuning;
int tgh_world_get_tuning(const tgh_world *world, tgh_tuning *out);
int tgh_world_set_tuning(tgh_world *world, const tgh_tuning *tuning);

/* The current HIP device (no counterpart in the reference): its name, its
 * architecture (gcnArchName) and compute-unit count, so that launch-shape
 * choices measured on one device are not reused on another. */
typedef struct tgh_device_info {
    char name[256];
    char arch[64];
    uint32_t compute_units;
    uint32_t reserved[7];
} tgh_device_info;
int tgh_device_info_get(tgh_device_info *out);
void tgh_world_destroy(tgh_world *world);
const char *tgh_last_error(void);

/* ---- learned warm start (no counterpart in the reference) -------------
 * A per-segment MLP 64 -> 64 -> 64 -> 16, ReLU, BF16 operands with FP32
 * accumulation on MFMA, trained on the GPU from trajectories recorded from
 * the plain relaxer. 55 inputs are used: for the 16 nearest listed
 * neighbours the unit closest-point direction times min((r1 + r2) /
 * distance, 8); the unit tangent; and the recent motion since the world's
 * last history mark (h = midpoint displacement / diameter, m = |h|): h / m
 * and log10(m) / 8 + 1, zero without a mark or motion. 3 outputs are used:
 * the predicted midpoint displacement over the next interval in units of
 * m diameters. Parameters are one f32 vector: W1 [64][64], b1 [64],
 * W2 [64][64], b2 [64], W3 [16][64], b3 [16] (weights [out][in]). */
typedef struct tgh_mlp tgh_mlp;
size_t tgh_mlp_parameter_count(void);
int tgh_mlp_create(uint64_t seed, tgh_mlp **out);
void tgh_mlp_destroy(tgh_mlp *mlp);
int tgh_mlp_get_parameters(const tgh_mlp *mlp, float *out, size_t count);
int tgh_mlp_set_parameters(tgh_mlp *mlp, const float *in, size_t count);
/* features: rows x 64 bf16 bit patterns; out: rows x 3 */
int tgh_mlp_forward(tgh_mlp *mlp, const uint16_t *features, size_t rows, float *out);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): tgh_mlp_forward with the features and predictions staged in a
 * static device array of the library instead of per-call allocations. It
 * calls no hipMalloc or hipFree. At most TGH_MLP_FORWARD_STATIC_MAX_ROWS rows
 * (TGH_ERR_SIZE beyond); the array is 9.75 MiB on each GPU the library runs
 * on, held until the process ends. Every call on one GPU shares that GPU's
 * array, so calls that overlap on one GPU (from several threads or models)
 * race: serialise them. Results equal tgh_mlp_forward's bit for bit. */
#define TGH_MLP_FORWARD_STATIC_MAX_ROWS 65536u
int tgh_mlp_forward_static(tgh_mlp *mlp, const uint16_t *features, size_t rows, float *out);
/* loss = mean over rows of the squared error summed over the 3 outputs;
 * grad: tgh_mlp_parameter_count() floats */
int tgh_mlp_gradients(tgh_mlp *mlp, const uint16_t *features, const float *targets, size_t rows, float *grad,
                      float *loss);
/* Marks the world's current segment midpoints as the reference of the
 * motion features (history); proposals and recordings read the motion since
 * the last mark. */
int tgh_world_mark_history(tgh_world *world);
/* Trajectory recording: features and segment midpoints of `world` now, then
 * (after the caller relaxes it) the midpoint displacement in units of the
 * recent motion as the targets; the pair of calls appends segment_count rows. */
int tgh_mlp_record_begin(tgh_mlp *mlp, tgh_world *world);
int tgh_mlp_record_end(tgh_mlp *mlp, tgh_world *world);
uint64_t tgh_mlp_dataset_rows(const tgh_mlp *mlp);
int tgh_mlp_dataset_clear(tgh_mlp *mlp);
/* Adam over shuffled minibatches of the recorded rows; epoch_losses: [epochs] or NULL */
int tgh_mlp_train(tgh_mlp *mlp, uint32_t epochs, uint32_t batch_rows, float learning_rate, uint64_t seed,
                  float *epoch_losses);
/* features of every segment at the world's current positions: segment_count x 64 bf16 */
int tgh_world_features(tgh_world *world, uint16_t *out, size_t count);
/* Moves every unpinned vertex by `scale` times the mean predicted
 * displacement of its segments (prediction times m diameters), keeping the
 * previous positions for tgh_world_revert_proposal; both request a
 * neighbour-list rebuild. Segments without recent motion do not move. */
int tgh_world_propose(tgh_world *world, tgh_mlp *mlp, float scale);
int tgh_world_revert_proposal(tgh_world *world);
/* Reverts the last proposal for the specimens with revert[s] != 0 only
 * (revert: [specimen_count]); the others keep it. Requests a rebuild. */
int tgh_world_revert_proposal_specimens(tgh_world *world, const uint32_t *revert, size_t count);
/* Average device time of one forward pass and of one gradient pass (forward,
 * backward and the fixed-order reduction) over `rows` feature rows, each
 * timed over `repeats` launches with HIP events. */
int tgh_mlp_time_kernels(tgh_mlp *mlp, uint32_t rows, uint32_t repeats, float *forward_ms, float *gradient_ms);
/* Matrix-core layout check: d (16x16 f32) = a (16x16 bf16) * b (16x16 bf16), row-major. */
int tgh_mfma_probe(const uint16_t *a, const uint16_t *b, float *d);

/* ---- seeded population generator (population.rs:246-357, 554-813) ---- */

#define TGH_ORIENT_ISOTROPIC 0   /* OrientationDistribution::Isotropic3d           */
#define TGH_ORIENT_PLANAR 1      /* ::Planar { normal, maximum_tilt }              */
#define TGH_ORIENT_ALIGNED 2     /* ::Aligned { axis, maximum_angle }              */
#define TGH_ORIENT_SCHLADITZ 3   /* Schladitz density about `axis`, parameter beta
                                    (no sampler in the reference; its density and
                                    quantile: tangle_characterize/src/shape.rs:26-30,
                                    552-556)                                       */
#define TGH_ORIENT_LAYERED_BIAXIAL 4 /* ::LayeredBiaxial { normal, primary_fraction,
                                    cross_fraction, maximum_in_plane_deviation,
                                    maximum_tilt, layer_seed } (:599-628); needs
                                    TGH_POSITION_LAYERED                           */

#define TGH_POSITION_UNIFORM 0   /* PositionDistribution::Uniform (population.rs:80-100) */
#define TGH_POSITION_LAYERED 1   /* ::Layered { axis, layers, jitter_fraction }    */
#define TGH_POSITION_GRADIENT 2  /* ::DensityGradient { axis, exponent, toward_high } */

typedef struct tgh_population_spec {
    uint32_t count;
    uint32_t segments_per_fiber;
    uint64_t seed;
    double length_min, length_max;        /* equal: ScalarDistribution::Constant */
    double radius_min, radius_max;
    double amplitude_min, amplitude_max;
    uint32_t orientation;                 /* TGH_ORIENT_* */
    double orientation_vector[3];         /* planar normal / aligned or Schladitz axis */
    double orientation_angle;             /* planar max tilt / aligned max angle */
    double orientation_beta;              /* Schladitz beta (shape.rs convention) */
    double min_bend_radius;               /* <= 0: none */
    uint32_t max_attempts_per_fiber;
    double cell_origin[3];
    double cell_lengths[3];
    uint32_t cell_periodic[3];
    uint32_t position;                    /* TGH_POSITION_* */
    uint32_t position_axis;               /* layered / gradient axis */
    uint32_t position_layers;             /* layered: layer count */
    uint32_t position_toward_high;        /* gradient: concentrate toward the high side */
    double position_jitter_fraction;      /* layered: jitter in layer spacings, [0, 1] */
    double position_exponent;             /* gradient: positive power */
    /* layered biaxial: orientation_vector is the normal, orientation_angle the maximum tilt */
    double orientation_primary_fraction;
    double orientation_cross_fraction;
    double orientation_in_plane_deviation;
    uint64_t orientation_layer_seed;
} tgh_population_spec;

/* Writes spec->count fibers of (segments_per_fiber+1) vertices each:
 * intrinsic/placed [3 * count * (segments_per_fiber+1)], radius [count] and,
 * when `formation_layer` is not NULL, each fiber's layer [count]
 * (0xffffffff: none; set by TGH_POSITION_LAYERED).
 * Returns TGH_ERR_INVALID with "could not place fiber i" when the placement
 * search is exhausted (PopulationGenerationError::PlacementExhausted). */
int tgh_generate_population(const tgh_population_spec *spec, double *intrinsic, double *placed,
                            double *radius, uint32_t *formation_layer, size_t vertex_capacity);

#ifdef __cplusplus
}
#endif

#endif /* TANGLE_HIP_H */
