// Host-side state of one world, shared by the relaxation and learned
// warm-start host sources.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "relax_device.h"

namespace tgh {

// Host copy of the packed assembly in the reference's f32 representation.
struct Packed {
    uint32_t F = 0, V = 0, S = 0;
    std::vector<float> pos;        // 3V
    std::vector<float> intr;       // 3V
    std::vector<uint32_t> fstart;  // F
    std::vector<uint32_t> fcount;  // F
    std::vector<float> radius;     // F
    std::vector<float> rest;       // S
    std::vector<float> kmax;       // V
    std::vector<uint8_t> pinned;   // V
    std::vector<float> arc;        // V
    std::vector<uint32_t> fstep;   // F: formation step
    std::vector<uint32_t> flayer;  // F: formation layer (0xffffffff: none)
    float lower[3], upper[3];
    uint32_t periodic[3];
};

struct Sizing {
    float skin = 0.f, cell_size = 0.f;
    float segment_cell_size = 0.f;  // whole-segment grid of the contact capture (world.rs:405)
    float max_radius = 0.f;
    float list_reach = 0.f;  // the list test's reach for a pair of mean half length plus radius, at creation
    uint32_t cells[3] = {1, 1, 1};
    uint32_t C = 1;
    size_t slots = 1;
    uint32_t block = 48;
};

}  // namespace tgh

struct tgh_world {
    int device = 0;
    hipStream_t stream = nullptr;
    // specimens relaxed together, each with its own grid; their fibers,
    // vertices, segments and cells follow one another on the device
    std::vector<tgh::Packed> specimens;
    std::vector<tgh::Sizing> sizes;
    std::vector<uint32_t> vbase, fbase, sbase, cbase;  // [nspec + 1] prefix offsets
    std::vector<tgh::GridParams> grids;
    tgh::WorldDev d{};
    tgh::SpecState *host_spec = nullptr;  // pinned [nspec], read after every batch
    tgh_config last_config{};
    bool ran = false;
    // formation targets (installed on the device, applied every iteration)
    tgh::FormationDev formation{};
    uint32_t formation_layer_capacity = 0, formation_vertex_capacity = 0;
    uint64_t device_bytes = 0;
    std::vector<void *> allocations;
    std::map<std::string, hipGraphExec_t> graphs;
    std::vector<hipGraph_t> graph_templates;
    bool use_graphs = true;
    // learned warm start: features, predictions and the pre-proposal positions
    uint16_t *proposal_x = nullptr;
    float4 *proposal_y = nullptr;
    float4 *proposal_backup = nullptr;
    float *proposal_scale = nullptr;  // per segment: history length (m) x diameter
    float *metrics_scratch = nullptr;  // compaction metrics: 13 outputs, then 5 S + 7 V terms
    float4 *history = nullptr;        // per segment: midpoint at the last history mark
    bool history_valid = false;
    bool proposal_pending = false;
    // Kernel-group timing: events bracket every group of every iteration of
    // a batch, launched directly on the stream.
    bool timing = false;
    std::vector<hipEvent_t> direct_events;
    double group_ms[tgh::kGroupCount] = {0, 0, 0, 0, 0};
    uint64_t group_launches[tgh::kGroupCount] = {0, 0, 0, 0, 0};

    template <typename T>
    hipError_t alloc(T **out, size_t count) {
        size_t bytes = std::max<size_t>(1, count) * sizeof(T);
        hipError_t e = hipMalloc((void **)out, bytes);
        if (e == hipSuccess) {
            allocations.push_back(*out);
            device_bytes += bytes;
        }
        return e;
    }

    ~tgh_world() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
        for (hipGraph_t g : graph_templates) (void)hipGraphDestroy(g);
        for (hipEvent_t e : direct_events) (void)hipEventDestroy(e);
        for (void *p : allocations) (void)hipFree(p);
        if (host_spec) (void)hipHostFree(host_spec);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

