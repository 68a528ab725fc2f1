// Seeded biased fiber population generator.
//
// Follows generate_biased_fiber_population
// (crates/tangle_generate/src/population.rs:246-357): per fiber, draws the
// formation layer of a layered position (sample_formation_layer :728-740),
// then per attempt (length, radius, amplitude, handedness) with
// ScalarDistribution::sample (constants draw nothing, :27-32), builds the
// local sinusoidal centreline (local_centerline :554-565), rejects it when its
// discrete curvature exceeds the bend limit (maximum_polyline_curvature,
// crates/tangle_core/src/curvature.rs:20-67), samples a direction
// (sample_orientation :567-655), rotates the centreline onto it and samples
// a contained center (sample_contained_center :657-726) with its uniform,
// layered or density-gradient coordinate along the position axis. SplitMix64 is the
// published generator (:795-813). f64 throughout; sin/cos/acos/pow are the C
// library's, the functions Rust's f64 methods call on Linux.
//
// TGH_ORIENT_SCHLADITZ has no sampler in the reference. It draws |cos θ|
// from the inverse CDF of the Schladitz density
// β / (4π (1 + (β² − 1) cos² θ)^{3/2}) about `axis` (quantile as
// schladitz_axis_cosine_quantile, crates/tangle_characterize/src/shape.rs:552-556),
// a sign and a uniform azimuth, in the draw order of the reference's own
// model-sampling test (shape.rs:788-797): quantile draw, sign draw, azimuth.

#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tangle_hip.h"
#include "errors.h"

namespace {

struct SplitMix64 {
    uint64_t state;
    double unit() {
        state += 0x9E3779B97F4A7C15ull;
        uint64_t v = state;
        v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
        v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
        v ^= v >> 31;
        return (double)(v >> 11) * (1.0 / (double)(uint64_t(1) << 53));
    }
};

// splitmix64_unit (population.rs:746-752): one SplitMix64 step from `value`.
inline double splitmix64_unit(uint64_t value) {
    SplitMix64 r{value};
    return r.unit();
}

// layer_reference_angle (population.rs:742-744)
inline double layer_reference_angle(uint64_t seed, uint32_t layer) {
    return M_PI * splitmix64_unit(seed ^ ((uint64_t)layer * 0x9E3779B97F4A7C15ull));
}

inline double sample_scalar(double lo, double hi, SplitMix64 &r) {
    if (lo == hi) return lo;  // ScalarDistribution::Constant draws nothing
    return lo + (hi - lo) * r.unit();
}

inline std::array<double, 3> add3(std::array<double, 3> a, std::array<double, 3> b) {
    return {a[0] + b[0], a[1] + b[1], a[2] + b[2]};
}
inline std::array<double, 3> scale3(std::array<double, 3> a, double s) { return {a[0] * s, a[1] * s, a[2] * s}; }
inline std::array<double, 3> cross3(std::array<double, 3> a, std::array<double, 3> b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
inline double norm3(std::array<double, 3> a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
inline std::array<double, 3> normalize3(std::array<double, 3> a) { return scale3(a, 1.0 / norm3(a)); }
inline double dot3(std::array<double, 3> a, std::array<double, 3> b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

std::array<double, 3> perpendicular(std::array<double, 3> d) {
    std::array<double, 3> axis;
    if (std::fabs(d[0]) <= std::fabs(d[1]) && std::fabs(d[0]) <= std::fabs(d[2]))
        axis = {1.0, 0.0, 0.0};
    else if (std::fabs(d[1]) <= std::fabs(d[2]))
        axis = {0.0, 1.0, 0.0};
    else
        axis = {0.0, 0.0, 1.0};
    return normalize3(cross3(d, axis));
}

// maximum over internal vertices of 2 sin(θ/2) / dual length (curvature.rs:20-67)
double max_polyline_curvature(const std::vector<std::array<double, 3>> &pts) {
    double best = 0.0;
    const double eps = 2.220446049250313e-16;
    for (size_t i = 0; i + 2 < pts.size(); ++i) {
        std::array<double, 3> in = {pts[i + 1][0] - pts[i][0], pts[i + 1][1] - pts[i][1], pts[i + 1][2] - pts[i][2]};
        std::array<double, 3> out = {pts[i + 2][0] - pts[i + 1][0], pts[i + 2][1] - pts[i + 1][1],
                                     pts[i + 2][2] - pts[i + 1][2]};
        double li = norm3(in), lo = norm3(out);
        if (li <= eps || lo <= eps) continue;
        double c = dot3(in, out) / (li * lo);
        if (c < -1.0) c = -1.0;
        if (c > 1.0) c = 1.0;
        double angle = std::acos(c);
        double dual = 0.5 * (li + lo);
        double k = 2.0 * std::sin(0.5 * angle) / dual;
        best = std::fmax(best, k);
    }
    return best;
}

std::array<double, 3> sample_direction(const tgh_population_spec &s, uint32_t layer, SplitMix64 &r) {
    const double pi = M_PI;
    switch (s.orientation) {
        case TGH_ORIENT_LAYERED_BIAXIAL: {
            auto n = normalize3({s.orientation_vector[0], s.orientation_vector[1], s.orientation_vector[2]});
            auto a = perpendicular(n);
            auto b = normalize3(cross3(n, a));
            const double reference = layer_reference_angle(s.orientation_layer_seed, layer);
            const double selection = r.unit();
            double azimuth;
            if (selection < s.orientation_primary_fraction)
                azimuth = reference + s.orientation_in_plane_deviation * (2.0 * r.unit() - 1.0);
            else if (selection < s.orientation_primary_fraction + s.orientation_cross_fraction)
                azimuth = reference + 0.5 * pi + s.orientation_in_plane_deviation * (2.0 * r.unit() - 1.0);
            else
                azimuth = 2.0 * pi * r.unit();
            double tilt = s.orientation_angle * (2.0 * r.unit() - 1.0);
            return add3(scale3(add3(scale3(a, std::cos(azimuth)), scale3(b, std::sin(azimuth))), std::cos(tilt)),
                        scale3(n, std::sin(tilt)));
        }
        case TGH_ORIENT_PLANAR: {
            auto n = normalize3({s.orientation_vector[0], s.orientation_vector[1], s.orientation_vector[2]});
            auto a = perpendicular(n);
            auto b = normalize3(cross3(n, a));
            double azimuth = 2.0 * pi * r.unit();
            double tilt = s.orientation_angle * (2.0 * r.unit() - 1.0);
            return add3(scale3(add3(scale3(a, std::cos(azimuth)), scale3(b, std::sin(azimuth))), std::cos(tilt)),
                        scale3(n, std::sin(tilt)));
        }
        case TGH_ORIENT_ALIGNED: {
            auto axis = normalize3({s.orientation_vector[0], s.orientation_vector[1], s.orientation_vector[2]});
            auto a = perpendicular(axis);
            auto b = normalize3(cross3(axis, a));
            double min_cos = std::cos(s.orientation_angle);
            double cosine = min_cos + (1.0 - min_cos) * r.unit();
            double sine = std::sqrt(std::fmax(1.0 - cosine * cosine, 0.0));
            double azimuth = 2.0 * pi * r.unit();
            return add3(scale3(axis, cosine),
                        scale3(add3(scale3(a, std::cos(azimuth)), scale3(b, std::sin(azimuth))), sine));
        }
        case TGH_ORIENT_SCHLADITZ: {
            auto axis = normalize3({s.orientation_vector[0], s.orientation_vector[1], s.orientation_vector[2]});
            auto a = perpendicular(axis);
            auto b = normalize3(cross3(axis, a));
            double beta2 = s.orientation_beta * s.orientation_beta;
            double p = r.unit();
            double c = std::fmin(p / std::sqrt(beta2 - (beta2 - 1.0) * p * p), 1.0);
            if (r.unit() < 0.5) c = -c;
            double sine = std::sqrt(1.0 - c * c);
            double phi = 2.0 * pi * r.unit();
            return add3(scale3(axis, c), scale3(add3(scale3(a, std::cos(phi)), scale3(b, std::sin(phi))), sine));
        }
        default: {  // Isotropic3d
            double z = 2.0 * r.unit() - 1.0;
            double azimuth = 2.0 * pi * r.unit();
            double radial = std::sqrt(1.0 - z * z);
            return {radial * std::cos(azimuth), radial * std::sin(azimuth), z};
        }
    }
}

bool unit_interval(double v) { return std::isfinite(v) && v >= 0.0 && v <= 1.0; }
bool quarter_turn(double v) { return std::isfinite(v) && v >= 0.0 && v <= 0.5 * M_PI; }

int generate(const tgh_population_spec *spec, double *intrinsic, double *placed, double *radius_out,
             uint32_t *layer_out, size_t vertex_capacity) {
    if (!spec || !intrinsic || !placed || !radius_out) {
        return tgh::fail(TGH_ERR_INVALID, "null argument");
    }
    const tgh_population_spec &s = *spec;
    if (s.count == 0) {
        return tgh::fail(TGH_ERR_INVALID, "fiber population count must be positive");
    }
    if (s.segments_per_fiber < 2) {
        return tgh::fail(TGH_ERR_INVALID, "at least two segments per fiber are required, got " + std::to_string(s.segments_per_fiber));
    }
    if (!(s.length_min > 0.0) || s.length_min > s.length_max || !(s.radius_min > 0.0) || s.radius_min > s.radius_max ||
        s.amplitude_min < 0.0 || s.amplitude_min > s.amplitude_max) {
        return tgh::fail(TGH_ERR_INVALID, "invalid length, radius or amplitude distribution");
    }
    if (s.orientation == TGH_ORIENT_SCHLADITZ && !(s.orientation_beta > 0.0 && std::isfinite(s.orientation_beta))) {
        return tgh::fail(TGH_ERR_INVALID, "Schladitz beta must be positive and finite");
    }
    // validate_spec (population.rs:392-492)
    if (s.orientation == TGH_ORIENT_LAYERED_BIAXIAL) {
        if (!(s.orientation_primary_fraction >= 0.0) || !(s.orientation_cross_fraction >= 0.0) ||
            !std::isfinite(s.orientation_primary_fraction) || !std::isfinite(s.orientation_cross_fraction) ||
            s.orientation_primary_fraction + s.orientation_cross_fraction > 1.0)
            return tgh::fail(TGH_ERR_INVALID, "layered biaxial fractions must be non-negative with sum <= 1");
        if (!quarter_turn(s.orientation_in_plane_deviation) || !quarter_turn(s.orientation_angle))
            return tgh::fail(TGH_ERR_INVALID, "layered biaxial angles must lie in [0, pi/2]");
        if (s.position != TGH_POSITION_LAYERED)
            return tgh::fail(TGH_ERR_INVALID, "layered biaxial orientation requires a layered position");
    } else if (s.orientation > TGH_ORIENT_LAYERED_BIAXIAL) {
        return tgh::fail(TGH_ERR_INVALID, "unknown orientation distribution");
    }
    if (s.position == TGH_POSITION_LAYERED) {
        if (s.position_axis > 2 || s.position_layers == 0 || !unit_interval(s.position_jitter_fraction))
            return tgh::fail(TGH_ERR_INVALID, "invalid layered position distribution");
    } else if (s.position == TGH_POSITION_GRADIENT) {
        if (s.position_axis > 2 || !(s.position_exponent > 0.0) || !std::isfinite(s.position_exponent))
            return tgh::fail(TGH_ERR_INVALID, "invalid density-gradient position distribution");
    } else if (s.position != TGH_POSITION_UNIFORM) {
        return tgh::fail(TGH_ERR_INVALID, "unknown position distribution");
    }
    for (int k = 0; k < 3; ++k)
        if (!(s.cell_lengths[k] > 0.0)) {
            return tgh::fail(TGH_ERR_INVALID, "biased-box generation requires a positive orthorhombic cell");
        }
    const uint32_t nv = s.segments_per_fiber + 1;
    if (vertex_capacity < (size_t)s.count * nv) {
        return tgh::fail(TGH_ERR_SIZE, "output buffers hold fewer than count*(segments_per_fiber+1) vertices");
    }
    const double pi = M_PI;
    double box_low[3], box_high[3];
    for (int k = 0; k < 3; ++k) {
        box_low[k] = s.cell_origin[k];
        box_high[k] = s.cell_origin[k] + s.cell_lengths[k];
    }
    SplitMix64 rng{s.seed};
    std::vector<std::array<double, 3>> local(nv), offsets(nv);
    const uint32_t attempts = s.max_attempts_per_fiber ? s.max_attempts_per_fiber : 128u;
    for (uint32_t fiber = 0; fiber < s.count; ++fiber) {
        uint32_t layer = 0xffffffffu;  // sample_formation_layer (population.rs:728-740)
        if (s.position == TGH_POSITION_LAYERED)
            layer = (uint32_t)std::fmin(std::floor(rng.unit() * (double)s.position_layers),
                                        (double)s.position_layers - 1.0);
        bool placed_ok = false;
        for (uint32_t attempt = 0; attempt < attempts && !placed_ok; ++attempt) {
            double length = sample_scalar(s.length_min, s.length_max, rng);
            double radius = sample_scalar(s.radius_min, s.radius_max, rng);
            double amplitude = sample_scalar(s.amplitude_min, s.amplitude_max, rng);
            double hand = rng.unit() < 0.5 ? -1.0 : 1.0;
            for (uint32_t i = 0; i < nv; ++i) {
                double centered = (double)i / (double)s.segments_per_fiber - 0.5;
                local[i] = {centered * length, amplitude * std::sin(2.0 * pi * centered),
                            0.5 * amplitude * hand * std::sin(4.0 * pi * centered)};
            }
            if (s.min_bend_radius > 0.0 &&
                max_polyline_curvature(local) > (1.0 / s.min_bend_radius) * (1.0 + 1.0e-10))
                continue;
            auto dir = sample_direction(s, layer, rng);
            auto ta = perpendicular(dir);
            auto tb = normalize3(cross3(dir, ta));
            for (uint32_t i = 0; i < nv; ++i)
                offsets[i] = add3(add3(scale3(dir, local[i][0]), scale3(ta, local[i][1])), scale3(tb, local[i][2]));
            double lower[3], upper[3];
            bool feasible = true;
            for (int k = 0; k < 3 && feasible; ++k) {
                if (s.cell_periodic[k]) {
                    lower[k] = box_low[k];
                    upper[k] = box_high[k];
                    continue;
                }
                double mn = INFINITY, mx = -INFINITY;
                for (uint32_t i = 0; i < nv; ++i) {
                    mn = std::fmin(mn, offsets[i][k]);
                    mx = std::fmax(mx, offsets[i][k]);
                }
                lower[k] = box_low[k] + radius - mn;
                upper[k] = box_high[k] - radius - mx;
                if (lower[k] > upper[k]) feasible = false;
            }
            if (!feasible) continue;
            double center[3];
            for (int k = 0; k < 3; ++k) center[k] = lower[k] + (upper[k] - lower[k]) * rng.unit();
            if (s.position == TGH_POSITION_LAYERED) {
                const uint32_t k = s.position_axis;
                const double spacing = (upper[k] - lower[k]) / (double)s.position_layers;
                const double jitter = s.position_jitter_fraction * (rng.unit() - 0.5) * spacing;
                center[k] = std::fmin(std::fmax(lower[k] + ((double)layer + 0.5) * spacing + jitter, lower[k]),
                                      upper[k]);
            } else if (s.position == TGH_POSITION_GRADIENT) {
                const uint32_t k = s.position_axis;
                const double u = rng.unit();
                const double biased = s.position_toward_high ? 1.0 - std::pow(1.0 - u, s.position_exponent)
                                                             : std::pow(u, s.position_exponent);
                center[k] = lower[k] + (upper[k] - lower[k]) * biased;
            }
            const size_t base = (size_t)fiber * nv;
            for (uint32_t i = 0; i < nv; ++i)
                for (int k = 0; k < 3; ++k) {
                    intrinsic[3 * (base + i) + k] = local[i][k];
                    placed[3 * (base + i) + k] = center[k] + offsets[i][k];
                }
            radius_out[fiber] = radius;
            if (layer_out) layer_out[fiber] = layer;
            placed_ok = true;
        }
        if (!placed_ok) {
            return tgh::fail(TGH_ERR_INVALID, "could not place fiber " + std::to_string(fiber) + " within max_attempts_per_fiber");
        }
    }
    return TGH_OK;
}

}  // namespace

extern "C" int tgh_generate_population(const tgh_population_spec *spec, double *intrinsic, double *placed,
                                       double *radius, uint32_t *formation_layer, size_t vertex_capacity) {
    return generate(spec, intrinsic, placed, radius, formation_layer, vertex_capacity);
}
