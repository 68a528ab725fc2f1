"""ORACLE — test infrastructure only.

Pure-Python restatement of the reference's seeded population generator,
generate_biased_fiber_population (crates/tangle_generate/src/population.rs:246-357)
with its helpers local_centerline (:554-565), sample_orientation (:567-655),
sample_contained_center (:657-726), sample_formation_layer (:728-740),
layer_reference_angle and splitmix64_unit (:742-752), perpendicular/normalize
(:754-790) and SplitMix64 (:795-813), plus maximum_polyline_curvature
(crates/tangle_core/src/curvature.rs:20-67). Python floats are IEEE doubles
and math.sin/cos/acos are the C library's, which is what Rust's f64 methods
call on Linux, so this restatement and the product's C++ generator must agree
bit for bit. Use it on small counts only (pure-Python loops).
"""

from __future__ import annotations

import math

MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK
        v = self.state
        v = ((v ^ (v >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        v = ((v ^ (v >> 27)) * 0x94D049BB133111EB) & MASK
        return v ^ (v >> 31)

    def unit(self) -> float:
        return float(self.next_u64() >> 11) * (1.0 / float(1 << 53))


def _sample(lo, hi, rng):
    return lo if lo == hi else lo + (hi - lo) * rng.unit()


def _cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def _scale(a, s):
    return [a[0] * s, a[1] * s, a[2] * s]


def _add(a, b):
    return [a[0] + b[0], a[1] + b[1], a[2] + b[2]]


def _norm(a):
    return math.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])


def _normalize(a):
    return _scale(a, 1.0 / _norm(a))


def _perpendicular(d):
    if abs(d[0]) <= abs(d[1]) and abs(d[0]) <= abs(d[2]):
        axis = [1.0, 0.0, 0.0]
    elif abs(d[1]) <= abs(d[2]):
        axis = [0.0, 1.0, 0.0]
    else:
        axis = [0.0, 0.0, 1.0]
    return _normalize(_cross(d, axis))


def max_polyline_curvature(points) -> float:
    best = 0.0
    eps = 2.220446049250313e-16
    for i in range(len(points) - 2):
        a, b, c = points[i], points[i + 1], points[i + 2]
        inc = [b[k] - a[k] for k in range(3)]
        out = [c[k] - b[k] for k in range(3)]
        li, lo = _norm(inc), _norm(out)
        if li <= eps or lo <= eps:
            continue
        cos = (inc[0] * out[0] + inc[1] * out[1] + inc[2] * out[2]) / (li * lo)
        cos = min(max(cos, -1.0), 1.0)
        angle = math.acos(cos)
        k = 2.0 * math.sin(0.5 * angle) / (0.5 * (li + lo))
        best = max(best, k)
    return best


def _splitmix64_unit(value: int) -> float:
    return SplitMix64(value).unit()


def _layer_reference_angle(seed: int, layer: int) -> float:
    return math.pi * _splitmix64_unit(seed ^ ((layer * 0x9E3779B97F4A7C15) & MASK))


def _direction(spec, rng, layer=None):
    kind = spec["orientation"]
    if kind == "layered_biaxial":
        n = _normalize(spec["vector"])
        a = _perpendicular(n)
        b = _normalize(_cross(n, a))
        reference = _layer_reference_angle(spec["layer_seed"], layer)
        selection = rng.unit()
        dev = spec["deviation"]
        if selection < spec["primary"]:
            azimuth = reference + dev * (2.0 * rng.unit() - 1.0)
        elif selection < spec["primary"] + spec["cross"]:
            azimuth = reference + 0.5 * math.pi + dev * (2.0 * rng.unit() - 1.0)
        else:
            azimuth = 2.0 * math.pi * rng.unit()
        tilt = spec["angle"] * (2.0 * rng.unit() - 1.0)
        return _add(_scale(_add(_scale(a, math.cos(azimuth)), _scale(b, math.sin(azimuth))), math.cos(tilt)),
                    _scale(n, math.sin(tilt)))
    if kind == "planar":
        n = _normalize(spec["vector"])
        a = _perpendicular(n)
        b = _normalize(_cross(n, a))
        azimuth = 2.0 * math.pi * rng.unit()
        tilt = spec["angle"] * (2.0 * rng.unit() - 1.0)
        return _add(_scale(_add(_scale(a, math.cos(azimuth)), _scale(b, math.sin(azimuth))), math.cos(tilt)),
                    _scale(n, math.sin(tilt)))
    if kind == "aligned":
        axis = _normalize(spec["vector"])
        a = _perpendicular(axis)
        b = _normalize(_cross(axis, a))
        mc = math.cos(spec["angle"])
        c = mc + (1.0 - mc) * rng.unit()
        s = math.sqrt(max(1.0 - c * c, 0.0))
        azimuth = 2.0 * math.pi * rng.unit()
        return _add(_scale(axis, c), _scale(_add(_scale(a, math.cos(azimuth)), _scale(b, math.sin(azimuth))), s))
    if kind == "schladitz":
        axis = _normalize(spec["vector"])
        a = _perpendicular(axis)
        b = _normalize(_cross(axis, a))
        beta2 = spec["beta"] * spec["beta"]
        p = rng.unit()
        c = min(p / math.sqrt(beta2 - (beta2 - 1.0) * p * p), 1.0)
        if rng.unit() < 0.5:
            c = -c
        s = math.sqrt(1.0 - c * c)
        phi = 2.0 * math.pi * rng.unit()
        return _add(_scale(axis, c), _scale(_add(_scale(a, math.cos(phi)), _scale(b, math.sin(phi))), s))
    z = 2.0 * rng.unit() - 1.0
    azimuth = 2.0 * math.pi * rng.unit()
    radial = math.sqrt(1.0 - z * z)
    return [radial * math.cos(azimuth), radial * math.sin(azimuth), z]


def generate(spec: dict):
    """spec keys: count, segments, seed, length (lo, hi), radius (lo, hi),
    amplitude (lo, hi), orientation ('isotropic'|'planar'|'aligned'|'schladitz'),
    vector, angle, beta, min_bend_radius (None or r), max_attempts, origin,
    lengths, periodic; for 'layered_biaxial' also primary, cross, deviation,
    layer_seed; optional position ('uniform'|'layered'|'gradient') with
    position_axis, layers, jitter, exponent, toward_high. Returns
    (intrinsic, placed, radius, layers) as lists; layers hold None for
    unlayered fibers."""
    rng = SplitMix64(spec["seed"])
    n = spec["segments"]
    lo_box = list(spec["origin"])
    hi_box = [spec["origin"][k] + spec["lengths"][k] for k in range(3)]
    intrinsic_all, placed_all, radii, layers = [], [], [], []
    position = spec.get("position", "uniform")
    for fiber in range(spec["count"]):
        layer = None
        if position == "layered":
            count = spec["layers"]
            layer = int(min(math.floor(rng.unit() * count), count - 1.0))
        done = False
        for _ in range(spec.get("max_attempts", 128)):
            length = _sample(*spec["length"], rng)
            radius = _sample(*spec["radius"], rng)
            amplitude = _sample(*spec["amplitude"], rng)
            hand = -1.0 if rng.unit() < 0.5 else 1.0
            local = []
            for i in range(n + 1):
                c = i / n - 0.5
                local.append([c * length, amplitude * math.sin(2.0 * math.pi * c),
                              0.5 * amplitude * hand * math.sin(4.0 * math.pi * c)])
            mbr = spec.get("min_bend_radius")
            if mbr and max_polyline_curvature(local) > (1.0 / mbr) * (1.0 + 1.0e-10):
                continue
            d = _direction(spec, rng, layer)
            ta = _perpendicular(d)
            tb = _normalize(_cross(d, ta))
            offsets = [_add(_add(_scale(d, p[0]), _scale(ta, p[1])), _scale(tb, p[2])) for p in local]
            lower, upper = [0.0] * 3, [0.0] * 3
            feasible = True
            for k in range(3):
                if spec["periodic"][k]:
                    lower[k], upper[k] = lo_box[k], hi_box[k]
                    continue
                mn = math.inf
                mx = -math.inf
                for o in offsets:
                    mn = min(mn, o[k])
                    mx = max(mx, o[k])
                lower[k] = lo_box[k] + radius - mn
                upper[k] = hi_box[k] - radius - mx
                if lower[k] > upper[k]:
                    feasible = False
                    break
            if not feasible:
                continue
            center = [lower[k] + (upper[k] - lower[k]) * rng.unit() for k in range(3)]
            if position == "layered":
                k = spec["position_axis"]
                spacing = (upper[k] - lower[k]) / spec["layers"]
                jitter = spec["jitter"] * (rng.unit() - 0.5) * spacing
                center[k] = min(max(lower[k] + (layer + 0.5) * spacing + jitter, lower[k]), upper[k])
            elif position == "gradient":
                k = spec["position_axis"]
                u = rng.unit()
                biased = 1.0 - math.pow(1.0 - u, spec["exponent"]) if spec["toward_high"] else \
                    math.pow(u, spec["exponent"])
                center[k] = lower[k] + (upper[k] - lower[k]) * biased
            intrinsic_all.extend(local)
            placed_all.extend(_add(center, o) for o in offsets)
            radii.append(radius)
            layers.append(layer)
            done = True
            break
        if not done:
            raise ValueError(f"could not place fiber {fiber} within max_attempts_per_fiber")
    return intrinsic_all, placed_all, radii, layers
