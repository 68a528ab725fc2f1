"""ORACLE — test infrastructure only.

Second, independent restatement (numpy float32, vectorised) of the broad
phase: the cell hash proxy_cell (crates/tangle_contact/src/device/cell_list.rs:32-79,
one proxy per segment), the ranked cell list (cell_list.rs:114-378) and the
neighbour-list predicate of build_segment_neighbor_lists
(crates/tangle_contact/src/device/neighbor.rs:155-364) evaluated by brute
force over every segment pair instead of the 27-cell stencil. numpy float32
arithmetic is IEEE single precision without fused multiply-add, as Rust's.
Use on small inputs only (O(S^2) memory).
"""

from __future__ import annotations

import numpy as np

f32 = np.float32


def rust_round(x: np.ndarray) -> np.ndarray:
    """f32::round: half away from zero (numpy's round is half to even)."""
    x = np.asarray(x, dtype=f32)
    t = np.trunc(x)
    frac = np.abs(x - t)
    return (t + np.where(frac >= f32(0.5), np.sign(x), f32(0.0))).astype(f32)


def sat_i32(x: np.ndarray) -> np.ndarray:
    x = np.asarray(x, dtype=np.float64)
    out = np.where(np.isnan(x), 0.0, np.clip(x, -2147483648.0, 2147483647.0))
    return out.astype(np.int64)


def segment_cells(pos, seg_vtx, lower, upper, cells, periodic):
    """Cell index of every segment midpoint, as proxy_cell with one proxy."""
    p = pos[seg_vtx[:, 0]].astype(f32)
    q = pos[seg_vtx[:, 1]].astype(f32)
    mid = (f32(0.5) * (p + q)).astype(f32)
    coords = []
    for k in range(3):
        width = f32(f32(upper[k]) - f32(lower[k])) / f32(cells[k])
        raw = sat_i32(np.floor((mid[:, k] - f32(lower[k])) / width))
        n = int(cells[k])
        c = np.clip(raw, 0, n - 1)
        if periodic[k]:
            c = np.mod(raw, n)  # ((raw % n) + n) % n for C-style remainders
        coords.append(c)
    return (coords[2] * cells[1] + coords[1]) * cells[0] + coords[0]


def cell_list(pos, seg_vtx, lower, upper, cells, periodic):
    """(counts, exclusive offsets, segments in ascending order per cell)."""
    cell = segment_cells(pos, seg_vtx, lower, upper, cells, periodic)
    C = int(cells[0] * cells[1] * cells[2])
    counts = np.bincount(cell, minlength=C).astype(np.uint32)
    offsets = (np.cumsum(counts) - counts).astype(np.uint32)
    order = np.lexsort((np.arange(cell.size), cell)).astype(np.uint32)
    return counts, offsets, order


def _clamp01(x):
    return np.minimum(np.maximum(x, f32(0.0)), f32(1.0)).astype(f32)


def axis_distance(p1, d1, p2, d2):
    """Vectorised proxy_pair_axis_distance (neighbor.rs:25-86), one proxy."""
    eps_len, eps_par = f32(1e-20), f32(1e-6)
    r = (p1 - p2).astype(f32)
    dot = lambda a, b: (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1] + a[..., 2] * b[..., 2]).astype(f32)
    a = dot(d1, d1)
    e = dot(d2, d2)
    f = dot(d2, r)
    c = dot(d1, r)
    b = dot(d1, d2)
    s = np.zeros_like(a)
    t = np.zeros_like(a)
    with np.errstate(all="ignore"):
        case1 = (a <= eps_len) & (e > eps_len)
        t = np.where(case1, _clamp01(f / e), t)
        case2 = (a > eps_len) & (e <= eps_len)
        s = np.where(case2, _clamp01(-c / a), s)
        case3 = (a > eps_len) & (e > eps_len)
        den = (a * e - b * b).astype(f32)
        nonpar = np.abs(den) > (eps_par * a * e).astype(f32)
        s3 = np.where(nonpar, _clamp01((b * f - c * e) / den), f32(0.0)).astype(f32)
        proj = ((b * s3 + f) / e).astype(f32)
        s3b = np.where(proj < 0, _clamp01(-c / a), np.where(proj > 1, _clamp01((b - c) / a), s3))
        t3 = np.where(proj < 0, f32(0.0), np.where(proj > 1, f32(1.0), proj))
        s = np.where(case3, s3b, s).astype(f32)
        t = np.where(case3, t3, t).astype(f32)
    q2 = (p2 + d2 * t[..., None]).astype(f32)
    q1 = (p1 + d1 * s[..., None]).astype(f32)
    delta = (q2 - q1).astype(f32)
    return np.sqrt(dot(delta, delta)).astype(f32)


def brute_force_lists(pos, seg_vtx, seg_fiber, seg_radius, arc, lower, upper, periodic, skin):
    """Every pair passing the neighbour-list predicate, by segment, ascending.

    The stencil only limits which pairs are examined; with a grid whose cells
    are at least one interaction span wide it never drops a pair that passes
    the predicate, so this brute-force set must equal the built lists."""
    pos = np.asarray(pos, dtype=f32)
    S = seg_vtx.shape[0]
    p = pos[seg_vtx[:, 0]]
    q = pos[seg_vtx[:, 1]]
    out = []
    lengths = [f32(f32(upper[k]) - f32(lower[k])) for k in range(3)]
    for i in range(S):
        p1, q1 = p[i], q[i]
        d1 = (q1 - p1).astype(f32)
        half = f32(0.5) * np.sqrt(f32(d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2]))
        p2 = p.copy()
        q2 = q.copy()
        for k in range(3):
            if periodic[k]:
                shift = (rust_round((f32(0.5) * (((p2[:, k] + q2[:, k]) - p1[k]) - q1[k])) / lengths[k])
                         * lengths[k]).astype(f32)
                p2[:, k] = p2[:, k] - shift
                q2[:, k] = q2[:, k] - shift
        d2 = (q2 - p2).astype(f32)
        a0, a1 = arc[seg_vtx[i, 0]], arc[seg_vtx[i, 1]]
        gap = np.maximum(arc[seg_vtx[:, 0]] - a1, a0 - arc[seg_vtx[:, 1]])
        near = (seg_fiber == seg_fiber[i]) & (gap < (seg_radius[i] + seg_radius).astype(f32))
        inter = ((seg_radius[i] + seg_radius + f32(skin)) * f32(1.0 + 1e-4)).astype(f32)
        md = (f32(0.5) * (((p2 + q2) - p1) - q1)).astype(f32)
        reach = (half + f32(0.5) * np.sqrt((d2[:, 0] * d2[:, 0] + d2[:, 1] * d2[:, 1] + d2[:, 2] * d2[:, 2])
                                           .astype(f32)) + inter).astype(f32)
        mid_ok = (md[:, 0] * md[:, 0] + md[:, 1] * md[:, 1] + md[:, 2] * md[:, 2]).astype(f32) <= reach * reach
        dist = axis_distance(np.broadcast_to(p1, p2.shape), np.broadcast_to(d1, d2.shape), p2, d2)
        ok = mid_ok & (dist <= inter) & ~near
        ok[i] = False
        out.append(np.nonzero(ok)[0].astype(np.uint32))
    return out
