"""The watertight mapping predicate against the reference's: which texels'
hit triangles differ, and whether those rays pass within an edge band of a
triangle edge (fornos_meshmap_set_predicate in include/fornos_hip.h).

A texel is in the band when the hit point of either predicate lies within
`band` x the scene's bounding-box diagonal of an edge of its own hit
triangle.  The reference's test accepts barycentrics down to -1e-5
(Shaders/meshmapping.comp:75-81), so near a shared edge it may take the
neighbouring triangle, or a triangle the ray only grazes; everywhere else the
two tests must pick the same triangle.  For texels with the same triangle the
hit distances are compared against the diagonal as well.
"""
import numpy as np

MISS = 0xFFFFFFFF


def edge_distance(coords, tidx, spos):
    """Distance of each hit point (barycentrics coords[:, 1:4] on soup
    triangle tidx / 3) to the nearest edge of its triangle; inf on a miss."""
    out = np.full(len(tidx), np.inf)
    hit = tidx != MISS
    if not hit.any():
        return out
    t = tidx[hit].astype(np.int64)
    a, b, c = (spos[t + k, :3].astype(np.float64) for k in range(3))
    twice_area = np.linalg.norm(np.cross(b - a, c - a), axis=1)
    opp = [np.linalg.norm(c - b, axis=1), np.linalg.norm(c - a, axis=1), np.linalg.norm(b - a, axis=1)]
    bc = coords[hit, 1:4].astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        d = np.stack([np.abs(bc[:, i]) * twice_area / opp[i] for i in range(3)], axis=1)
    out[hit] = np.nan_to_num(d.min(axis=1), nan=0.0)
    return out


def compare(ref, wt, scene, band=1e-6):
    """ref, wt: (coords, tidx) of the same texels mapped with the reference
    and the watertight predicate; scene: the fornos_amd.Scene they share."""
    (c0, t0), (c1, t1) = ref, wt
    spos, _ = scene.soup()
    p = spos[:, :3].astype(np.float64)
    diag = float(np.linalg.norm(p.max(0) - p.min(0)))
    eps = band * diag
    diff = t0 != t1
    idx = np.flatnonzero(diff)
    d0 = edge_distance(c0[idx], t0[idx], spos)
    d1 = edge_distance(c1[idx], t1[idx], spos)
    in_band = np.minimum(d0, d1) <= eps
    same = (~diff) & (t0 != MISS)
    dt = np.abs(c0[same, 0].astype(np.float64) - c1[same, 0].astype(np.float64))
    return {"texels": int(len(t0)), "hits_reference": int(np.sum(t0 != MISS)), "hits_watertight": int(np.sum(t1 != MISS)),
            "differing": int(diff.sum()), "differing_in_edge_band": int(in_band.sum()),
            "differing_outside_edge_band": int((~in_band).sum()), "edge_band": eps,
            "max_t_difference_over_diagonal": float(dt.max() / diag) if len(dt) else 0.0,
            "outside_band_texels": idx[~in_band][:16].tolist()}
