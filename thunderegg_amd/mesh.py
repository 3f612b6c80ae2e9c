"""Patch tables and the mesh JSON schema.

A level of a mesh is a numpy structured array of ``PATCH_DTYPE`` records, the
binary twin of ``te_patch_info`` in ``include/thunderegg_amd.h``.  Neighbour
slots follow the reference's ``PatchInfo::nbr_infos`` indexing,
``Face<D,M>::sum_of_faces + face value`` (reference ``Face.h:67-79``,
``PatchInfo.h:222-289``): 2D corners 0-3 and sides 4-7; 3D corners 0-7,
edges 8-19 and sides 20-25.

``read_mesh_json`` reads the reference's mesh format (``json.cpp:161-331``,
loaded the way ``test/utils/DomainReader.h:25-77`` does: ``lengths`` are divided
by the cells per side to give spacings); ``write_mesh_json`` writes it back.
"""
from __future__ import annotations

import json

import numpy as np

NBR_NONE, NBR_NORMAL, NBR_COARSE, NBR_FINE = -1, 0, 1, 2
NBR_NAMES = {"NORMAL": NBR_NORMAL, "COARSE": NBR_COARSE, "FINE": NBR_FINE}
NBR_NAMES_INV = {v: k for k, v in NBR_NAMES.items()}

PATCH_DTYPE = np.dtype(
    [
        ("id", "<i4"),
        ("rank", "<i4"),
        ("parent_id", "<i4"),
        ("parent_rank", "<i4"),
        ("orth_on_parent", "<i4"),
        ("refine_level", "<i4"),
        ("child_ids", "<i4", (8,)),
        ("child_ranks", "<i4", (8,)),
        ("starts", "<f8", (3,)),
        ("spacings", "<f8", (3,)),
        ("nbr_type", "<i4", (26,)),
        ("nbr_orth", "<i4", (26,)),
        ("nbr_ids", "<i4", (26, 4)),
        ("nbr_ranks", "<i4", (26, 4)),
    ],
    align=False,
)
assert PATCH_DTYPE.itemsize == 1176

# Face names of the reference's JSON schema (Face.cpp, Orthant.cpp).
SIDE_NAMES = {
    2: ["WEST", "EAST", "SOUTH", "NORTH"],
    3: ["WEST", "EAST", "SOUTH", "NORTH", "BOTTOM", "TOP"],
}
EDGE_NAMES = ["BS", "BN", "TS", "TN", "BW", "BE", "TW", "TE", "SW", "SE", "NW", "NE"]
CORNER_NAMES = {2: ["SW", "SE", "NW", "NE"], 3: ["BSW", "BSE", "BNW", "BNE", "TSW", "TSE", "TNW", "TNE"]}
ORTH_NAMES = {
    1: ["LOWER", "UPPER"],
    2: ["SW", "SE", "NW", "NE"],
    3: ["BSW", "BSE", "BNW", "BNE", "TSW", "TSE", "TNW", "TNE"],
}


def num_faces(dim: int, m: int) -> int:
    from math import comb

    return (1 << (dim - m)) * comb(dim, m)


def sum_of_faces(dim: int, m: int) -> int:
    return sum(num_faces(dim, k) for k in range(m))


def side_slot(dim: int, side: int) -> int:
    return sum_of_faces(dim, dim - 1) + side


def edge_slot(edge: int) -> int:
    return sum_of_faces(3, 1) + edge


def corner_slot(corner: int) -> int:
    return corner


def empty_patches(n: int) -> np.ndarray:
    a = np.zeros(n, dtype=PATCH_DTYPE)
    a["orth_on_parent"] = -1
    a["refine_level"] = -1
    a["parent_id"] = -1
    a["parent_rank"] = -1
    a["child_ids"] = -1
    a["child_ranks"] = -1
    a["nbr_type"] = NBR_NONE
    a["nbr_orth"] = -1
    a["nbr_ids"] = -1
    a["nbr_ranks"] = -1
    return a


def _orth(name, m):
    if name is None:
        return -1
    return ORTH_NAMES[m].index(name)


def _patch_from_json(pj: dict, dim: int, ns) -> np.ndarray:
    rec = empty_patches(1)[0]
    rec["id"] = pj["id"]
    rec["parent_id"] = pj["parent_id"]
    rec["parent_rank"] = pj["parent_rank"]
    rec["rank"] = pj["rank"]
    rec["orth_on_parent"] = _orth(pj.get("orth_on_parent"), dim)
    rec["refine_level"] = pj.get("refine_level", -1)
    for a in range(dim):
        rec["starts"][a] = pj["starts"][a]
        rec["spacings"][a] = pj["lengths"][a] / ns[a]
    if "child_ids" in pj:
        rec["child_ids"][: 1 << dim] = pj["child_ids"]
        rec["child_ranks"][: 1 << dim] = pj["child_ranks"]

    def put(slot, nj, m):
        t = NBR_NAMES[nj["type"]]
        rec["nbr_type"][slot] = t
        ids = nj["ids"]
        ranks = nj["ranks"]
        rec["nbr_ids"][slot][: len(ids)] = ids
        rec["nbr_ranks"][slot][: len(ranks)] = ranks
        if t == NBR_COARSE:
            rec["nbr_orth"][slot] = _orth(nj.get("orth_on_coarse"), m) if m > 0 else 0

    for nj in pj.get("nbrs", []):
        put(side_slot(dim, SIDE_NAMES[dim].index(nj["side"])), nj, dim - 1)
    if dim == 3:
        for nj in pj.get("edge_nbrs", []):
            put(edge_slot(EDGE_NAMES.index(nj["edge"])), nj, 1)
    for nj in pj.get("corner_nbrs", []):
        put(corner_slot(CORNER_NAMES[dim].index(nj["corner"])), nj, 0)
    return rec


def read_mesh_json(path: str, ns, rank: int | None = 0):
    """Read a reference-format mesh JSON.  Returns (dim, [level arrays]), finest first.

    Only patches owned by ``rank`` are kept (``None`` keeps all), as
    ``DomainReader`` does (reference ``test/utils/DomainReader.h:61-62``)."""
    with open(path) as fh:
        j = json.load(fh)
    levels_j = j["levels"]
    dim = len(levels_j[0][0]["starts"])
    levels = []
    for lj in levels_j:
        recs = [_patch_from_json(pj, dim, ns) for pj in lj if rank is None or pj["rank"] == rank]
        arr = empty_patches(len(recs))
        for i, r in enumerate(recs):
            arr[i] = r
        levels.append(arr)
    return dim, levels


def _nbr_to_json(rec, slot, dim, m):
    t = int(rec["nbr_type"][slot])
    cnt = (1 << m) if t == NBR_FINE else 1
    out = {
        "type": NBR_NAMES_INV[t],
        "ids": [int(v) for v in rec["nbr_ids"][slot][:cnt]],
        "ranks": [int(v) for v in rec["nbr_ranks"][slot][:cnt]],
    }
    if t == NBR_COARSE and m > 0 and rec["nbr_orth"][slot] >= 0:
        out["orth_on_coarse"] = ORTH_NAMES[m][int(rec["nbr_orth"][slot])]
    return out


def patch_to_json(rec, dim: int, ns) -> dict:
    """One patch in the reference schema (``json.cpp:161-247``)."""
    pj = {
        "id": int(rec["id"]),
        "parent_id": int(rec["parent_id"]),
        "parent_rank": int(rec["parent_rank"]),
        "rank": int(rec["rank"]),
        "starts": [float(rec["starts"][a]) for a in range(dim)],
        "lengths": [float(rec["spacings"][a] * ns[a]) for a in range(dim)],
        "refine_level": int(rec["refine_level"]),
    }
    if rec["orth_on_parent"] >= 0:
        pj["orth_on_parent"] = ORTH_NAMES[dim][int(rec["orth_on_parent"])]
    pj["nbrs"] = []
    for s in range(2 * dim):
        slot = side_slot(dim, s)
        if rec["nbr_type"][slot] != NBR_NONE:
            nj = _nbr_to_json(rec, slot, dim, dim - 1)
            nj["side"] = SIDE_NAMES[dim][s]
            pj["nbrs"].append(nj)
    if dim == 3:
        pj["edge_nbrs"] = []
        for e in range(12):
            slot = edge_slot(e)
            if rec["nbr_type"][slot] != NBR_NONE:
                nj = _nbr_to_json(rec, slot, dim, 1)
                nj["edge"] = EDGE_NAMES[e]
                pj["edge_nbrs"].append(nj)
    pj["corner_nbrs"] = []
    for c in range(1 << dim):
        if rec["nbr_type"][c] != NBR_NONE:
            nj = _nbr_to_json(rec, c, dim, 0)
            nj["corner"] = CORNER_NAMES[dim][c]
            pj["corner_nbrs"].append(nj)
    if rec["child_ids"][0] != -1:
        pj["child_ids"] = [int(v) for v in rec["child_ids"][: 1 << dim]]
        pj["child_ranks"] = [int(v) for v in rec["child_ranks"][: 1 << dim]]
    return pj


def write_mesh_json(path: str, dim: int, ns, levels) -> None:
    with open(path, "w") as fh:
        json.dump({"levels": [[patch_to_json(r, dim, ns) for r in lv] for lv in levels]}, fh)


def cell_centers(level: np.ndarray, dim: int, ns, with_ghost: bool = True):
    """Cell-centre coordinates of every padded cell, shape (P, [nz+2,] ny+2, nx+2, dim).

    ``DomainTools::GetRealCoordGhost`` (reference ``DomainTools.h:150-160``):
    start + h/2 + h*i for i in [-1, n]."""
    g = 1 if with_ghost else 0
    axes = []
    for a in range(dim):
        i = np.arange(-g, ns[a] + g, dtype=np.float64)
        axes.append(level["starts"][:, a][:, None] + level["spacings"][:, a][:, None] / 2.0
                    + level["spacings"][:, a][:, None] * i[None, :])
    P = len(level)
    if dim == 2:
        X = np.broadcast_to(axes[0][:, None, :], (P, ns[1] + 2 * g, ns[0] + 2 * g))
        Y = np.broadcast_to(axes[1][:, :, None], (P, ns[1] + 2 * g, ns[0] + 2 * g))
        return np.stack([X, Y], axis=-1)
    shp = (P, ns[2] + 2 * g, ns[1] + 2 * g, ns[0] + 2 * g)
    X = np.broadcast_to(axes[0][:, None, None, :], shp)
    Y = np.broadcast_to(axes[1][:, None, :, None], shp)
    Z = np.broadcast_to(axes[2][:, :, None, None], shp)
    return np.stack([X, Y, Z], axis=-1)


def padded_shape(num_patches: int, dim: int, ns, ncomp: int = 1):
    """Vector layout: [patch][component][z][y][x], each axis padded by one ghost cell
    (reference ``Vector.h:125-154``)."""
    return (num_patches, ncomp) + tuple(ns[a] + 2 for a in reversed(range(dim)))


def interior(v: np.ndarray, dim: int) -> np.ndarray:
    if dim == 2:
        return v[..., 1:-1, 1:-1]
    return v[..., 1:-1, 1:-1, 1:-1]
