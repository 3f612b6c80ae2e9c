"""CPU tests of the host side: the C ABI library loads and exports every function
include/thunderegg_amd.h declares (no device calls), the forest generator
reproduces the reference's mesh fixtures, the mesh JSON reader/writer round-trips
the reference schema, and the synthetic workloads are what they claim."""
import json
import os
import re

import numpy as np
import pytest

from helpers import MESHES, load
from thunderegg_amd import mesh
from thunderegg_amd import runtime as R
from thunderegg_amd import solver

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "thunderegg_amd.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(te_[a-z0-9_]+)\s*\(", text)))


def test_cabi_exports_every_declared_function():
    names = _declared_functions()
    assert len(names) > 60
    lib = R.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding types every one of them
    untyped = [n for n in names if n not in R.SIGNATURES]
    assert not untyped, untyped


def test_cabi_errors_without_device_work():
    """Argument errors are reported through the status code and te_last_error."""
    lib = R.lib()
    rc = lib.te_domain_create(2, None, 1, 0, None, None)
    assert rc == -1
    assert len(lib.te_last_error()) > 0


def _sig2(level):
    bystart = {(round(r["starts"][0], 6), round(r["starts"][1], 6), round(r["spacings"][0], 8)): r for r in level}
    id2 = {int(r["id"]): r for r in level}
    out = {}
    for k, r in bystart.items():
        items = []
        for slot in range(8):
            t = int(r["nbr_type"][slot])
            if t < 0:
                items.append(None)
                continue
            cnt = 2 if (t == 2 and slot >= 4) else 1
            geo = tuple((round(id2[int(i)]["starts"][0], 6), round(id2[int(i)]["starts"][1], 6))
                        for i in r["nbr_ids"][slot][:cnt])
            items.append((t, int(r["nbr_orth"][slot]) if (t == 1 and slot >= 4) else -1, geo))
        out[k] = (tuple(items), int(r["orth_on_parent"]))
    return out


def test_forest_matches_2d_cross_fixture():
    """The fixture 2d_uniform_8x8_refined_cross (reference test/mesh_inputs) regenerated:
    same patches on every level, same neighbour types, orthants and geometry."""
    ns = (10, 10)
    dim, fx = load("2d_uniform_8x8_refined_cross_mpi1.json", ns)
    f = solver.Forest(2, 3)
    lo, _ = f.leaf_boxes()
    fine = fx[0]
    targets = set((round(s[0], 6), round(s[1], 6)) for s, l in zip(fine["starts"], fine["spacings"][:, 0] * 10)
                  if abs(l - 1 / 16) < 1e-9)
    parents = set((round(np.floor(x * 8) / 8, 6), round(np.floor(y * 8) / 8, 6)) for x, y in targets)
    flags = np.array([(round(a, 6), round(b, 6)) in parents for a, b in lo], dtype=np.uint8)
    f.refine(flags)
    f.balance()
    mine = f.build_domains(ns)
    assert [len(l) for l in mine] == [len(l) for l in fx]
    for k in range(len(fx)):
        a, b = _sig2(mine[k]), _sig2(fx[k])
        assert a == b, k


def _sig3(level):
    bystart = {tuple(np.round(r["starts"][:3], 6)) + (round(r["spacings"][0], 8),): r for r in level}
    id2 = {int(r["id"]): r for r in level}
    out = {}
    for k, r in bystart.items():
        items = []
        for slot in range(26):
            t = int(r["nbr_type"][slot])
            if t < 0:
                continue
            m = 0 if slot < 8 else (1 if slot < 20 else 2)
            cnt = (1 << m) if t == 2 else 1
            geo = tuple(tuple(np.round(id2[int(i)]["starts"][:3], 6)) for i in r["nbr_ids"][slot][:cnt])
            items.append((slot, t, int(r["nbr_orth"][slot]) if (t == 1 and m > 0) else -1, geo))
        out[k] = (tuple(items), int(r["orth_on_parent"]))
    return out


@pytest.mark.parametrize("name,base", [("3d_refined_bnw_2x2x2_mpi1.json", 1), ("3d_mid_refine_4x4x4_mpi1.json", 2),
                                       ("3d_uniform_2x2x2_mpi1.json", 1)])
def test_forest_matches_3d_fixtures(name, base):
    ns = (4, 4, 4)
    dim, fx = load(name, ns)
    fine = fx[0]
    f = solver.Forest(3, base)
    hmin = (fine["spacings"][:, 0] * 4).min()
    lo, _ = f.leaf_boxes()
    small = [tuple(np.round(s, 6)) for s, l in zip(fine["starts"], fine["spacings"][:, 0] * 4)
             if abs(l - hmin) < 1e-9 and hmin < 1 / (1 << base) - 1e-9]
    par = set(tuple(np.round(np.floor(np.array(s) * (1 << base) + 1e-9) / (1 << base), 6)) for s in small)
    flags = np.array([tuple(np.round(b, 6)) in par for b in lo], dtype=np.uint8)
    f.refine(flags)
    f.balance()
    mine = f.build_domains(ns)
    assert len(mine) >= len(fx)
    for k in range(len(fx)):
        assert _sig3(mine[k]) == _sig3(fx[k]), k


def _morton_key(dim, starts, lmax):
    """z-order key of a cell's lower corner at the finest resolution 2^-lmax: the bits of
    the integer coordinates interleaved, x lowest (p4est's quadrant order in one tree)."""
    c = [int(round(starts[a] * (1 << lmax))) for a in range(dim)]
    key = 0
    for b in range(lmax):
        for a in range(dim):
            key |= ((c[a] >> b) & 1) << (b * dim + a)
    return key


@pytest.mark.parametrize("dim,base,seed", [(2, 2, 3), (2, 3, 4), (3, 1, 5), (3, 2, 6)])
def test_forest_ids_follow_p4est_rules(dim, base, seed):
    """Ids and patch order of every level follow the reference generator's rules: the
    finest level's ids are the leaves' positions in Morton order (SetIds,
    P4estDomainGenerator.cpp:180-198: an MPI_Scan prefix plus the position in p4est's
    z-order iteration, one rank); a coarser level merges each complete family into one
    patch that takes the id of its first (SW / orthant-0) child and lists the family's ids in
    orthant order, while a patch that is not merged keeps its id and is its own parent
    (coarsenTree, :300-319).  The key is computed here from the patch geometry alone."""
    rng = np.random.default_rng(seed)
    f = solver.Forest(dim, base)
    for _ in range(2):
        lv, _ = f.leaves()
        f.refine((rng.random(len(lv)) < 0.3).astype(np.uint8))
    f.balance()
    ns = (4,) * dim
    levels = f.build_domains(ns)
    h0 = np.array([r["spacings"][0] * ns[0] for r in levels[0]])
    lmax = int(round(-np.log2(h0.min())))
    keys0 = [_morton_key(dim, r["starts"], lmax) for r in levels[0]]
    assert list(levels[0]["id"]) == list(range(len(levels[0])))
    assert all(a < b for a, b in zip(keys0, keys0[1:]))
    for k in range(1, len(levels)):
        fine, coarse = levels[k - 1], levels[k]
        fid = {int(r["id"]): r for r in fine}
        keys = [_morton_key(dim, r["starts"], lmax) for r in coarse]
        assert all(a < b for a, b in zip(keys, keys[1:])), k
        for r in coarse:
            kids = [int(c) for c in r["child_ids"][: 1 << dim]]
            if kids[1] == -1:  # not merged: the same patch one level down
                q = fid[int(r["id"])]
                assert kids[0] == int(r["id"]) and int(q["parent_id"]) == int(r["id"])
                assert int(q["orth_on_parent"]) == -1
                assert np.allclose(q["starts"][:dim], r["starts"][:dim])
                continue
            assert int(r["id"]) == kids[0]
            for o, c in enumerate(kids):
                q = fid[c]
                assert int(q["parent_id"]) == int(r["id"]) and int(q["orth_on_parent"]) == o
                hq = q["spacings"][0] * ns[0]
                off = [(o >> a) & 1 for a in range(dim)]
                assert np.allclose(q["starts"][:dim], np.asarray(r["starts"][:dim]) + hq * np.array(off))


def test_forest_balance_is_2to1():
    """After balance(), face and corner neighbours differ by at most one level."""
    f = solver.Forest(2, 2)
    rng = np.random.default_rng(1)
    for _ in range(4):
        lv, _ = f.leaves()
        f.refine((rng.random(len(lv)) < 0.2).astype(np.uint8))
        f.balance()
    levels = f.build_domains((4, 4))
    fine = levels[0]
    rl = {int(r["id"]): int(r["refine_level"]) for r in fine}
    for r in fine:
        for slot in range(8):
            t = int(r["nbr_type"][slot])
            if t < 0:
                continue
            for i in r["nbr_ids"][slot][: (2 if (t == 2 and slot >= 4) else 1)]:
                assert abs(rl[int(i)] - int(r["refine_level"])) <= 1
    # every coarser level has one patch per complete family of the finer level
    for a, b in zip(levels[:-1], levels[1:]):
        assert len(b) < len(a)
        assert set(int(p) for p in a["parent_id"]) == set(int(i) for i in b["id"])
    assert len(levels[-1]) == 1


@pytest.mark.parametrize("name", sorted(os.listdir(MESHES)))
def test_mesh_json_round_trip(name, tmp_path):
    ns = (8, 8, 8)
    path = os.path.join(MESHES, name)
    dim = len(json.load(open(path))["levels"][0][0]["starts"])
    nsd = ns[:dim]
    dim, levels = mesh.read_mesh_json(path, nsd, rank=None)
    out = tmp_path / "m.json"
    mesh.write_mesh_json(str(out), dim, nsd, levels)
    dim2, levels2 = mesh.read_mesh_json(str(out), nsd, rank=None)
    assert dim2 == dim and len(levels2) == len(levels)
    for a, b in zip(levels, levels2):
        assert a.tobytes() == b.tobytes()
    # and the reference file's own content is preserved field by field
    j0 = json.load(open(path))
    j1 = json.load(open(out))
    for l0, l1 in zip(j0["levels"], j1["levels"]):
        for p0, p1 in zip(l0, l1):
            assert p0["id"] == p1["id"] and p0["parent_id"] == p1["parent_id"]
            assert sorted(n["side"] for n in p0.get("nbrs", [])) == sorted(n["side"] for n in p1.get("nbrs", []))


def test_uniform_levels_shape():
    lv = solver.uniform_levels(2, 3, (32, 32))
    assert [len(l) for l in lv] == [64, 16, 4, 1]
    p = lv[0]
    assert np.allclose(p["spacings"][:, :2], 1 / 256)
    assert (p["nbr_type"][:, 4:8] >= -1).all()


def _build_api_test(tmp, name="api_test"):
    """g++ build of tests/cpp/<name>.cpp against include/ThunderEgg (the reference-shaped
    C++ API) and the C ABI library; host compiler only."""
    import subprocess
    exe = os.path.join(str(tmp), name)
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", name + ".cpp"), "-L" + os.path.join(ROOT, "thunderegg_amd"),
           "-lthunderegg_amd", "-Wl,-rpath," + os.path.join(ROOT, "thunderegg_amd"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_cpp_api_compiles(tmp_path):
    _build_api_test(tmp_path)
    _build_api_test(tmp_path, "patch_api_test")
    # every forwarding header compiles on its own
    import subprocess
    inc = os.path.join(ROOT, "include")
    hdrs = sorted(os.path.relpath(os.path.join(dp, f), inc) for dp, _, fs in os.walk(os.path.join(inc, "ThunderEgg"))
                  for f in fs if f.endswith(".h"))
    src = tmp_path / "all.cpp"
    src.write_text("".join(f"#include <{h}>\n" for h in hdrs) + "int main() { return 0; }\n")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I" + inc, str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
