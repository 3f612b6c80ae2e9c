"""Timer on the host clock, through the C ABI (no device work).

The cases restate the reference's Timer tests (test/Timer_MPI1.cpp): the error cases,
the JSON shape of to_json, the line counts of the text report and saveToFile.  The
domains a timer carries are checked against the mesh writer (mesh.patch_to_json) and
the reference's own mesh files.  Stream timings are covered by the GPU tests.
"""
import json
import os
import types

import numpy as np
import pytest

from thunderegg_amd import mesh
from thunderegg_amd import runtime as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MESHES = os.path.join(ROOT, "tests", "golden", "mesh_inputs")


def _domain():
    """GetDomain of the reference's Timer tests: one 10x10 patch, id 0, spacing 0.01."""
    p = mesh.empty_patches(1)
    p["id"] = 0
    p["rank"] = 0
    p["spacings"] = 0.01
    return types.SimpleNamespace(dim=2, ns=(10, 10), patches=p)


def _occ(s, target):
    n, pos = 0, 0
    while True:
        pos = s.find(target, pos)
        if pos < 0:
            return n
        n += 1
        pos += len(target)


def _two_domains_nested(t, c_in_second=None):
    t.add_domain(0, _domain())
    t.add_domain(1, _domain())
    t.start_domain_timing(0, "A")
    if c_in_second is False:
        t.start_domain_timing(0, "C")
        t.stop_domain_timing(0, "C")
    t.start_domain_timing(0, "B")
    t.stop_domain_timing(0, "B")
    t.stop_domain_timing(0, "A")
    t.start_domain_timing(1, "A")
    if c_in_second:
        t.start_domain_timing(0, "C")
        t.stop_domain_timing(0, "C")
    t.start_domain_timing(1, "B")
    t.stop_domain_timing(1, "B")
    t.stop_domain_timing(1, "A")


# ---- errors (Timer_MPI1.cpp:60-180) ----

def test_stop_before_start_and_wrong_order():
    t = R.Timer()
    t.start("A")
    t.stop("A")
    with pytest.raises(R.TEError, match="expecting to end"):
        t.stop("B")
    with pytest.raises(R.TEError):
        t.stop("")  # the root is never stopped
    t.start("A")
    t.start("B")
    with pytest.raises(R.TEError, match='expecting to end "B", instead got "A"'):
        t.stop("A")


def test_report_fails_with_unfinished_timing():
    t = R.Timer()
    t.start("A")
    with pytest.raises(R.TEError, match="unfinished"):
        t.report()


def test_domain_timing_errors():
    t = R.Timer()
    t.add_domain(0, _domain())
    t.add_domain(1, _domain())
    with pytest.raises(R.TEError, match="already added"):
        t.add_domain(0, _domain())
    with pytest.raises(R.TEError, match="not added"):
        t.start_domain_timing(5, "A")
    t.start_domain_timing(0, "A")
    for dom, name in ((1, "A"), (0, "B"), (1, "B")):
        with pytest.raises(R.TEError):
            t.stop_domain_timing(dom, name)
    with pytest.raises(R.TEError):
        t.stop("A")  # same name, no domain
    t.stop_domain_timing(0, "A")
    t.start_patch_timing(0, 0, "P")
    with pytest.raises(R.TEError):
        t.stop_patch_timing(1, 0, "P")
    t.stop_patch_timing(0, 0, "P")


def test_info_errors():
    t = R.Timer()
    with pytest.raises(R.TEError, match="No timing"):
        t.add_int_info("Example", 1)
    with pytest.raises(R.TEError, match="No timing"):
        t.add_double_info("Example", 1.0)
    t.start("A")
    t.add_int_info("I", 1)
    t.add_double_info("D", 1.0)
    with pytest.raises(R.TEError):
        t.add_double_info("I", 1.0)
    with pytest.raises(R.TEError):
        t.add_int_info("D", 1)
    t.stop("A")


# ---- to_json (Timer_MPI1.cpp:181-514) ----

def test_to_json_empty():
    assert R.Timer().to_json() is None


def _check_timing(tj, name, calls):
    assert tj["name"] == name and tj["rank"] == 0 and tj["num_calls"] == calls
    for k in ("min", "max", "sum"):
        assert isinstance(tj[k], float)
    assert tj["min"] <= tj["sum"] / calls <= tj["max"]


def test_to_json_unassociated_and_infos():
    t = R.Timer()
    t.start("A")
    t.stop("A")
    j = t.to_json()
    assert set(j) == {"comm_size", "timings"} and j["comm_size"] == 1
    assert len(j["timings"]) == 1
    _check_timing(j["timings"][0], "A", 1)
    assert "infos" not in j["timings"][0] and "domain_id" not in j["timings"][0]

    t = R.Timer()
    t.start("A")
    t.add_int_info("Example", 10)
    t.add_int_info("Example", 1)
    t.add_double_info("Other", 0.5)
    t.add_double_info("Other", 2.5)
    t.stop("A")
    j = t.to_json()
    infos = j["timings"][0]["infos"]
    assert [i["name"] for i in infos] == ["Example", "Other"]
    assert infos[0] == {"name": "Example", "num_calls": 2, "sum": 11, "min": 1, "max": 10}
    assert all(isinstance(infos[0][k], int) for k in ("sum", "min", "max"))
    assert infos[1] == {"name": "Other", "num_calls": 2, "sum": 3.0, "min": 0.5, "max": 2.5}


def test_to_json_sequential_and_nested():
    t = R.Timer()
    t.start("A")
    t.stop("A")
    t.start("B")
    t.stop("B")
    t.start("A")
    t.stop("A")
    j = t.to_json()
    assert [x["name"] for x in j["timings"]] == ["A", "B"]
    _check_timing(j["timings"][0], "A", 2)

    t = R.Timer()
    t.start("A")
    t.start("B")
    t.stop("B")
    t.stop("A")
    j = t.to_json()
    assert len(j["timings"]) == 1
    a = j["timings"][0]
    assert len(a["timings"]) == 1
    _check_timing(a["timings"][0], "B", 1)
    assert a["sum"] >= a["timings"][0]["sum"]


def test_to_json_domain_and_patch_timing():
    for patch in (False, True):
        t = R.Timer()
        t.add_domain(0, _domain())
        if patch:
            t.start_patch_timing(0, 0, "A")
            t.stop_patch_timing(0, 0, "A")
        else:
            t.start_domain_timing(0, "A")
            t.stop_domain_timing(0, "A")
        j = t.to_json()
        assert set(j) == {"comm_size", "domains", "timings"}
        assert len(j["domains"]) == 1 and len(j["domains"][0]) == 2 and j["domains"][0][0] == 0
        tj = j["timings"][0]
        _check_timing(tj, "A", 1)
        assert tj["domain_id"] == 0
        assert ("patch_id" in tj) == patch


def test_to_json_keys_sorted_and_compact():
    """The reference's JSON library writes object keys sorted and, compact, no whitespace."""
    t = R.Timer()
    t.add_domain(3, _domain())
    t.start_domain_timing(3, "Z")
    t.add_int_info("b", 1)
    t.stop_domain_timing(3, "Z")
    s = t.dumps()
    assert " " not in s and "\n" not in s
    j = json.loads(s)
    assert list(j) == sorted(j)
    assert list(j["timings"][0]) == sorted(j["timings"][0])
    assert list(j["timings"][0]["infos"][0]) == sorted(j["timings"][0]["infos"][0])


# ---- text report (Timer_MPI1.cpp:515-784) ----

def test_report_empty_and_single():
    assert _occ(R.Timer().report(), "No timings to report") == 1
    t = R.Timer()
    t.start("A")
    t.add_int_info("Example", 10)
    t.stop("A")
    s = t.report()
    assert s.startswith("\nTIMING RESULTS\n==============\n\n")
    assert (_occ(s, "A"), _occ(s, "time (sec)"), _occ(s, "average (sec)"), _occ(s, "min (sec)"),
            _occ(s, "max (sec)"), _occ(s, "average calls per rank")) == (1, 1, 0, 0, 0, 0)
    assert (_occ(s, "Example"), _occ(s, "Example avg"), _occ(s, "Example min"), _occ(s, "Example max")) == (1, 0, 0, 0)
    # the layout: the name, a dashed line, then "label: value" lines right-aligned to the
    # longest label ("average calls per rank"), then a blank line
    lines = s.split("\n")
    k = lines.index("A")
    assert lines[k + 1] == "-"
    assert lines[k + 2].startswith(" " * (23 - len("time (sec)")) + "time (sec): ")
    assert lines[k + 3] == " " * (23 - len("Example")) + "Example: 10"


def test_report_two_infos_and_sequential():
    t = R.Timer()
    t.start("A")
    t.add_int_info("Example", 10)
    t.add_int_info("Example", 11)
    t.stop("A")
    s = t.report()
    assert (_occ(s, "Example"), _occ(s, "Example avg"), _occ(s, "Example min"), _occ(s, "Example max")) == (3, 1, 1, 1)
    assert "Example avg: 10.5" in s
    t = R.Timer()
    for _ in range(2):
        t.start("A")
        t.stop("A")
    s = t.report()
    assert (_occ(s, "A"), _occ(s, "time (sec)"), _occ(s, "average (sec)"), _occ(s, "min (sec)"),
            _occ(s, "max (sec)"), _occ(s, "average calls per rank")) == (1, 0, 1, 1, 1, 1)
    assert "average calls per rank: 2\n" in s


def test_report_nested_and_patch():
    t = R.Timer()
    t.start("A")
    t.start("B")
    t.stop("B")
    t.stop("A")
    s = t.report()
    assert (_occ(s, "A"), _occ(s, "B"), _occ(s, "time (sec)")) == (2, 1, 2)
    assert "\nA -> B\n------\n" in s
    t = R.Timer()
    t.add_domain(0, _domain())
    for _ in range(2):
        t.start_patch_timing(0, 0, "A")
        t.stop_patch_timing(0, 0, "A")
    s = t.report()
    assert (_occ(s, "A"), _occ(s, "time (sec)"), _occ(s, "average (sec)")) == (1, 0, 1)


def test_report_merges_domains():
    t = R.Timer()
    t.add_domain(0, _domain())
    t.add_domain(1, _domain())
    for d in (0, 1):
        t.start_domain_timing(d, "A")
        t.add_int_info("Example", d)
        t.stop_domain_timing(d, "A")
    s = t.report()
    assert (_occ(s, "A"), _occ(s, "time (sec)"), _occ(s, "average (sec)"), _occ(s, "average calls per rank")) == (
        1, 0, 1, 1)
    assert (_occ(s, "Example"), _occ(s, "Example avg"), _occ(s, "Example min"), _occ(s, "Example max")) == (3, 1, 1, 1)
    for c_in_second, expect in ((None, (2, 1, 0, 0, 2)), (False, (3, 1, 1, 1, 2)), (True, (3, 1, 1, 1, 2))):
        t = R.Timer()
        _two_domains_nested(t, c_in_second)
        s = t.report()
        got = (_occ(s, "A"), _occ(s, "B"), _occ(s, "C"), _occ(s, "time (sec)"), _occ(s, "average (sec)"))
        assert got == expect, (c_in_second, s)
        assert _occ(s, "min (sec)") == 2 and _occ(s, "max (sec)") == 2 and _occ(s, "average calls per rank") == 2


# ---- saveToFile (Timer_MPI1.cpp:785-864) ----

def test_save_to_file(tmp_path):
    t = R.Timer()
    _two_domains_nested(t, True)
    path = str(tmp_path / "timer.json")
    with open(path, "w") as fh:
        fh.write("x" * 100000)  # overwritten, not appended to
    t.save_to_file(path)
    t.save_to_file(path)
    text = open(path).read()
    assert json.loads(text) == t.to_json()  # exactly one JSON value in the file
    assert text.startswith('{\n    "comm_size": 1,\n    "domains": [\n')
    with pytest.raises(R.TEError, match="Failed to open file"):
        t.save_to_file(str(tmp_path / "surely" / "not" / "there" / "timer.json"))


# ---- domains in the mesh schema (json.cpp:161-247) ----

def _strip_nulls(x):
    if isinstance(x, dict):
        return {k: _strip_nulls(v) for k, v in x.items() if v is not None}
    if isinstance(x, list):
        return [_strip_nulls(v) for v in x]
    return x


@pytest.mark.parametrize("name", ["2d_uniform_8x8_refined_cross_mpi1.json", "2d_refined_complicated_mpi2.json",
                                  "3d_mid_refine_4x4x4_mpi1.json"])
def test_domain_json_matches_mesh_schema(name):
    path = os.path.join(MESHES, name)
    dim = len(json.load(open(path))["levels"][0][0]["starts"])
    ns = (8, 8, 8)[:dim]
    dim, levels = mesh.read_mesh_json(path, ns, rank=None)
    t = R.Timer()
    for k, lv in enumerate(levels):
        t.add_domain(k, types.SimpleNamespace(dim=dim, ns=ns, patches=lv))
    j = t.to_json()
    assert [d[0] for d in j["domains"]] == list(range(len(levels)))
    ref = json.load(open(path))["levels"]
    for (k, pj), lv, lref in zip(j["domains"], levels, ref):
        assert len(pj) == len(lv)
        for p, rec, pr in zip(pj, lv, lref):
            # the reference writes orth_on_parent always (null for none) and orth_on_coarse
            # on every coarse neighbour (null across a corner)
            assert "orth_on_parent" in p
            for n in p["nbrs"] + p["corner_nbrs"] + p.get("edge_nbrs", []):
                assert ("orth_on_coarse" in n) == (n["type"] == "COARSE")
            assert _strip_nulls(p) == _strip_nulls(mesh.patch_to_json(rec, dim, ns))
            # and the reference's own file, field by field
            assert p["id"] == pr["id"] and p["parent_id"] == pr["parent_id"]
            assert np.allclose(p["starts"], pr["starts"]) and np.allclose(p["lengths"], pr["lengths"])
            mine = sorted((n["side"], n["type"], tuple(n["ids"]), n.get("orth_on_coarse")) for n in p["nbrs"])
            theirs = sorted((n["side"], n["type"], tuple(n["ids"]), n.get("orth_on_coarse")) for n in pr["nbrs"])
            assert mine == theirs
