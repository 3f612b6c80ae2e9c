"""The autotuner's table on the host: loading, its line format and errors, set / get /
clear.  No GPU calls (the table is host state; tests/test_gpu_tune.py runs the variants)."""
import pytest

from thunderegg_amd import runtime as R


@pytest.fixture
def clean_table():
    R.tune_clear()
    yield
    R.tune_clear()


def test_variant_counts():
    assert R.tune_num_variants(R.TUNE_DFT_GENERIC) == 7
    assert R.tune_num_variants(R.TUNE_DFT_MFMA_2D) == 4
    assert R.tune_num_variants(R.TUNE_C3_SWEEP) == 10
    assert R.tune_num_variants(7) == 0


def test_load_set_get_clear(tmp_path, clean_table):
    p = tmp_path / "t.txt"
    p.write_text("# comment\n0 2 8 1 4   # trailing comment\n\n1 2 16 0 2\n")
    R.tune_load(p)
    assert R.tune_table_path() == str(p)
    assert R.tune_get(R.TUNE_DFT_GENERIC, 2, 8, R.F32) == 4
    assert R.tune_get(R.TUNE_DFT_MFMA_2D, 2, 16, R.F64) == 2
    assert R.tune_get(R.TUNE_DFT_MFMA_2D, 2, 32, R.F64) == 0  # no entry: the hand default
    R.tune_set(R.TUNE_DFT_GENERIC, 3, 8, R.F64, 6)
    assert R.tune_get(R.TUNE_DFT_GENERIC, 3, 8, R.F64) == 6
    R.tune_clear()
    assert R.tune_get(R.TUNE_DFT_GENERIC, 2, 8, R.F32) == 0


@pytest.mark.parametrize("text", ["0 2 8 1\n", "0 2 8 1 99\n", "5 2 8 1 0\n", "0 4 8 1 0\n", "0 2 8 7 0\n"])
def test_bad_lines_are_rejected(tmp_path, clean_table, text):
    p = tmp_path / "bad.txt"
    p.write_text(text)
    with pytest.raises(RuntimeError, match="line 1"):
        R.tune_load(p)


def test_bad_set_is_rejected(clean_table):
    with pytest.raises(RuntimeError):
        R.tune_set(R.TUNE_DFT_MFMA_2D, 2, 16, R.F32, 4)


def test_missing_file(tmp_path, clean_table):
    with pytest.raises(RuntimeError, match="cannot open"):
        R.tune_load(tmp_path / "none.txt")


def test_shipped_table_is_loaded_by_a_fresh_process():
    """tune_table.txt beside the library is what a new process uses"""
    import os
    import subprocess
    import sys

    path = os.path.join(os.path.dirname(R.__file__), "tune_table.txt")
    if not os.path.exists(path):
        pytest.skip("no table shipped beside the library")
    entries = []
    for line in open(path):
        line = line.split("#")[0].split()
        if line:
            entries.append([int(v) for v in line])
    code = ("from thunderegg_amd import runtime as R; import json, sys; "
            "e = json.loads(sys.argv[1]); "
            "print(R.tune_table_path()); "
            "print(all(R.tune_get(f, d, n, t) == v for f, d, n, t, v in e))")
    env = dict(os.environ)
    env.pop("TE_TUNE_TABLE", None)
    import json
    out = subprocess.run([sys.executable, "-c", code, json.dumps(entries)], capture_output=True, text=True,
                         env=env, cwd=os.path.dirname(os.path.dirname(R.__file__)), timeout=120)
    assert out.returncode == 0, out.stderr
    got_path, ok = out.stdout.split()
    assert os.path.realpath(got_path) == os.path.realpath(path) and ok == "True"
