"""The autotuner's table file (so3x_tune_save / so3x_tune_load) without a
GPU: entries load, save back unchanged, and invalid entries are refused."""
import pytest

import so3_amd.api as A


@pytest.fixture(autouse=True)
def _fresh_table():
    A.tune_clear()
    yield
    A.tune_clear()


def test_table_round_trip(tmp_path):
    src = tmp_path / "in.txt"
    # seven columns (no gamma engine: the default rule) and eight
    src.write_text("# comment\n64 32 0 0 1 0 0\n300 12 1 1 0 1 1536\n96 40 0 1 0 0 0 2\n")
    assert A.tune_load(src) == 3
    out = tmp_path / "out.txt"
    A.tune_save(out)
    lines = [l for l in out.read_text().splitlines() if not l.startswith("#")]
    assert "64 32 0 0 1 0 0 0" in lines and "300 12 1 1 0 1 1536 0" in lines and "96 40 0 1 0 0 0 2" in lines


@pytest.mark.parametrize("bad", ["64 32 0 0 2 0 0", "64 32 0 0 0 0 100", "300 4 0 0 0 0 8192",
                                 "300 4 0 0 0 0 1200", "64 32 0 0 0 0 0 3", "64 32 0 0 0 0 0 -1"])
def test_invalid_entries_are_refused(tmp_path, bad):
    f = tmp_path / "bad.txt"
    f.write_text(bad + "\n")
    with pytest.raises(RuntimeError):
        A.tune_load(f)
