"""The progress record of a run's checkpoint directory (host code, no GPU):
"#$<index>" lines and "#$DONE", read as the reference reads them when it
loads a backup configuration (get_completed_indices, conf/src/conf.cpp)."""
from symphas_amd.solver import checkpoint_progress


def test_progress_record(tmp_path):
    assert checkpoint_progress(tmp_path) == (-1, 0, False)  # nothing recorded
    d = tmp_path / "checkpoint"
    d.mkdir()
    (d / "progress.txt").write_text("#$100\n  #$200  \nnot an index\n#$x1\n#$300\n")
    assert checkpoint_progress(tmp_path) == (300, 3, False)
    with open(d / "progress.txt", "a") as f:
        f.write("#$400\n#$DONE\n")
    assert checkpoint_progress(tmp_path) == (400, 4, True)
