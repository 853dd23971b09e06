"""Checkpoint directory of a run on the GPU: file names, the shared file's
appended blocks with the short header, the progress record, and resuming
from a recorded index bit for bit (modules-io.h:306-327, io/inc/io.h:185-188,
io/inc/readdefines.h:340-420)."""
import numpy as np
import pytest

from symphas_amd import abi, inputs, models
from symphas_amd.solver import FDPlan, MultiPlan, checkpoint_progress

pytestmark = pytest.mark.gpu


def _desc():
    return models.model_b((40, 24), h=0.8, dt=0.005, dtype=abi.F32, path=abi.PATH_FUSED)


def _state_after(desc, psi, steps):
    with FDPlan(desc) as p:
        p.upload(psi)
        p.step(steps)
        return p.download()


def test_single_file_layout_and_resume(gpu_lib, orc, tmp_path):
    print("[case] checkpoint dir, single file", flush=True)
    d = _desc()
    psi = inputs.uniform((24, 40), seed=91, lo=-0.5, hi=0.5).astype(np.float32)
    with FDPlan(d) as p:
        p.upload(psi)
        states = {}
        for idx in (10, 20, 30):
            p.step(10)
            states[idx] = p.download()
            p.checkpoint_save(tmp_path, idx, done=(idx == 30))
    # bytes: the first block under the full header, later ones under "<index> @"
    single = models.make_desc((40, 24), models.Stencil2d2h(5, 6, 13), 0.8, 0.005, [])
    texts = {i: orc.checkpoint_text(single, orc.pad(single, states[i].astype(np.float64)), i)
             for i in states}
    want = texts[10] + b"".join(b"%d @\n" % i + texts[i].split(b"\n", 1)[1] for i in (20, 30))
    assert (tmp_path / "checkpoint" / "data0").read_bytes() == want
    assert (tmp_path / "checkpoint" / "progress.txt").read_text() == "#$10\n#$20\n#$30\n#$DONE\n"
    assert checkpoint_progress(tmp_path) == (30, 3, True)
    # resume: each recorded index, and an index between two (the last block
    # whose index does not pass it)
    for ask, got in ((10, 10), (20, 20), (25, 20), (30, 30), (99, 30), (5, 10)):
        with FDPlan(d) as q:
            assert q.checkpoint_load(tmp_path, ask) == got
            assert np.array_equal(q.download(), states[got])
    # continuing from index 20 equals the uninterrupted run
    with FDPlan(d) as q:
        q.checkpoint_load(tmp_path, 20)
        q.step(17)
        assert np.array_equal(q.download(), _state_after(d, psi, 37))


def test_file_per_index_and_coupled_fields(gpu_lib, orc, tmp_path):
    print("[case] checkpoint dir, file per index, coupled", flush=True)
    d = _desc()
    psi = inputs.uniform((24, 40), seed=92, lo=-0.5, hi=0.5).astype(np.float32)
    with FDPlan(d) as p:
        p.upload(psi)
        p.step(5)
        s5 = p.download()
        p.checkpoint_save(tmp_path, 5, field_id=3, single_file=False)
        text5 = p.checkpoint_text(5)
        p.step(5)
        p.checkpoint_save(tmp_path, 10, field_id=3, single_file=False)
    assert (tmp_path / "checkpoint" / "data3_5").read_bytes() == text5
    assert (tmp_path / "checkpoint" / "data3_10").exists()
    with FDPlan(d) as q:
        assert q.checkpoint_load(tmp_path, 5, field_id=3, single_file=False) == 5
        assert np.array_equal(q.download(), s5)
        with pytest.raises(abi.SpsError):
            q.checkpoint_load(tmp_path, 7, field_id=3, single_file=False)  # no such file
    md = models.model_c((36, 20), h=1.0, dt=0.002, dtype=abi.F32, path=abi.PATH_FUSED)
    fields = [inputs.uniform((20, 36), seed=93 + f, lo=-0.4, hi=0.4).astype(np.float32)
              for f in range(2)]
    cdir = tmp_path / "coupled"
    with MultiPlan(md) as p:
        for f, a in enumerate(fields):
            p.upload(f, a)
        p.step(8)
        p.checkpoint_save(cdir, 8)
        p.step(8)
        want = [p.download(f) for f in range(2)]
    assert sorted(x.name for x in (cdir / "checkpoint").iterdir()) == \
        ["data0", "data1", "progress.txt"]
    with MultiPlan(md) as q:
        assert q.checkpoint_load(cdir, 8) == 8
        q.step(8)
        assert all(np.array_equal(q.download(f), want[f]) for f in range(2))
