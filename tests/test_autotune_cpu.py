"""The chemistry autotuner's protocol (no GPU): warm-up call, one call per
candidate scored by time per accepted step, then the best configuration,
cached per (device, network, cell-count bucket)."""

from prism3d_amd.autotune import ChemTuner


def test_tuner_picks_lowest_time_per_step(tmp_path, monkeypatch):
    monkeypatch.setenv("P3D_TUNE_CACHE", str(tmp_path / "tune.json"))
    cands = ((4, 8, 0.95), (4, 8, 0.99), (2, 0, 0.0))
    t = ChemTuner(262144, "MI355X", candidates=cands)
    assert t.next_config() == cands[0]          # warm-up call
    t.record(cands[0], 7.0, 1000)
    # per-step cost: 2.0e-3, 1.0e-3, 4.0e-3 -> the second wins
    for cfg, sec, steps in ((cands[0], 4.0, 2000), (cands[1], 3.0, 3000), (cands[2], 8.0, 2000)):
        assert t.tuning and t.next_config() == cfg
        t.record(cfg, sec, steps)
    assert not t.tuning and t.next_config() == cands[1]
    # a new tuner for the same device / size bucket starts from the cache
    t2 = ChemTuner(300000, "MI355X", candidates=cands)
    assert not t2.tuning and t2.next_config() == cands[1]
    # another size bucket tunes again
    t3 = ChemTuner(4096, "MI355X", candidates=cands)
    assert t3.tuning
