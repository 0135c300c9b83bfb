"""The autotuner's table without a GPU: entries written by autotune.write_entry
are read back by read_tuning (and replace an entry with the same key), and
the C ABI exports the tuning entry point."""
import rmblast_amd
from rmblast_amd import autotune


def test_tuning_table_round_trip(tmp_path):
    path = str(tmp_path / "tuning.tsv")
    a = {"batches": 4, "prelim_ring": 16, "prelim_wg_per_cu": 8, "tb_ring": 64, "tb_wg_per_cu": 7, "tb_wave_halves": -1}
    b = dict(a, batches=12)
    autotune.write_entry(path, ("synth.matrix", 7, 30, 100), a, 812.5, "uneven 120 x 1500")
    autotune.write_entry(path, ("18p43g.matrix", 7, 30, 100), a, 700.0, "x")
    autotune.write_entry(path, ("synth.matrix", 7, 30, 100), b, 800.0, "uneven 120 x 1500")
    t = rmblast_amd.read_tuning(path)
    assert t == {("synth.matrix", 7, 30, 100): b, ("18p43g.matrix", 7, 30, 100): a}
    assert rmblast_amd.tuning_key("/a/b/synth.matrix", 7, 30, 100) in t


def test_committed_table_parses():
    for key, cfg in rmblast_amd.read_tuning().items():
        assert set(cfg) == set(rmblast_amd.TUNING_FIELDS)
        assert cfg["prelim_ring"] in (0, 16, 64) and cfg["tb_ring"] in (0, 64, 256)
