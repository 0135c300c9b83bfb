"""rmb_search_report (the outfmt 6 rows written while the search runs, a
batch of whole queries at a time) against rmb_search followed by
rmb_report_tabular over the same search: identical bytes and identical HSPs.
The search spans several batches (60 uneven records of 1-500 kbp, 5.5 Mbp,
against 300 families), with mask level on and off, E-value and bit-score
fields from override statistics, and the --evalue reap before mask level
(rmblastn/src/main.rs:1224-1234).  The rows themselves are pinned against the
oracle through the front end, which writes them the same way."""
import numpy as np
import pytest

import rmblast_amd
from rmblast_amd import synth
from conftest import RM_FIELDS

pytestmark = pytest.mark.gpu

FIELDS = RM_FIELDS.split(" ", 1)[1] + " evalue bitscore"


@pytest.fixture(scope="module")
def searcher(tmp_path_factory):
    w = synth.uneven_config(n_records=60, n_families=300)
    mpath = str(tmp_path_factory.mktemp("stream") / "synth.matrix")
    with open(mpath, "w") as f:
        f.write(synth.MATRIX_TEXT)
    s = rmblast_amd.Searcher(rmblast_amd.load_matrix(mpath), gap_open=20, gap_extend=5, word_size=7, xdrop_ungap=20,
                             xdrop_gap=30, xdrop_gap_final=100, min_raw_gapped_score=200, complexity_adjust=True,
                             matrix_name=mpath)
    s.load_library([f[1] for f in w.families])
    ka = rmblast_amd.ka_stats(mpath, 20, 5, matrix_lambda=0.11, matrix_k=0.05, matrix_alpha=0.9, matrix_beta=-20)
    yield w, s, ka
    s.close()


@pytest.mark.parametrize("mask,reap,stats", [(80, 0.0, False), (101, 0.0, True), (80, 1e-30, True)])
def test_streamed_report_is_the_report_after_search(searcher, mask, reap, stats):
    w, s, ka = searcher
    qs = [r[1] for r in w.records]
    qids = [r[0] for r in w.records]
    sids = [f[0] for f in w.families]
    kw = dict(ka=ka if stats else None, evalue_max=reap)
    streamed = s.search_report(qs, FIELDS, qids, sids, mask, **kw)
    st = s.stats()
    h1, o1 = s.results()
    h2, o2 = s.search(qs)
    after = s.report_tabular(FIELDS, qids, sids, mask, **kw)
    assert st["phase1_launches"] >= 2  # several batches, each handed to the report as it completed
    assert streamed.count(b"\n") > 1000
    assert streamed == after
    assert np.array_equal(h1, h2) and np.array_equal(o1, o2)


def test_streamed_report_refuses_unknown_fields(searcher):
    w, s, _ = searcher
    with pytest.raises(rmblast_amd.RmbError, match="unknown outfmt 6 field"):
        s.search_report([w.records[0][1]], "score nosuchfield", [w.records[0][0]], [f[0] for f in w.families])
