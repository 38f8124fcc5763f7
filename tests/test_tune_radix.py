"""Host logic of the offline radix tuner (dns_amd/tune_radix.py): the table
it reads back and the candidate schedules it would build (no GPU)."""
import math

from dns_amd import build, roofline, tune_radix


def test_table_covers_every_grid_size():
    t = tune_radix.table()
    assert sorted(t) == sorted(build.grid_sizes())
    for n, (e, rad) in t.items():
        assert math.prod(rad) == n, (n, rad)
        assert all(e % r == 0 for r in rad), (n, e, rad)
        assert n % e == 0 and roofline.elements_per_thread(n) == e


def test_candidates_are_the_fewest_stage_orderings():
    t = tune_radix.table()
    for n in (32, 48, 96, 128, 192, 256, 512, 768, 1024):
        e, cands = tune_radix.candidates(n)
        assert e == t[n][0]
        assert cands[0] == t[n][1]                      # the table's own schedule first
        assert len(set(cands)) == len(cands)
        stages = {len(c) for c in cands}
        assert stages == {len(t[n][1])}, (n, stages)   # one stage count: the fewest
        for c in cands:
            assert math.prod(c) == n and all(e % r == 0 for r in c)
    # counts that follow from the radix multisets (n = 2^a 3^b)
    assert len(tune_radix.candidates(512)[1]) == 1      # 8-8-8 only
    assert len(tune_radix.candidates(1024)[1]) == 10    # orderings of 8,8,4,4 and 8,8,8,2
    assert len(tune_radix.candidates(96)[1]) == 12      # orderings of 4,4,2,3


def test_candidate_library_names_are_distinct():
    e, cands = tune_radix.candidates(1024)
    names = {tune_radix.name(1024, c) for c in cands}
    assert len(names) == len(cands)
