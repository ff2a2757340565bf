"""The persistent lane solver's operand-cache schedules (mechgen/lane_cache.py):
executed on the CPU as the kernel executes them (versioned symbolic memory,
reads of a batch before its fills), every cache hit returns the current
value of the address it stands for; the schedules save a sizeable part of
the slot-block reads; a planted wrong slot is caught; the generated header
holds the schedules the generator builds now."""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "geos-chem_amd" / "mechgen"))


def raw(m):
    return {"nvar": m.nvar, "nfix": m.nfix, "nspec": m.nspec, "nreact": m.nreact,
            "lu_nonzero": m.lu_nonzero, "reactions": m.reactions, "lu_irow": list(m.lu_irow),
            "lu_icol": list(m.lu_icol), "lu_crow": list(m.lu_crow), "lu_diag": list(m.lu_diag)}


def setup(m):
    from emit_tables import build_tables
    import lane_cache as LC
    r = raw(m)
    return LC, r, build_tables(r)


def test_batches_are_the_kernel_sizes():
    import lane_cache as LC
    assert LC.batches(0) == []
    assert LC.batches(13) == [8, 4, 1]
    assert LC.batches(16) == [16]
    assert LC.batches(47) == [16, 16, 8, 4, 2, 1]
    for n in range(200):
        assert sum(LC.batches(n)) == n


@pytest.mark.parametrize("name", ["fullchem", "Hg"])
@pytest.mark.parametrize("slots", [16, 32])
def test_schedules_hit_only_current_values(name, slots, mech_of):
    LC, r, t = setup(mech_of(name))
    lw, sw = LC.build(r, t, slots)
    for ev, words, nread_min in ((LC.lu_events(r, t), lw, 0.2), (LC.solve_events(r), sw, 0.2)):
        reads, hits, writes, errors = LC.emulate(ev, words, slots)
        assert errors == []
        assert hits / (reads + hits) > nread_min
        assert writes == sum(1 for x in ev if x[0] == "W")


def test_event_streams_follow_the_tables(fullchem):
    """Every LU position appears once as a result, with all its Crout and
    Ghimj operands; every row of both substitutions once."""
    LC, r, t = setup(fullchem)
    ev = LC.lu_events(r, t)
    res = [x[1] for x in ev if x[0] == "W"]
    assert res == list(range(fullchem.lu_nonzero))
    n_cr = sum(len(x[1]) for x in ev if x[0] == "C")
    n_jv = sum(len(x[1]) for x in ev if x[0] == "G")
    assert n_cr == 2 * len(t["cr_l"]) and n_jv == len(t["jv_b"])
    sv = LC.solve_events(r)
    rows = [x[1][1] for x in sv if x[0] == "W"]
    fwd = [i for i in range(fullchem.nvar) if fullchem.lu_crow[i] != fullchem.lu_diag[i]]
    assert rows == fwd + list(range(fullchem.nvar - 1, -1, -1))


def test_planted_wrong_slot_is_caught(fullchem):
    LC, r, t = setup(fullchem)
    ev = LC.solve_events(r)
    words = list(LC.build(r, t, 16)[1])
    # move the first hit of the stream to another slot
    for k, w in enumerate(words):
        for sh in (0, 8, 16, 24):
            b = (w >> sh) & 255
            if b >> 5 == LC.HIT:
                words[k] = w ^ (1 << sh)
                break
        else:
            continue
        break
    assert LC.emulate(ev, words, 16)[3]


def test_emitted_schedules_match_generator(fullchem):
    LC, r, t = setup(fullchem)
    hdr = (ROOT / "geos-chem_amd" / "csrc" / "gen" / "mech_fullchem.h").read_text()
    for slots in LC.SLOTS:
        lw, sw = LC.build(r, t, slots)
        for nm, w in (("lu", lw), ("sv", sw)):
            head = "static __constant__ unsigned lc%d_%s[%d] = {\n" % (slots, nm, len(w) + 16)
            assert head in hdr
            i = hdr.index(head) + len(head)
            first = [int(x) for x in hdr[i:i + 400].replace("\n", " ").split(",")[:16]]
            assert first == w[:16]
