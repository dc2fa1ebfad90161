"""The tuner's choice rule and the shipped tuning table (CPU only)."""
import json
import os

from periodfind_amd import tune

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_pick_takes_the_fastest_when_no_claim_ties():
    t = {(0, 4): 1.00, (0, 16): 1.05, (0, 64): 1.10, (2, 4): 0.90}
    assert tune.pick(t) == ((2, 4), 0.90)


def test_pick_breaks_a_near_tie_toward_the_smaller_claim():
    # CE 4096-16384 points in the shipped report: claim 64 leads claim 16 by
    # 0.3%, claim 4 is 6% behind.
    t = {(0, 4): 0.00745, (0, 16): 0.00706, (0, 64): 0.00704, (2, 4): 0.00983}
    (g, c), s = tune.pick(t)
    assert (g, c) == (0, 16) and s == 0.00706


def test_pick_keeps_the_fastest_group():
    # A smaller claim of another group inside the tie band does not move the
    # group: plain launches take the group alone.
    t = {(1, 64): 1.000, (1, 16): 1.020, (2, 4): 1.005}
    assert tune.pick(t)[0] == (1, 64)


def test_shipped_table_is_the_rule_applied_to_the_shipped_report():
    report = json.load(open(os.path.join(ROOT, "profiles", "tuning_report_gfx950.json")))
    want = tune.table_from_report(report)
    got = []
    for line in open(os.path.join(ROOT, "periodfind_amd", "tuning_gfx950.txt")):
        if line.strip() and not line.startswith("#"):
            a, lo, g, c = line.split()
            got.append((a, int(lo), int(g), int(c)))
    assert sorted(got) == sorted(want)
    # The long-curve CE claim is no longer the L2-thrashing 64.
    ce = {lo: c for a, lo, g, c in got if a == "ce"}
    assert ce[4096] == 16


def test_from_report_cli_writes_the_shipped_entries(tmp_path):
    out = tmp_path / "tuning.txt"
    rc = tune.main(["--from-report",
                    os.path.join(ROOT, "profiles", "tuning_report_gfx950.json"),
                    "--out", str(out)])
    assert rc == 0
    body = [l for l in out.read_text().splitlines() if l and not l.startswith("#")]
    shipped = [l.strip() for l in open(os.path.join(ROOT, "periodfind_amd", "tuning_gfx950.txt"))
               if l.strip() and not l.startswith("#")]
    assert body == shipped
