"""Decks used by the tests: the reference's own decks (copied to
tests/golden/decks by tests/golden/make_fixtures.py) and the synthetic
workload decks of pennant_amd.decks."""
from pathlib import Path

from pennant_amd import decks as wd

HERE = Path(__file__).resolve().parent


def ref_deck(name: str) -> str:
    return (HERE / "golden" / "decks" / f"{name}.pnt").read_text()


DECKS = {name: ref_deck(name) for name in
         ("sedovsmall", "sedov", "noh", "nohsmall", "leblanc")}
