"""Decks used by the tests: the reference's own decks (copied to
tests/golden/decks by tests/golden/make_fixtures.py) and the synthetic
workload decks of pennant_amd.decks."""
from pathlib import Path

from pennant_amd import decks as wd

HERE = Path(__file__).resolve().parent


def ref_deck(name: str) -> str:
    return (HERE / "golden" / "decks" / f"{name}.pnt").read_text()


def override(text: str, **kv) -> str:
    """The deck with the keys kv set (replaced where present, else added)."""
    lines = [ln for ln in text.splitlines() if ln.split() and ln.split()[0] not in kv]
    return "\n".join(lines + [f"{k} {v!r}" for k, v in kv.items()]) + "\n"


# Parameters no reference deck sets (gamma, qgamma, q2 and alfa away from
# their defaults, ssmin binding, a tight cflv, dtmax and dtfac)
PARAMS = dict(gamma=1.4, qgamma=1.4, q1=0.3, q2=1.5, alfa=0.3, ssmin=0.05, cflv=0.05, dtmax=5e-4,
              dtfac=1.05)


DECKS = {name: ref_deck(name) for name in
         ("sedovsmall", "sedov", "noh", "nohsmall", "leblanc")}
