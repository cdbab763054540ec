"""The divergences of the batched bump form are the scheme's: a member that
diverged on the device (tests/batched64_members.json, recorded on an MI355X
by the batched form) diverges on the CPU oracle at the same residual check.
Member 48 (16x9 quads, P4) is the one small enough for the CPU suite; the
others are run by tests/batched_divergence.py."""
import json
from pathlib import Path

from batched_divergence import member, run_member

REC = json.loads((Path(__file__).parent / "batched64_members.json").read_text())["members"]


def test_record_matches_the_draws():
    for m in REC[:8] + [REC[48]]:
        _, _, (nx, ny, P, M) = member(m["i"])
        assert (nx, ny, P) == (m["nx"], m["ny"], m["P"]) and abs(M - m["M"]) < 1e-15


def test_member_48_diverges_on_the_oracle_where_the_device_did():
    dev = REC[48]
    assert dev["diverged"]
    out = run_member(48, dev["steps"] + 300)
    assert out["status"] == "diverged" and out["steps"] == dev["steps"], (out["status"], out["steps"], dev["steps"])
