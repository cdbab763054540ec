"""Host logic of the serving loop (no GPU): how far a tick may advance a
request."""
from flurrypp_amd.serve import Request, Server


def test_room_stops_at_checks_and_at_the_last_step():
    srv = Server(steps_per_tick=7, check_every=10)
    un = Request(0, "vortex3d", (4,), 2, 0.0, n_steps=13)
    assert srv._room(un) == 13
    un.steps = 10
    assert srv._room(un) == 3
    st = Request(1, "bump2d", (8, 4), 2, 0.0, steady_tol=1e-6, max_steps=25)
    assert srv._room(st) == 10          # next check at step 10
    st.steps = 14
    assert srv._room(st) == 6           # next check at step 20
    st.steps = 20
    assert srv._room(st) == 5           # max_steps comes before the check at 30
