"""The serving loop on the device: a request ends in the same state (to the
FP64 atomics-order bar 1e-12 per cell) and after the same number of steps
whether it is served alone or in a mix that arrives over time."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class TickClock:
    """Simulated time: every reading advances by 1 ms, so arrivals are
    admitted at loop iterations that do not depend on the machine's speed."""

    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 1e-3
        return self.t

    def sleep(self, dt):
        self.t += max(0.0, dt)


def requests(pkg):
    rng = np.random.default_rng(21)
    out = []
    for rid, (edge, arrival) in enumerate(((11, 0.0), (13, 0.004), (9, 0.02), (15, 0.05))):
        cfg, f = pkg.mixture_instance(rng, edge, samples_per_node=10, seed=900 + rid)
        out.append(pkg.Request(rid, arrival, cfg, f, "homogeneous", 1e-4, 60))
    cfg, f = pkg.shock_instance(rng, 7, 9, 1.8, 400, 950, 0)
    out.append(pkg.Request(4, 0.01, cfg, f, "shock", 1e-9, 40))
    return out


def serve(pkg, reqs):
    clock = TickClock()
    species = reqs[0].cfg
    with pkg.Server(species, node_capacity=100_000, cell_capacity=64, max_slots=8, check_every=10,
                    keep_f=True, clock=clock, sleep=clock.sleep) as srv:
        return srv.run(reqs)


def test_same_alone_and_in_the_mix(pkg):
    reqs = requests(pkg)
    print("[serve] 4 homogeneous + 1 shock arriving over time", flush=True)
    mix = {o.rid: o for o in serve(pkg, reqs).outcomes}
    assert all(o.reason in ("tol", "plateau", "max_steps") for o in mix.values())
    for r in reqs:
        print(f"[serve] request {r.rid} alone", flush=True)
        alone = serve(pkg, [r]).outcomes[0]
        assert alone.steps == mix[r.rid].steps and alone.reason == mix[r.rid].reason
        a, b = alone.f, mix[r.rid].f
        err = np.max(np.abs(a - b), axis=1) / np.max(np.abs(a), axis=1)
        assert np.all(err <= 1e-12), (r.rid, err.max())


def test_requests_larger_than_the_pool_are_refused(pkg):
    reqs = requests(pkg)
    clock = TickClock()
    print("[serve] pool of 2000 nodes", flush=True)
    with pkg.Server(reqs[0].cfg, node_capacity=2000, cell_capacity=8, max_slots=2, clock=clock,
                    sleep=clock.sleep) as srv:
        st = srv.run(reqs)
    o = {x.rid: x for x in st.outcomes}
    assert o[0].reason != "too_large" and o[2].reason != "too_large"  # 1331 and 729 nodes
    assert o[1].reason == "too_large" and o[3].reason == "too_large" and o[4].reason == "too_large"
    assert st.max_resident <= 2
