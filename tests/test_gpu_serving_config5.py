"""The serving workload's shape on the GPU: Poisson-arriving trajectory
requests (mean 20 per second) whose systems are drawn from the uneven
benchmark distribution (chain counts log-uniform in [4, 192], chain lengths
in {8, 12, 16, 24, 32} residues, rho* in [0.005, 0.02], T* in [0.15, 0.50]),
each asking for a log-uniform number of collisions, submitted to the
continuous-batching server as they arrive, with device-time slices.  Every
request completes with the collisions it asked for, and a sample of them --
the largest and smallest system among them -- ends bit for bit where the same
request ends alone.  Collision counts are scaled down (1e3 to 3e4 instead of
1e5 to 1e8) so that the test runs in seconds."""
import math
import time

import numpy as np
import pytest

import pdmd

pytestmark = pytest.mark.gpu


def _draw(rng, n):
    out = []
    for k in range(n):
        nch = int(round(math.exp(rng.uniform(math.log(4), math.log(192)))))
        L = int(rng.choice([8, 12, 16, 24, 32]))
        dens = float(rng.uniform(0.005, 0.02))
        tstar = float(rng.uniform(0.15, 0.50))
        ncoll = int(math.exp(rng.uniform(math.log(1e3), math.log(3e4))))
        out.append(dict(system=pdmd.generate([L] * nch, dens, tstar, seed=500 + k), ncoll=ncoll, seed=900 + k,
                        nbeads=4 * nch * L))
    return out


def _alone(req):
    e = pdmd.Engine(waves=4)
    try:
        sid = e.add_system(req["system"], seed=req["seed"])
        e.advance([sid], [req["ncoll"]])
        return e.download(sid)["sv"]
    finally:
        e.close()


def test_poisson_requests_from_the_uneven_distribution():
    rng = np.random.default_rng(20261203)
    n = 14
    reqs = _draw(rng, n)
    arrive = np.cumsum(rng.exponential(1.0 / 20.0, n))
    print(f"serving: {n} requests of {min(r['nbeads'] for r in reqs)}..{max(r['nbeads'] for r in reqs)} beads",
          flush=True)
    e = pdmd.Engine(waves=4)
    srv = pdmd.Server(e, mem_budget_bytes=8 << 30, quantum=1 << 30, slice_ms=4.0)
    rid = [None] * n
    try:
        t0 = time.perf_counter()
        k = 0
        for _ in range(100000):
            now = time.perf_counter() - t0
            while k < n and arrive[k] <= now:
                r = reqs[k]
                rid[k] = srv.submit(r["system"], r["ncoll"], seed=r["seed"])
                k += 1
            srv.step()
            q, res, _ = srv.pending()
            if k == n and q == 0 and res == 0:
                break
        infos = [srv.result(x, with_state=True) for x in rid]
    finally:
        srv.close()
        e.close()
    for (info, _), r in zip(infos, reqs):
        assert info.state == 2 and info.status == 0 and info.done == r["ncoll"]
        assert info.t_done >= info.t_admit >= info.t_submit
    order = np.argsort([r["nbeads"] for r in reqs])
    for q in (order[0], order[-1], order[n // 2]):
        sv_alone = _alone(reqs[q])
        assert np.array_equal(infos[q][1].view(np.int64), sv_alone.view(np.int64)), f"request {q} differs"
