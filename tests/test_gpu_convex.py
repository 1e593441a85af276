"""Ellipsoids and cylinders on the GPU (the step kernel built with the GJK /
EPA narrow phase) against the FP64 oracle, through the C ABI."""
from __future__ import annotations

import numpy as np
import pytest

import states
from mjmlx_hip import capi
from test_convex_collision import DROP, drop_states

pytestmark = pytest.mark.gpu


def test_convex_drop_single_step_parity():
    model = capi.Model(DROP)
    B = 32
    q, qv, qa, orc = drop_states(model.field, B, 7)
    sim = capi.BatchedSim(model, B)
    sim.set_state(q, qv)
    sim.set_qacc_warmstart(qa)
    dump = sim.debug_step(None)
    st = sim.stats()
    assert not (st["status"] & ~3).any(), st["status"]
    ncon = 0
    worst = 0.0
    for e in range(B):
        states.oracle_step(orc, q[e], qv[e], qa[e], np.zeros(0, np.float32))
        con = dump.get(1, e, "contact").reshape(-1, 16)[: st["ncon"][e]]
        ok, msg = states.compare_contacts(con[:, 13].view(np.int32), con[:, 12], orc.contacts(), margin_tol=1e-4)
        assert ok, f"env {e}: {msg}"
        oc = sorted((c["pair"], c["dist"]) for c in orc.contacts())
        kc = sorted((int(c[13:14].view(np.int32)[0]), float(c[12])) for c in con)
        for (p1, d1), (p2, d2) in zip(oc, kc):
            assert p1 == p2 and abs(d1 - d2) < 1e-5, (p1, d1, d2)
        ncon += len(con)
        b = orc.get("qacc")
        worst = max(worst, np.max(np.abs(dump.get(2, e, "qacc") - b)) / max(1.0, np.max(np.abs(b))))
    print(f"convex drop: contacts {ncon}, worst qacc rel err {worst:.2e}")
    assert ncon > B
    assert worst < 1e-2
