"""GJK / EPA narrow phase for ellipsoids and cylinders (any pair with one of
them that is not plane-based).

Pinned by the support-function characterisation of convex shapes: with
f(n) = h_A(n) + h_B(-n) (h the support functions), the signed distance is
-min over unit n of f(n) (the separating-axis duality; for overlapping shapes
its magnitude is the minimum translation distance). The test minimises f
over a dense direction sampling refined by local optimisation, with exact
support functions written here, and compares distance, normal and contact
points. Plane pairs use the deepest support point, pinned in closed form.

Parity with MuJoCo's own convex collision (its exact contact points and its
multi-contact cases) is unpinned: MuJoCo is not available here.
"""
from __future__ import annotations

import numpy as np
import pytest
from scipy.optimize import minimize

import emu
import ref64

SCENE = """<mujoco><option gravity="0 0 0"/><worldbody>
<geom name="floor" type="plane" size="0 0 1" margin="0.1"/>
<body><freejoint/><geom type="sphere" size="0.07" margin="0.1"/></body>
<body><freejoint/><geom type="capsule" size="0.05 0.12" margin="0.1"/></body>
<body><freejoint/><geom type="ellipsoid" size="0.12 0.07 0.05" margin="0.1"/></body>
<body><freejoint/><geom type="cylinder" size="0.06 0.09" margin="0.1"/></body>
<body><freejoint/><geom type="box" size="0.08 0.05 0.1" margin="0.1"/></body>
<body><freejoint/><geom type="ellipsoid" size="0.05 0.09 0.08" margin="0.1"/></body>
<body><freejoint/><geom type="cylinder" size="0.04 0.12" margin="0.1"/></body>
</worldbody></mujoco>"""

TYPES = {0: "plane", 1: "sphere", 2: "capsule", 3: "ellipsoid", 4: "cylinder", 5: "box", 6: "ellipsoid", 7: "cylinder"}


def support(kind, size, x, R, d):
    dl = R.T @ d
    n = np.linalg.norm(dl)
    dl = dl / n
    if kind == "sphere":
        p = size[0] * dl
    elif kind == "capsule":
        p = size[0] * dl + np.array([0, 0, size[1] if dl[2] >= 0 else -size[1]])
    elif kind == "ellipsoid":
        q = size ** 2 * dl
        p = q / np.sqrt(np.sum(size ** 2 * dl ** 2))
    elif kind == "cylinder":
        r = np.hypot(dl[0], dl[1])
        p = np.array([size[0] * dl[0] / r if r > 0 else 0, size[0] * dl[1] / r if r > 0 else 0,
                      size[1] if dl[2] >= 0 else -size[1]])
    else:
        p = np.where(dl >= 0, size, -size)
    return x + R @ p


def fdir(u):
    th, ph = u
    return np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])


def signed_distance(ka, sa, xa, Ra, kb, sb, xb, Rb):
    f = lambda n: n @ support(ka, sa, xa, Ra, n) - n @ support(kb, sb, xb, Rb, -n)
    k = np.arange(20000) + 0.5
    th = np.arccos(1 - 2 * k / 20000)
    ph = np.pi * (1 + 5 ** 0.5) * k
    vals = np.array([f(fdir((a, b))) for a, b in zip(th[::4], ph[::4])])
    best = np.argsort(vals)[:6]
    out = []
    for i in best:
        r = minimize(lambda u: f(fdir(u)), [th[::4][i], ph[::4][i]], method="Nelder-Mead",
                     options=dict(xatol=1e-10, fatol=1e-13, maxiter=4000))
        out.append((r.fun, fdir(r.x)))
    v, n = min(out, key=lambda t: t[0])
    return -v, n


def rand_rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


@pytest.fixture(scope="module")
def orc():
    e = emu.CompiledModel(SCENE)
    o = ref64.Oracle(ref64.fields_from(e.field))
    o._e = e
    return o


def _collide(orc, pi, xa, Ra, xb, Rb):
    out = (ref64.OrcContact * 4)()
    arr = lambda a: ref64._dp(np.ascontiguousarray(np.asarray(a, np.float64).ravel()))
    n = ref64.lib().orc_collide_pair(orc.p, pi, arr(xa), arr(Ra), arr(xb), arr(Rb), out, 4)
    return [dict(dist=out[k].dist, pos=np.array(out[k].pos), n=np.array(out[k].frame[:3])) for k in range(n)]


PAIRS = [(1, 3), (2, 3), (3, 6), (3, 4), (4, 7), (4, 5), (3, 5), (2, 4), (1, 7)]


@pytest.mark.parametrize("ga,gb", PAIRS, ids=[f"{TYPES[a]}-{TYPES[b]}" for a, b in PAIRS])
def test_gjk_epa_against_support_duality(orc, ga, gb):
    pairs = orc.pairs()
    pi = [i for i, p in enumerate(pairs) if (p[0], p[1]) == (ga, gb)][0]
    size = orc._e.field("geom_size").reshape(-1, 3)
    ka, kb = TYPES[ga], TYPES[gb]
    rng = np.random.default_rng(ga * 10 + gb)
    nsep = npen = 0
    for _ in range(12):
        Ra, Rb = rand_rot(rng), rand_rot(rng)
        xa = np.zeros(3)
        xb = rng.normal(size=3)
        xb *= rng.uniform(0.05, 0.3) / np.linalg.norm(xb)
        d_ref, n_ref = signed_distance(ka, size[ga], xa, Ra, kb, size[gb], xb, Rb)
        c = _collide(orc, pi, xa, Ra, xb, Rb)
        if d_ref > 0.1 + 1e-6:
            assert c == []
            continue
        if d_ref > 0.1 - 1e-6:
            continue
        assert len(c) == 1
        c = c[0]
        tol = 2e-6 if d_ref > 0 else 5e-5  # EPA approximates curved surfaces by a polytope
        assert c["dist"] == pytest.approx(d_ref, abs=tol), (d_ref, c["dist"])
        # the returned normal attains the optimum of the duality (to 5e-5 m: a
        # support with a flat or straight part, like a capsule's side, turns a
        # 1e-5 rad normal error into a few 1e-6 m)
        n = c["n"]
        f = n @ support(ka, size[ga], xa, Ra, n) - n @ support(kb, size[gb], xb, Rb, -n)
        assert -f == pytest.approx(d_ref, abs=5e-5)
        # the contact position sits midway between support points of A and B
        pa, pb = c["pos"] - 0.5 * c["dist"] * n, c["pos"] + 0.5 * c["dist"] * n
        assert n @ pa == pytest.approx(n @ support(ka, size[ga], xa, Ra, n), abs=5e-5)
        assert n @ pb == pytest.approx(n @ support(kb, size[gb], xb, Rb, -n), abs=5e-5)
        nsep += d_ref > 0
        npen += d_ref < 0
    assert nsep + npen >= 6


@pytest.mark.parametrize("ga,gb", PAIRS, ids=[f"{TYPES[a]}-{TYPES[b]}" for a, b in PAIRS])
def test_kernel_narrow_phase_matches_oracle(orc, ga, gb):
    """The kernel's FP32 GJK/EPA (its source run on the CPU with fused
    multiply-adds, as on the GPU) against the FP64 oracle on 300 random poses:
    separated, touching and overlapping up to 4 cm (deeper overlaps than that
    do not arise between resting geoms of these sizes)."""
    pairs = orc.pairs()
    pi = [i for i, p in enumerate(pairs) if (p[0], p[1]) == (ga, gb)][0]
    size = orc._e.field("geom_size").reshape(-1, 3)
    gt = orc._e.field("geom_type").astype(int)
    rng = np.random.default_rng(100 + ga * 10 + gb)
    nd, na = [], []
    for _ in range(300):
        Ra, Rb = rand_rot(rng), rand_rot(rng)
        xa = rng.normal(size=3) * 0.1
        xb = rng.normal(size=3)
        xb = xa + xb * rng.uniform(0.02, 0.3) / np.linalg.norm(xb)
        Ra, Rb, xa, xb = (np.float32(a).astype(np.float64) for a in (Ra, Rb, xa, xb))
        c = _collide(orc, pi, xa, Ra, xb, Rb)
        k = emu.narrow(gt[ga], gt[gb], size[ga], size[gb], xa, Ra, xb, Rb, 0.1)
        if len(c) != len(k):
            d = (c or [dict(dist=k[0][0])])[0]["dist"]
            assert abs(d - 0.1) < 1e-5, (c, k)  # only at the margin's edge
            continue
        if c and c[0]["dist"] >= -0.04:
            nd.append(abs(c[0]["dist"] - k[0][0]))
            na.append(np.degrees(np.arccos(np.clip(np.dot(c[0]["n"], k[0][2]), -1, 1))))
    nd, na = np.array(nd), np.array(na)
    assert len(nd) > 150
    # EPA stops at a support gap of 1e-6 of the shapes' extent in FP32: depths
    # agree to a few 1e-6 m (1e-5 at worst), normals to 0.5 degree (1.5 worst)
    assert nd.max() < 3e-5 and np.quantile(nd, 0.99) < 5e-6, (nd.max(), np.quantile(nd, 0.99))
    assert na.max() < 1.5 and np.quantile(na, 0.99) < 0.5, (na.max(), np.quantile(na, 0.99))


def test_plane_ellipsoid_and_cylinder_closed_form(orc):
    pairs = orc.pairs()
    size = orc._e.field("geom_size").reshape(-1, 3)
    I = np.eye(3)
    for g, kind in ((3, "ellipsoid"), (4, "cylinder")):
        pi = [i for i, p in enumerate(pairs) if (p[0], p[1]) == (0, g)][0]
        rng = np.random.default_rng(g)
        for _ in range(10):
            R = rand_rot(rng)
            x = np.array([0.1, -0.2, rng.uniform(0.0, 0.15)])
            low = support(kind, size[g], x, R, np.array([0, 0, -1.0]))
            c = _collide(orc, pi, np.zeros(3), I, x, R)
            if low[2] > 0.1:
                assert c == []
                continue
            assert len(c) == 1
            assert c[0]["dist"] == pytest.approx(low[2], abs=1e-12)
            assert np.allclose(c[0]["n"], [0, 0, 1])
            assert np.allclose(c[0]["pos"][:2], low[:2])


DROP = """<mujoco><option solver="Newton" iterations="50"/><worldbody>
<geom type="plane" size="0 0 1"/>
<body pos="0 0 0.3"><freejoint/><geom type="ellipsoid" size="0.1 0.06 0.05"/></body>
<body pos="0.05 0.02 0.5" euler="30 20 0"><freejoint/><geom type="cylinder" size="0.05 0.08"/></body>
<body pos="-0.04 0 0.7"><freejoint/><geom type="box" size="0.05 0.05 0.05"/></body>
<body pos="0.02 -0.03 0.9"><freejoint/><geom type="sphere" size="0.05"/></body>
<body pos="0 0.04 1.1" euler="0 80 0"><freejoint/><geom type="capsule" size="0.03 0.08"/></body>
</worldbody></mujoco>"""


def drop_states(model_field, B, seed):
    orc = ref64.Oracle(ref64.fields_from(model_field))
    rng = np.random.default_rng(seed)
    nv = int(model_field("nv")[0])
    q = np.tile(model_field("qpos0"), (B, 1))
    qv = np.zeros((B, nv))
    qa = np.zeros((B, nv))
    lens = rng.integers(40, 160, size=B)
    for k in range(int(lens.max())):
        idx = np.nonzero(lens > k)[0]
        qs, vs, as_ = q[idx].copy(), qv[idx].copy(), qa[idx].copy()
        orc.batch_step(qs, vs, as_, np.zeros((len(idx), 0)), 2)
        q[idx], qv[idx], qa[idx] = qs, vs, as_
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return f32(q), f32(qv), f32(qa), orc


def test_interpreter_matches_oracle_on_convex_drop():
    import states
    B = 16
    sim = emu.EmuSim(DROP, B, maxcon=64, maxefc=256)
    q, qv, qa, orc = drop_states(sim.field, B, 3)
    sim.set_state(q, qv, qa)
    sim.step(None, nsub=1, dump=True)
    st = sim.stats()
    assert not (st["status"] & ~3).any(), st["status"]
    ncon = 0
    worst = 0.0
    for e in range(B):
        states.oracle_step(orc, q[e], qv[e], qa[e], np.zeros(0, np.float32))
        con = sim.dump(1, e, "contact").reshape(-1, 16)[: st["ncon"][e]]
        ok, msg = states.compare_contacts(con[:, 13].view(np.int32), con[:, 12], orc.contacts(), margin_tol=1e-4, ncon=st["ncon"][e])
        assert ok, f"env {e}: {msg}"
        # FP32 EPA against FP64 EPA: same pair list, depths to 1e-5 m, normals to 0.2 degree
        oc = sorted((c["pair"], c["dist"], tuple(c["frame"][:3])) for c in orc.contacts())
        kc = sorted((int(c[13:14].view(np.int32)[0]), float(c[12]), tuple(c[3:6])) for c in con)
        for (p1, d1, n1), (p2, d2, n2) in zip(oc, kc):
            assert p1 == p2 and abs(d1 - d2) < 1e-5, (p1, d1, d2)
            assert np.degrees(np.arccos(np.clip(np.dot(n1, n2), -1, 1))) < 0.2
        ncon += len(con)
        b = orc.get("qacc")
        worst = max(worst, np.max(np.abs(sim.dump(2, e, "qacc") - b)) / max(1.0, np.max(np.abs(b))))
    print(f"contacts {ncon}, worst qacc rel err {worst:.2e}")
    assert ncon > B
    # EPA normals of curved pairs agree to ~0.1 degree between FP32 and FP64;
    # stiff contacts turn that into up to ~5e-3 of qacc
    assert worst < 1e-2
