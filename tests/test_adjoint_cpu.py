"""The adjoint's reverse passes, run on the host (swe_adj_host_* of the C ABI),
checked against central finite differences of the oracle's forward routines:
hllc (swe.mjs:974-1011), residual (:308-629), applyBC (:289-303), dryClean and
applyFriction (:659-745).  Each check is a dot-product identity
<lambda, (F(x + e d) - F(x - e d)) / 2e> = <F'^T lambda, d> for random lambda and
directions d that cross no branch of the forward, so a dropped term, a wrong
sign or a transposed index in a reverse pass fails it.  BC is linear, so its
identity holds to rounding."""
import math

import numpy as np
import pytest

import oracle
import swe_amd

G = 9.80665


def vjp_fd(f, x, lam, d, eps):
    return float(np.dot(lam, (f(x + eps * d) - f(x - eps * d)) / (2 * eps)))


# ---- hllc^T ----------------------------------------------------------------------
HLLC_CASES = [
    ("subcritical rarefaction", (2.0, 0.5, 0.3, 1.5, -0.2, 0.1)),
    ("two-shock (hStar above both)", (2.0, 3.0, 0.4, 1.8, -2.5, -0.3)),
    ("shock left only", (1.0, 1.5, 0.2, 3.0, 0.0, 0.5)),
    ("supercritical right-going", (1.0, 8.0, 0.5, 0.8, 6.0, 0.2)),
    ("supercritical left-going", (0.7, -5.0, 0.1, 1.2, -9.0, 0.4)),
    ("dry left", (5e-5, 0.0, 0.0, 1.3, -0.4, 0.2)),
    ("dry right", (0.9, 0.6, -0.3, 2e-5, 0.0, 0.0)),
    ("contact moving left", (1.1, -0.8, 0.9, 1.0, -0.7, -0.6)),
]


@pytest.mark.parametrize("name,x", HLLC_CASES, ids=[c[0] for c in HLLC_CASES])
def test_hllc_rev_matches_fd(name, x):
    x = np.array(x, dtype=np.float64)
    dry = 1e-4
    rng = np.random.default_rng(7)
    lam = rng.normal(size=3)
    f = lambda v: oracle.hllc(*v, dry=dry, g=G)
    g = swe_amd.hllc_rev(*x, dry, G, *lam)
    wet = np.array([x[0] > dry] * 3 + [x[3] > dry] * 3)
    for i in range(6):
        d = np.zeros(6)
        d[i] = 1.0
        if not wet[i] and i % 3 != 0:
            continue        # momentum of a dry side is not read
        if not wet[i]:
            continue        # depth of a dry side: a kink of the dry threshold, not probed
        eps = 1e-6 * max(1.0, abs(x[i]))
        fd = vjp_fd(f, x, lam, d, eps)
        assert abs(g[i] - fd) <= 1e-6 * max(1.0, abs(fd)), (name, i, g[i], fd)


# ---- residual^T --------------------------------------------------------------------
def basin(nx, ny, seed, *, dx=5.0, wetdry=False):
    rng = np.random.default_rng(seed)
    W, H = nx + 4, ny + 4
    x = (np.arange(W) - 1.5) * dx
    y = (np.arange(H) - 1.5) * dx
    X, Y = np.meshgrid(x, y)
    L = nx * dx
    b = -3.0 + (4.2 * X / L if wetdry else 0.0) + 0.3 * np.sin(Y / (3 * dx)) + 0.2 * np.cos(X / (2.3 * dx))
    eta = 0.3 * np.exp(-(((X - 0.4 * L) ** 2 + (Y - 0.5 * ny * dx) ** 2) / (0.05 * L * L)))
    # surface roughness well above the finite-difference step, so that no limiter
    # changes branch between the two perturbed evaluations
    eta = eta + rng.normal(0, 0.03, X.shape)
    h = np.maximum(0.0, eta - b)
    hu = h * (0.3 * np.sin(Y / (2.1 * dx)) + rng.normal(0, 0.1, X.shape))
    hv = h * (0.2 * np.cos(X / (1.7 * dx)) + rng.normal(0, 0.1, X.shape))
    hu[h <= 1e-3] = 0
    hv[h <= 1e-3] = 0
    return b.ravel(), h.ravel(), hu.ravel(), hv.ravel()


def rows_for(kind, nx, ny, dx, coriolis):
    if kind == "sphere":
        return swe_amd.geometry_spherical(nx, ny, lat0=-60.0, lat1=60.0, omega=7.2921e-5)
    return swe_amd.geometry_cartesian(nx, ny, dx, coriolis=coriolis)


RES_CASES = [
    dict(id="mc-coriolis", order=2, limiter="mc", coriolis=1e-3),
    dict(id="minmod", order=2, limiter="minmod"),
    dict(id="central", order=2, limiter="central"),
    dict(id="order1", order=1, limiter="mc"),
    dict(id="mc-wetdry", order=2, limiter="mc", wetdry=True),
    dict(id="positivity", order=2, limiter="mc", wetdry=True, dt_scale=40.0),
    dict(id="sphere-mc", order=2, limiter="mc", kind="sphere"),
    dict(id="sphere-order1", order=1, limiter="minmod", kind="sphere"),
]


@pytest.mark.parametrize("case", RES_CASES, ids=[c["id"] for c in RES_CASES])
def test_residual_rev_dot_product(case):
    nx, ny, dx = 22, 17, 5.0
    kind = case.get("kind", "cart")
    coriolis = case.get("coriolis", 0.0)
    rows = rows_for(kind, nx, ny, dx, coriolis)
    kw = dict(order=case["order"], limiter=case["limiter"], coriolis=coriolis, manning=0.0)
    if kind == "sphere":
        kw.update(sphere=True, R=swe_amd.R_EARTH)
    o = oracle.OracleSim(nx, ny, rows, dx=dx, dy=dx, **kw)
    desc, keep = swe_amd.make_desc(nx, ny, rows, **kw)
    b, h, hu, hv = basin(nx, ny, 11, dx=dx, wetdry=case.get("wetdry", False))
    if kind == "sphere":
        h = h + 3000.0          # an ocean on the sphere; same relief
        hu, hv = hu * 50, hv * 50
    W = nx + 4
    # the ghost ring keeps the basin's own values: residual() reads it as stored,
    # and a mirrored ring would put every limiter at the wall exactly on its kink
    o.load_state(b, h, hu, hv)
    dt = o.max_dt() * case.get("dt_scale", 1.0)
    rng = np.random.default_rng(3)
    interior = np.zeros((ny + 4, W), bool)
    interior[2:-2, 2:-2] = True
    interior = interior.ravel()
    lam = [rng.normal(size=h.size) * interior for _ in range(3)]
    # directions: zero where the depth sits near a threshold of the forward
    safe = h > 0.05
    d = [rng.normal(size=h.size) * safe * (1.0 if q == 0 else np.abs(hu) + np.abs(hv) + 1e-3) for q in range(3)]

    def R(state):
        o.load_state(b, *state)
        o.residual(dt)
        return np.concatenate([o.field(oracle.RH).copy(), o.field(oracle.RHU).copy(), o.field(oracle.RHV).copy()])

    x = np.concatenate([h, hu, hv])
    dd = np.concatenate(d)
    N = h.size

    def f(v):
        return R((v[:N], v[N:2 * N], v[2 * N:]))

    eps = 1e-7
    fd = vjp_fd(f, x, np.concatenate(lam), dd, eps)
    lh, lhu, lhv = swe_amd.residual_rev_host(desc, h, hu, hv, b, dt, *lam)
    ad = float(np.dot(np.concatenate([lh, lhu, lhv]), dd))
    assert abs(ad - fd) <= 1e-6 * max(abs(fd), 1e-9 * np.abs(dd).sum()), (case["id"], ad, fd)
    if case["id"] == "positivity":
        o.load_state(b, h, hu, hv)
        o.residual(dt)
        assert (o.field(oracle.THETA) < 1).any(), "the positivity limiter should be active here"


# ---- BC^T ------------------------------------------------------------------------
BC_SETS = [
    ("reflect",) * 4,
    ("outflow",) * 4,
    ("periodic",) * 4,
    ("periodic", "periodic", "reflect", "reflect"),
    ("reflect", "outflow", "periodic", "periodic"),
    ("outflow", "reflect", "reflect", "outflow"),
]


@pytest.mark.parametrize("kinds", BC_SETS, ids=["-".join(k) for k in BC_SETS])
@pytest.mark.parametrize("shape", [(9, 7), (2, 1), (3, 5)])
def test_bc_rev_is_the_transpose(kinds, shape):
    nx, ny = shape
    rows = swe_amd.geometry_cartesian(nx, ny, 2.0)
    o = oracle.OracleSim(nx, ny, rows, dx=2.0, dy=2.0, manning=0.0)
    desc, keep = swe_amd.make_desc(nx, ny, rows, manning=0.0)
    rng = np.random.default_rng(sum(shape))
    N = (nx + 4) * (ny + 4)
    S = [rng.normal(size=N) for _ in range(4)]
    l = [rng.normal(size=N) for _ in range(3)]
    o.load_state(*S)
    o.set_bc(kinds)
    o.apply_bc()
    _, h, hu, hv = o.state()
    lhs = sum(float(np.dot(a, c)) for a, c in zip(l, (h, hu, hv)))
    lt = swe_amd.bc_rev_host(desc, kinds, *l)
    rhs = sum(float(np.dot(a, c)) for a, c in zip(lt, S[1:]))
    assert abs(lhs - rhs) <= 1e-12 * sum(float(np.abs(a).sum()) for a in l) * 3
    W = nx + 4
    for a in lt:                  # the transpose leaves nothing on the ghost ring
        g = a.reshape(ny + 4, W).copy()
        g[2:-2, 2:-2] = 0
        assert not g.any()


FLATHER_SETS = [
    ("flather", "reflect", "periodic", "periodic"),
    ("flather", "flather", "flather", "flather"),
    ("reflect", "flather", "outflow", "flather"),
]


@pytest.mark.parametrize("kinds", FLATHER_SETS, ids=["-".join(k) for k in FLATHER_SETS])
@pytest.mark.parametrize("shape", [(9, 7), (3, 5)])
def test_flather_bc_rev_matches_fd(kinds, shape):
    """Flather's ghost fill (swe.mjs:1084-1116) is nonlinear in its source cell:
    its transpose against central differences of the oracle's applyBC, with a
    dry source cell on every side (its ghost depends on the far field only)."""
    nx, ny = shape
    rows = swe_amd.geometry_cartesian(nx, ny, 2.0)
    o = oracle.OracleSim(nx, ny, rows, dx=2.0, dy=2.0, manning=0.0)
    desc, keep = swe_amd.make_desc(nx, ny, rows, manning=0.0)
    rng = np.random.default_rng(3 + sum(shape))
    W, H = nx + 4, ny + 4
    N = W * H
    b = -1.5 + 0.2 * rng.normal(size=N)
    h = np.maximum(0.0, 0.1 * rng.normal(size=N) - b)
    hu = 0.3 * rng.normal(size=N) * h
    hv = 0.3 * rng.normal(size=N) * h
    # a dry corner of the interior, next to the west and south sides
    dry = np.zeros((H, W), bool)
    dry[2, 2] = True
    h[dry.ravel()] = 0.0
    hu[dry.ravel()] = hv[dry.ravel()] = 0.0
    ext = np.array([0.12, 0.03, -0.05, 0.02, 0.07, -0.04, 0.02, 0.05])
    o.set_bc(kinds)

    def F(v):
        o.load_state(b, v[:N], v[N:2 * N], v[2 * N:])
        o.apply_bc(ext)
        return np.concatenate([o.h.copy(), o.hu.copy(), o.hv.copy()])

    inter = np.zeros((H, W), bool)
    inter[2:-2, 2:-2] = True
    inter = inter.ravel()
    lam = [rng.normal(size=N) for _ in range(3)]
    d = np.concatenate([rng.normal(size=N) * inter * sc for sc in (0.05, 0.02, 0.02)])
    x = np.concatenate([h, hu, hv])
    fd = vjp_fd(F, x, np.concatenate(lam), d, 1e-7)
    lt = swe_amd.bc_rev_host(desc, kinds, *lam, ext=ext, state=(b, h, hu, hv))
    ad = float(np.dot(np.concatenate(lt), d))
    assert abs(ad - fd) <= 1e-7 * sum(float(np.abs(a).sum()) for a in lam), (ad, fd)
    for a in lt:                  # the transpose leaves nothing on the ghost ring
        g = a.reshape(H, W).copy()
        g[2:-2, 2:-2] = 0
        assert not g.any()


GATHER_SETS = [
    ("reflect",) * 4,
    ("outflow",) * 4,
    ("periodic", "periodic", "reflect", "reflect"),
    ("reflect", "outflow", "outflow", "reflect"),
    ("periodic", "periodic", "outflow", "reflect"),
]


@pytest.mark.parametrize("kinds", GATHER_SETS, ids=["-".join(k) for k in GATHER_SETS])
@pytest.mark.parametrize("shape", [(9, 7), (4, 4), (17, 5)])
def test_bc_rev_gather_form_equals_the_loop(kinds, shape):
    """The device's gather form of BC^T (one work item per source column or row,
    no atomics) and the same transpose folded on read by each interior cell give
    the ghost-by-ghost loop's result to the bit; both are refused for a periodic
    south or north side and for Flather sides (those run the ghost-by-ghost
    form)."""
    nx, ny = shape
    rows = swe_amd.geometry_cartesian(nx, ny, 2.0)
    desc, keep = swe_amd.make_desc(nx, ny, rows, manning=0.0)
    rng = np.random.default_rng(7 + 3 * nx + ny)
    W, H = nx + 4, ny + 4
    N = W * H
    b = -1.5 + 0.2 * rng.normal(size=N)
    h = np.maximum(0.0, 0.1 * rng.normal(size=N) - b)
    h[rng.random(N) < 0.15] = 0.0                       # some dry source cells
    hu = 0.3 * rng.normal(size=N) * h
    hv = 0.3 * rng.normal(size=N) * h
    ext = np.array([0.12, 0.03, -0.05, 0.02, 0.07, -0.04, 0.02, 0.05])
    lam = [rng.normal(size=N) for _ in range(3)]
    loop = swe_amd.bc_rev_host(desc, kinds, *lam, ext=ext, state=(b, h, hu, hv))
    gath = swe_amd.bc_rev_host(desc, kinds, *lam, ext=ext, state=(b, h, hu, hv), gather=True)
    for x, y in zip(loop, gath):
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64))
    # folded on read, as the device's reverse epilogues apply it
    onread = swe_amd.bc_rev_host(desc, kinds, *lam, onread=True)
    for x, y in zip(loop, onread):
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64))
    for refused in (("reflect", "reflect", "periodic", "periodic"), ("flather", "reflect", "reflect", "reflect")):
        for form in ("gather", "onread"):
            with pytest.raises(swe_amd.SweError):
                swe_amd.bc_rev_host(desc, refused, *lam, ext=ext, state=(b, h, hu, hv), **{form: True})


# ---- dryClean^T / friction^T / sponge^T -----------------------------------------------
@pytest.mark.parametrize("stage,sponge", [(1, False), (2, False), (2, True)])
@pytest.mark.parametrize("field", [False, True])
def test_epilogue_rev_dot_product(stage, field, sponge):
    nx, ny, dx = 15, 12, 4.0
    rows = swe_amd.geometry_cartesian(nx, ny, dx)
    man = 0.03
    o = oracle.OracleSim(nx, ny, rows, dx=dx, dy=dx, manning=man, use_manning_field=field)
    desc, keep = swe_amd.make_desc(nx, ny, rows, manning=man)
    b, h, hu, hv = basin(nx, ny, 5, dx=dx, wetdry=True)
    rng = np.random.default_rng(9)
    h = h + rng.normal(0, 0.002, h.size) * (h < 0.01)         # some negative depths to floor
    N = h.size
    nf = rng.uniform(0.0, 0.05, N) if field else None
    if field:
        nf[::5] = 0.0
        o.set_manning_field(nf)
    dt = 0.7
    interior = np.zeros((ny + 4, nx + 4), bool)
    interior[2:-2, 2:-2] = True
    interior = interior.ravel()
    lam = [rng.normal(size=N) * interior for _ in range(3)]      # ghosts of the output are not read
    # makeSponge strengths (swe.mjs:1174-1186): a band of the domain, one cell relaxed fully (f = 1)
    strength = None
    if sponge:
        st2 = np.zeros((ny + 4, nx + 4))
        st2[2:-2, -8:-2] = np.linspace(0.002, 0.02, 6)
        st2[5, -3] = 1.0
        strength = st2.ravel()
        o.set_sponge(strength)

    def out(state, n=nf):
        o.load_state(b, *state)
        if field:
            o.field(oracle.MANNING)[:] = n
        o.dry_clean()
        if stage == 2:
            o.apply_friction(dt)
            if sponge:
                o.sponge(dt, 0.05)
        return np.concatenate([o.h.copy(), o.hu.copy(), o.hv.copy()])

    safe = (np.abs(h) > 0.01) & (h > 2e-3)
    d = [rng.normal(size=N) * safe * s for s in (0.01, 0.05, 0.05)]
    x = np.concatenate([h, hu, hv])
    dd = np.concatenate(d)
    f = lambda v: out((v[:N], v[N:2 * N], v[2 * N:]))
    eps = 1e-6
    fd = vjp_fd(f, x, np.concatenate(lam), dd, eps)
    lh, lhu, lhv, lN = swe_amd.epilogue_rev_host(desc, stage, (h, hu, hv), nf, dt, lam, sponge=strength)
    ad = float(np.dot(np.concatenate([lh, lhu, lhv]), dd))
    assert abs(ad - fd) <= 1e-6 * max(abs(fd), 1e-12), (ad, fd)
    # floored and dry cells pass nothing back
    assert not lh[h < 0].any()
    assert not lh[~interior].any()
    if stage == 2 and field:
        dn = rng.uniform(0, 1e-3, N) * (nf > 0)
        fdn = vjp_fd(lambda n: out((h, hu, hv), n), nf, np.concatenate(lam), dn, 1e-4)
        # out() returns all three fields; lambda spans them, dn perturbs n only
        assert abs(float(np.dot(lN, dn)) - fdn) <= 1e-6 * max(abs(fdn), 1e-12)
    elif stage == 2:
        # a scalar n is read at every cell: its per-cell adjoints are those of a uniform field
        ref = swe_amd.epilogue_rev_host(desc, stage, (h, hu, hv), np.full(N, man), dt, lam, sponge=strength)
        assert np.array_equal(lN, ref[3]) and lN.any()
    else:
        assert not lN.any()


# ---- one whole step^T, composed from the host passes --------------------------------
def step_rev_composed(o, desc, bc, b, A, dt, nf, lamC, ext=None, sponge=None):
    """The reverse of one step(dt) (swe.mjs:754-814) in the order the device
    driver runs it: forward recompute with the oracle, then epilogue^T,
    residual^T and BC^T of stage 2 and of stage 1.  ext: the step's Flather
    values; sponge: the strengths of a sponge applied after the step."""
    N = b.size
    W = desc.nx + 4
    inter = np.zeros((desc.ny + 4, W), bool)
    inter[2:-2, 2:-2] = True
    inter = inter.ravel()

    def bc_of(S, bed):
        o.load_state(bed, *S)
        o.apply_bc(ext)
        st = o.state()
        return st[0], np.array(st[1:])

    def res(bed, S):
        o.load_state(bed, *S)
        o.residual(dt)
        return np.array([o.field(oracle.RH).copy(), o.field(oracle.RHU).copy(), o.field(oracle.RHV).copy()])

    bg, Ap = bc_of(A, b)
    R1 = res(bg, Ap)
    pre1 = Ap + dt * R1
    B = pre1.copy()
    hB = B[0]
    hB[inter & (hB < 0)] = 0.0
    dry = inter & (hB <= desc.min_depth)
    B[1][dry] = 0.0
    B[2][dry] = 0.0
    bg, Bp = bc_of(B, bg)
    R2 = res(bg, Bp)
    pre2 = 0.5 * (Ap + Bp + dt * R2)
    l2 = swe_amd.epilogue_rev_host(desc, 2, pre2, nf, dt, lamC, sponge=sponge)
    lpre2, lN = np.array(l2[:3]), l2[3]
    lAp, lBp, lR2 = 0.5 * lpre2, 0.5 * lpre2, 0.5 * dt * lpre2
    lBp = lBp + np.array(swe_amd.residual_rev_host(desc, *Bp, bg, dt, *lR2))
    lBp = np.array(swe_amd.bc_rev_host(desc, bc, *lBp, ext=ext, state=(bg, *Bp)))
    lpre1 = np.array(swe_amd.epilogue_rev_host(desc, 1, pre1, nf, dt, lBp)[:3])
    lAp = lAp + lpre1
    lAp = lAp + np.array(swe_amd.residual_rev_host(desc, *Ap, bg, dt, *(dt * lpre1)))
    return np.array(swe_amd.bc_rev_host(desc, bc, *lAp, ext=ext, state=(bg, *Ap))), lN


STEP_CASES = [
    dict(id="mc-reflect", limiter="mc", bc=("reflect",) * 4),
    dict(id="minmod-mixed", limiter="minmod", bc=("outflow", "reflect", "periodic", "periodic")),
    dict(id="order1-periodic", order=1, limiter="mc", bc=("periodic",) * 4),
    dict(id="mc-wetdry", limiter="mc", bc=("reflect",) * 4, wetdry=True),
    dict(id="sphere-central", limiter="central", bc=("periodic", "periodic", "reflect", "reflect"), sphere=True),
    dict(id="sphere-mc", limiter="mc", bc=("periodic", "periodic", "reflect", "reflect"), sphere=True),
    dict(id="mc-flather", limiter="mc", bc=("flather", "reflect", "periodic", "periodic"), flather=True),
    dict(id="mc-flather-sponge", limiter="mc", bc=("flather", "flather", "reflect", "reflect"), flather=True,
         sponge=True),
]


@pytest.mark.parametrize("case", STEP_CASES, ids=[c["id"] for c in STEP_CASES])
def test_step_rev_composition(case):
    nx, ny, dx = 18, 14, 5.0
    order = case.get("order", 2)
    kw = dict(order=order, limiter=case["limiter"], manning=0.0)
    if case.get("sphere"):
        rows = swe_amd.geometry_spherical(nx, ny, R=2.0e6, lat0=-50.0, lat1=50.0, omega=7.2921e-5)
        kw.update(sphere=True, R=2.0e6)
    else:
        rows = swe_amd.geometry_cartesian(nx, ny, dx)
    o = oracle.OracleSim(nx, ny, rows, dx=dx, dy=dx, use_manning_field=True, **kw)
    desc, keep = swe_amd.make_desc(nx, ny, rows, **kw)
    b, h, hu, hv = basin(nx, ny, 21, dx=dx, wetdry=case.get("wetdry", False))
    N = h.size
    rng = np.random.default_rng(4)
    nf = rng.uniform(0.01, 0.04, N)
    o.set_manning_field(nf)
    o.set_bc(case["bc"])
    o.load_state(b, h, hu, hv)
    dt = o.max_dt()
    inter = np.zeros((ny + 4, nx + 4), bool)
    inter[2:-2, 2:-2] = True
    inter = inter.ravel()
    lamC = np.array([rng.normal(size=N) * inter for _ in range(3)])

    ext = np.array([0.08, 0.02, -0.05, 0.01, 0, 0, 0, 0]) if case.get("flather") else None
    strength = None
    if case.get("sponge"):
        st2 = np.zeros((ny + 4, nx + 4))
        st2[2:-2, -6:-2] = np.linspace(0.002, 0.01, 4)
        strength = st2.ravel()
        o.set_sponge(strength)

    def step(S, n=nf):
        o.field(oracle.MANNING)[:] = n
        o.load_state(b, *S)
        o.step(dt, ext)
        if strength is not None:
            o.sponge(dt, 0.0)
        return np.concatenate([o.h.copy(), o.hu.copy(), o.hv.copy()])

    A = np.array([h, hu, hv])
    lA, lN = step_rev_composed(o, desc, case["bc"], b, A, dt, nf, lamC, ext=ext, sponge=strength)
    safe = inter & (h > 0.05)
    d = np.array([rng.normal(size=N) * safe, rng.normal(size=N) * safe * 0.1, rng.normal(size=N) * safe * 0.1])
    f = lambda v: step(v.reshape(3, N))
    fd = vjp_fd(f, A.ravel(), lamC.ravel(), d.ravel(), 1e-7)
    ad = float(np.dot(lA.ravel(), d.ravel()))
    assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)
    dn = rng.uniform(-1, 1, N) * inter * 1e-3
    fdn = vjp_fd(lambda n: step(A, n), nf, lamC.ravel(), dn, 1e-4)
    assert abs(float(np.dot(lN, dn)) - fdn) <= 1e-6 * abs(fdn), (float(np.dot(lN, dn)), fdn)
