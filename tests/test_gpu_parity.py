"""Parity of the HIP engine (through the C ABI) with the FP64 oracle and the
reference's stats fixtures.

Tolerances: the FP64 GPU path is held to the reference's own device-vs-CPU
bars (crates/stl-gpu/tests/stats.rs:248-275), relative to the largest entry:
statistics 1e-13 per family (max |dS| / max |S|), variances 1e-12, gradient
1e-12 (max |dg| / max |g|); the reference's fixture tolerances
(stats.rs:183-201) for its fixtures.  The mixed path (FP32 bulk, FP64 sums) is
held to 1e-5 per family and 1e-4 on the gradient at these small sizes; its
error at the benchmark's size is checked against oracle vectors elsewhere.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from fixtures import cases

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FAMILY_NAMES = ["mean", "var", "S1", "S2", "S3", "S4"]
STAT_BAR, VAR_BAR, GRAD_BAR = 1e-13, 1e-12, 1e-12  # crates/stl-gpu/tests/stats.rs:248-275


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


_ENGINES = {}


def engine(nside, J, batch=1, precision="fp64"):
    from stl_port import NativeEngine
    key = (nside, J, batch, precision)
    if key not in _ENGINES:
        print(f"[gpu] NativeEngine nside={nside} J={J} batch={batch} precision={precision}", flush=True)
        _ENGINES[key] = NativeEngine(nside, J, batch=batch, backend="gpu", precision=precision)
    return _ENGINES[key]


def family_errors(values, ref, plan):
    """max |dS| / max |S| per family; the mean is scaled by the map RMS sqrt(var)
    (a mean-subtracted map has a mean of rounding size)."""
    fam = np.array([FAMILY_NAMES.index(c["family"]) for c in plan])
    out = {}
    for f, name in enumerate(FAMILY_NAMES):
        sel = fam == f
        if sel.any():
            scale = np.abs(ref[..., sel]).max()
            if name == "mean":
                scale = np.sqrt(np.abs(ref[..., fam == 1]).max())
            out[name] = float(np.abs(values[..., sel] - ref[..., sel]).max() / max(scale, 1e-300))
    return out


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("nside", [2, 4, 8, 16, 64])
def test_operators_match_oracle(nside):
    J = int(np.log2(nside)) + 1
    eng = engine(nside, J)
    ref = O.Engine(nside, J)
    rng = np.random.default_rng(nside)
    for n in range(J):
        N = 12 * (nside >> n) ** 2
        print(f"[gpu] operators nside={nside} level={n}", flush=True)
        x = rng.standard_normal((2, N))
        y = torch.empty((2, 4, N, 2), dtype=torch.float64, device="cuda")
        eng.apply_operator(0, n, torch.from_numpy(x).cuda(), y)
        yc = y.cpu().numpy()
        yc = yc[..., 0] + 1j * yc[..., 1]
        assert rel(yc, ref.conv(n, x)) < 1e-12
        gy = rng.standard_normal((2, 4, N)) + 1j * rng.standard_normal((2, 4, N))
        gyt = torch.from_numpy(np.stack([gy.real, gy.imag], -1).copy()).cuda()
        gx = torch.empty((2, N), dtype=torch.float64, device="cuda")
        eng.apply_operator(1, n, gyt, gx)
        assert rel(gx.cpu().numpy(), ref.conv_adjoint(n, gy)) < 1e-12
        if n + 1 < J:
            ns = nside >> n
            indptr, cols, vals = O.decimation_rows(ns, 0, 3 * ns * ns)
            D = sp.csr_matrix((vals, cols, indptr), shape=(3 * ns * ns, N))
            yd = torch.empty((2, N // 4), dtype=torch.float64, device="cuda")
            eng.apply_operator(2, n, torch.from_numpy(x).cuda(), yd)
            assert rel(yd.cpu().numpy(), (D @ x.T).T) < 1e-13
            g = rng.standard_normal((2, N // 4))
            gd = torch.empty((2, N), dtype=torch.float64, device="cuda")
            eng.apply_operator(3, n, torch.from_numpy(g).cuda(), gd)
            assert rel(gd.cpu().numpy(), (D.T @ g.T).T) < 1e-13


@pytest.mark.parametrize("case", [c for c in cases("stats", "stats")
                                  if c["L"] == 4 and c["kernel_sz"] == 5 and c["channels"] == 1
                                  and not c["fewer"] and c["j_to_dg"] == list(range(c["J"]))],
                         ids=lambda c: c.name)
def test_stats_fixtures(case):
    """The reference's fixtures of the hot-path configuration, through the C ABI."""
    eng = engine(case["nside"], case["J"], case["batch"])
    x = case.arr("map0").reshape(case["batch"], -1)
    print(f"[gpu] fixture {case.name}", flush=True)
    values, _ = eng.forward(x)
    ref = case.complex("values").reshape(values.shape)
    flat = lambda z: np.concatenate([z.real.ravel(), z.imag.ravel()])
    assert np.max(np.abs(flat(values) - flat(ref)) / (1e-11 + 1e-10 * np.abs(flat(ref)))) <= 1.0
    grad = eng.vjp(case.complex("gradient").reshape(values.shape))
    refg = case.arr("band_gradient0").reshape(grad.shape)
    assert np.max(np.abs(grad - refg) / (1e-10 + 1e-9 * np.abs(refg))) <= 1.0


@pytest.mark.parametrize("nside, J", [(32, 4), (64, 5)])
def test_forward_vjp_vs_oracle(nside, J):
    """Random N(0,1) maps (seed 0), cotangent 2 (S(x) - S(y)) (the plain-method workload)."""
    eng = engine(nside, J)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(12 * nside * nside)
    y = np.random.default_rng(1).standard_normal(12 * nside * nside)
    ref = O.Engine(nside, J)
    ref.load(y)
    sy = ref.forward()
    ref.load(x)
    sx = ref.forward()
    cot = 2.0 * (sx - sy)
    g_ref = ref.gradient(cot)
    print(f"[gpu] forward+vjp nside={nside} J={J}", flush=True)
    values, _ = eng.forward(x, generation=5)
    errs = family_errors(values, sx, eng.plan)
    print(f"[gpu] fp64 vs oracle per family {errs}", flush=True)
    assert max(errs.values()) < STAT_BAR, errs
    grad = eng.vjp(cot, generation=5)
    gerr = rel(grad, g_ref)
    print(f"[gpu] fp64 vs oracle gradient {gerr:.3e}", flush=True)
    assert gerr < GRAD_BAR


def test_mixed_precision_vs_oracle():
    nside, J = 64, 5
    eng = engine(nside, J, precision="mixed")
    g = np.random.default_rng(7).standard_normal(12 * 16 * 16)
    x = np.repeat(g, 16) + 0.1 * np.random.default_rng(8).standard_normal(12 * nside * nside)
    x = np.exp(0.5 * x)
    x -= x.mean()
    ref = O.Engine(nside, J)
    ref.load(x)
    sx = ref.forward()
    cot = np.random.default_rng(9).standard_normal(sx.shape) * np.abs(sx)
    g_ref = ref.gradient(cot)
    print("[gpu] mixed precision nside=64", flush=True)
    values, _ = eng.forward(x)
    errs = family_errors(values, sx, eng.plan)
    print("mixed errors per family", errs)
    assert max(errs.values()) < 1e-5, errs
    grad = eng.vjp(cot)
    gerr = rel(grad, g_ref)
    print("mixed gradient error", gerr)
    assert gerr < 1e-4


def test_batch_equals_single_maps():
    nside, J = 16, 3
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 12 * nside * nside))
    e2 = engine(nside, J, batch=2)
    e1 = engine(nside, J, batch=1)
    v2, _ = e2.forward(x)
    cot = rng.standard_normal(v2.shape) + 1j * rng.standard_normal(v2.shape)
    g2 = e2.vjp(cot)
    for b in range(2):
        v1, _ = e1.forward(x[b])
        assert rel(v2[b], v1[0]) < 1e-14
        g1 = e1.vjp(cot[b])
        assert rel(g2[b], g1[0]) < 1e-14


def test_moments_match_oracle():
    nside, J = 16, 3
    eng = engine(nside, J)
    x = np.random.default_rng(4).standard_normal(12 * nside * nside)
    ref = O.Engine(nside, J)
    ref.load(x)
    sx, second = ref.forward(moments=True)
    var_ref = np.maximum(second - np.abs(sx) ** 2, 0.0)
    values, var = eng.forward(x, moments=True)
    assert max(family_errors(values, sx, eng.plan).values()) < STAT_BAR
    assert np.abs(var - var_ref).max() <= VAR_BAR * np.abs(var_ref).max()


@pytest.mark.parametrize("case", [c for c in cases("stats", "stats")
                                  if c["L"] == 4 and c["kernel_sz"] == 5 and c["channels"] == 1
                                  and not c["fewer"] and c["j_to_dg"] == list(range(c["J"]))],
                         ids=lambda c: c.name)
def test_forward_levels_fixtures(case):
    """stlp_forward_levels (lib.rs:583-623): the fixture's map of every level in (map{d}, the
    reference's own decimations), its statistics out, at the fixture tolerance; no VJP may
    follow.  Also against the oracle fed the same levels' decimation of level 0."""
    eng = engine(case["nside"], case["J"], case["batch"])
    levels = [case.arr(f"map{d}").reshape(case["batch"], -1) for d in range(case["J"])]
    print(f"[gpu] forward_levels {case.name}", flush=True)
    values = eng.forward_levels(levels)
    ref = case.complex("values").reshape(values.shape)
    flat = lambda z: np.concatenate([z.real.ravel(), z.imag.ravel()])
    assert np.max(np.abs(flat(values) - flat(ref)) / (1e-11 + 1e-10 * np.abs(flat(ref)))) <= 1.0
    full, _ = eng.forward(levels[0])
    assert max(family_errors(values, full, eng.plan).values()) < 1e-12
    with pytest.raises(RuntimeError, match="stale"):
        eng.forward_levels(levels)
        eng.vjp(values)


def test_identity_is_checked():
    nside, J = 16, 3
    eng = engine(nside, J)
    x = np.random.default_rng(5).standard_normal(12 * nside * nside)
    values, _ = eng.forward(x, generation=11, member=2, objective="a")
    with pytest.raises(RuntimeError):
        eng.vjp(values, generation=12, member=2, objective="a")
    with pytest.raises(RuntimeError):
        eng.vjp(values, generation=11, member=2, objective="b")
    lib = eng._lib.cdll
    key = (11, 3, 0)
    out = np.empty(eng.npix)
    c = np.ascontiguousarray(values.astype(np.complex128))
    assert lib.stlp_vjp(eng._h.ptr, c.ctypes.data_as(lib.stlp_vjp.argtypes[1]), c.size, *key,
                        out.ctypes.data_as(lib.stlp_vjp.argtypes[1])) == -1
    assert "stale" in lib.stlp_last_error().decode()
    eng.vjp(values, generation=11, member=2, objective="a")


def test_checkpoint_restore():
    """A VJP after restoring a checkpoint equals the oracle's gradient of the checkpointed
    forward (and a fresh forward's, bit for bit); the reference's bridge test of checkpoint and
    restore against fresh handles (crates/stl-bridge/tests/bridge.rs:202-240)."""
    nside, J = 16, 3
    eng = engine(nside, J)
    rng = np.random.default_rng(6)
    x, y = rng.standard_normal((2, 12 * nside * nside))
    vx, _ = eng.forward(x, generation=1, objective="x")
    snap = eng.checkpoint()
    assert snap.nbytes() > 0
    vy, _ = eng.forward(y, generation=1, objective="y")
    gy = eng.vjp(vy, generation=1, objective="y")
    eng.restore(snap, generation=1, objective="x")
    gx = eng.vjp(vx, generation=1, objective="x")
    ref = O.Engine(nside, J)
    ref.load(x)
    sx = ref.forward()
    assert max(family_errors(vx, sx, eng.plan).values()) < STAT_BAR
    assert rel(gx, ref.gradient(vx)) < GRAD_BAR
    vx2, _ = eng.forward(x, generation=2, objective="x")
    assert np.array_equal(vx, vx2)
    assert np.array_equal(eng.vjp(vx2, generation=2, objective="x"), gx)
    with pytest.raises(RuntimeError):
        eng.restore(snap, generation=3, objective="x")
    snap.close()
    assert not np.array_equal(gx, gy)


def test_torch_autograd_matches_vjp_and_finite_differences():
    """torch.autograd through the engine equals vjp; central differences of the
    pairing Re sum(conj(c) S(x)) along unit directions (eps 1e-4) match the
    directional derivative (the reference's stats.rs:77-110 check)."""
    from stl_port.torch_fn import scattering
    nside, J = 16, 3
    eng = engine(nside, J)
    rng = np.random.default_rng(10)
    x0 = torch.from_numpy(rng.standard_normal((1, 12 * nside * nside))).cuda()
    c = rng.standard_normal(eng.ncoeff) + 1j * rng.standard_normal(eng.ncoeff)
    ct = torch.from_numpy(c).cuda()

    def pairing(x):
        return (torch.conj(ct) * scattering(eng, x)).real.sum()

    x = x0.clone().requires_grad_(True)
    pairing(x).backward()
    g = x.grad.detach()
    eng.forward(x0.cpu().numpy(), generation=99)
    assert rel(g.cpu().numpy(), eng.vjp(c, generation=99)) < 1e-13
    eps = 1e-4
    for _ in range(4):
        v = torch.from_numpy(rng.standard_normal(x0.shape)).cuda()
        v /= torch.linalg.norm(v)
        with torch.no_grad():
            fd = (pairing(x0 + eps * v) - pairing(x0 - eps * v)).item() / (2 * eps)
        exact = float((g * v).sum())
        assert abs(fd - exact) <= 1e-6 * abs(exact), (fd, exact)


@pytest.mark.parametrize("J", [1, 2, 3, 4])
def test_graph_replays_match_eager_launches(J):
    """The forward and the VJP run as HIP graphs (captured per device pointer
    and moments flag; with 2 levels or more, levels 0 and 1's branch work on
    side streams) unless per-kernel timing is on; replays, interleaved with
    other work and other inputs, give the eager launches' bits (J = 1: one
    graph; J = 2: level 0 on a side stream; J >= 3: levels 0 and 1)."""
    from stl_port import NativeEngine
    nside = 32
    print(f"[gpu] graph vs eager nside={nside} J={J}", flush=True)
    rng = np.random.default_rng(21)
    xs = [rng.standard_normal((1, 12 * nside * nside)) for _ in range(2)]
    with NativeEngine(nside, J, precision="mixed") as g, NativeEngine(nside, J, precision="mixed",
                                                                      timing=True) as e:
        cots = rng.standard_normal((1, g.ncoeff)) + 1j * rng.standard_normal((1, g.ncoeff))
        xd = [torch.from_numpy(x).cuda() for x in xs]
        out_g, out_e = torch.empty_like(xd[0]), torch.empty_like(xd[0])
        for it, k in enumerate([0, 1, 0, 0, 1]):
            moments = it % 2 == 1
            vg, sg = g.forward(xs[k], generation=it, moments=moments)
            ve, se = e.forward(xs[k], generation=it, moments=moments)
            np.testing.assert_array_equal(vg, ve)
            np.testing.assert_array_equal(sg, se)
            np.testing.assert_array_equal(g.vjp(cots * (it + 1), generation=it),
                                          e.vjp(cots * (it + 1), generation=it))
            torch.cuda.synchronize()
            _ = (xd[k] * 2).sum().item()  # other stream work between replays
            vg, _ = g.forward_device(xd[k], generation=100 + it)
            ve, _ = e.forward_device(xd[k], generation=100 + it)
            np.testing.assert_array_equal(vg, ve)
            g.vjp_device(cots, out_g, generation=100 + it)
            e.vjp_device(cots, out_e, generation=100 + it)
            torch.cuda.synchronize()
            assert torch.equal(out_g, out_e)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_forward_conv_variants_agree(precision):
    """The convolution kernel variants (16 x 16 tiles, several faces per
    workgroup, 8 x 8 tiles, 32 x 32 tiles, slots split over lanes, several
    faces per workgroup on 32 x 32 tiles; chosen per level and kernel by
    timing unless forced) agree to rounding."""
    from stl_port import NativeEngine
    nside, J = 64, 5
    rng = np.random.default_rng(31)
    x = rng.standard_normal((1, 12 * nside * nside))
    out = []
    for v in (0, 1, 2, 3, 4, 5, None):
        print(f"[gpu] conv_variant={v} precision={precision}", flush=True)
        with NativeEngine(nside, J, precision=precision, conv_variant=v) as e:
            vals, _ = e.forward(x)
            cot = np.conj(vals) * 0.5 + 1.0
            out.append((vals, e.vjp(cot), e.info()["conv_variants"]))
            plan = e.plan
    for vals, grad, variants in out[1:]:
        if precision == "fp64":
            np.testing.assert_allclose(vals, out[0][0], rtol=1e-12, atol=0)
        else:  # FP32 sums in another order: per-family error, as for the mixed engine
            assert max(family_errors(vals, out[0][0], plan).values()) <= 1e-6
        assert rel(grad, out[0][1]) <= (1e-12 if precision == "fp64" else 1e-5)
    # levels of nside >= 8 are tiled; FP64 has no multi-face branch adjoint;
    # the split gather serves every level of nside <= 64
    assert all(row == [1, 1, 1, 1 if precision == "mixed" else 0] for row in out[1][2][:4])
    assert all(row == [4, 4, 4, 4] for row in out[4][2])
    # 32 x 32 multi-face kernels at nside >= 32 (none for FP64's branch kernels)
    assert out[5][2][:2] == [[5, 5, 5, 5] if precision == "mixed" else [5, 0, 5, 0]] * 2


def test_two_row_adjoint_with_odd_rows_agrees():
    """The FP32 32 x 32 branch adjoint takes two input rows per workgroup; an odd
    row count sends the last row to the one-row kernel on faces of >= 256 x 256
    pixels. Three maps at nside 256 give 3 rows at level 0 and 6 at level 1:
    forced to that variant, the mixed engine agrees with the 16 x 16 tiles and
    each map with its own one-map evaluation."""
    from stl_port import NativeEngine
    nside, J = 256, 3
    rng = np.random.default_rng(41)
    x = rng.standard_normal((3, 12 * nside * nside))
    out = {}
    for v in (0, 3):
        print(f"[gpu] two-row adjoint nside={nside} batch=3 conv_variant={v}", flush=True)
        with NativeEngine(nside, J, batch=3, precision="mixed", conv_variant=v) as e:
            assert e.info()["conv_variants"][0][3] == v
            vals, _ = e.forward(x)
            cot = np.conj(vals) * 0.5 + 1.0
            out[v] = (vals, e.vjp(cot))
            plan = e.plan
    assert max(family_errors(out[3][0], out[0][0], plan).values()) <= 1e-6
    assert rel(out[3][1], out[0][1]) <= 1e-5
    with NativeEngine(nside, J, batch=1, precision="mixed", conv_variant=3) as one:
        for b in range(3):
            vb, _ = one.forward(x[b:b + 1])
            assert max(family_errors(vb, out[3][0][b:b + 1], plan).values()) <= 1e-6
            gb = one.vjp(np.conj(vb) * 0.5 + 1.0)
            assert rel(gb, out[3][1][b:b + 1]) <= 1e-5


@pytest.mark.skip(reason="the work-queue path is switched off: its first two GPU runs did not finish "
                         "(the stats-fixture forward at nside 16 hung inside the queue launch)")
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_work_queue_levels_agree_with_level_launches(precision):
    """The coarse levels as one persistent work-queue launch per sweep (levels of nside <= 32
    from level 2 on, the default) against a launch per kernel and level (queue_nside=0), on
    a batch of 2 maps at nside 128 (J 6: levels 2..5 in the queue); and the queued FP64 engine
    against the oracle at the reference's device bars."""
    from stl_port import NativeEngine
    nside, J = 128, 6
    rng = np.random.default_rng(51)
    x = rng.standard_normal((2, 12 * nside * nside))
    out = {}
    for q in (32, 0):
        print(f"[gpu] work queue nside<={q} {precision}", flush=True)
        with NativeEngine(nside, J, batch=2, precision=precision, queue_nside=q) as e:
            assert e.info()["queue_level"] == (2 if q else J)
            vals, var = e.forward(x, moments=True)
            cot = rng.standard_normal(vals.shape) + 1j * rng.standard_normal(vals.shape) if q else out[32][3]
            out[q] = (vals, var, e.vjp(cot), cot)
            plan = e.plan
    if precision == "fp64":
        assert max(family_errors(out[32][0], out[0][0], plan).values()) < STAT_BAR
        assert rel(out[32][1], out[0][1]) < VAR_BAR
        assert rel(out[32][2], out[0][2]) < GRAD_BAR
        ref = O.Engine(nside, J, B=2)
        ref.load(x)
        sx = ref.forward()
        assert max(family_errors(out[32][0], sx, plan).values()) < STAT_BAR
        assert rel(out[32][2], ref.gradient(out[32][3])) < GRAD_BAR
    else:
        assert max(family_errors(out[32][0], out[0][0], plan).values()) < 1e-6
        assert rel(out[32][2], out[0][2]) < 1e-5


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_levels_agree_with_level_launches(precision):
    """The coarse levels batched (batch_nside=64: from level 2 on, the levels of at most
    12 * 64^2 pixels over the batch; each step of a sweep one launch for all of them) against a
    launch per kernel and level (batch_nside=0), on a batch of 2 maps at nside 128 (J 6: levels
    2..5 batched), graph replays and eager launches; and the batched FP64 engine against the
    oracle at the reference's device bars."""
    from stl_port import NativeEngine
    nside, J = 128, 6
    rng = np.random.default_rng(51)
    x = rng.standard_normal((2, 12 * nside * nside))
    cot = None
    out = {}
    for bn, timing in ((64, False), (64, True), (0, False)):
        print(f"[gpu] batched levels nside<={bn} timing={timing} {precision}", flush=True)
        with NativeEngine(nside, J, batch=2, precision=precision, batch_nside=bn, timing=timing) as e:
            assert e.info()["batch_level"] == (2 if bn else J)
            vals, var = e.forward(x, moments=True)
            if cot is None:
                cot = rng.standard_normal(vals.shape) + 1j * rng.standard_normal(vals.shape)
            out[(bn, timing)] = (vals, var, e.vjp(cot))
            plan = e.plan
    for a, b in zip(out[(64, False)], out[(64, True)]):
        np.testing.assert_array_equal(a, b)
    got, ref0 = out[(64, False)], out[(0, False)]
    if precision == "fp64":
        assert max(family_errors(got[0], ref0[0], plan).values()) < STAT_BAR
        assert rel(got[1], ref0[1]) < VAR_BAR
        assert rel(got[2], ref0[2]) < GRAD_BAR
        ref = O.Engine(nside, J, B=2)
        ref.load(x)
        sx = ref.forward()
        assert max(family_errors(got[0], sx, plan).values()) < STAT_BAR
        assert rel(got[2], ref.gradient(cot)) < 2 * GRAD_BAR
    else:
        assert max(family_errors(got[0], ref0[0], plan).values()) < 1e-6
        assert rel(got[2], ref0[2]) < 1e-5
