"""Several plans and batches on one engine's fields (stlx_share, NativeEngine.share): the
reference's GpuEvaluator::share (crates/stl-gpu/src/engine.rs:1909-1916) with its holder rule
(a gradient needs the last forward on the fields to be its own, :1938-1943).

Every shared engine must give what a standalone engine of its batch, cross flag and families
gives (FP64, 1e-14 relative to the largest entry: the sums may be split differently over
workgroups), refuse a VJP after another engine's forward, and outlive its owner.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

NSIDE, J = 32, 4
NPIX = 12 * NSIDE ** 2


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def rel(a, b):
    return float(np.abs(np.ravel(a) - np.ravel(b)).max() / np.abs(b).max())


def test_shared_engines_match_standalone_ones():
    from stl_port import NativeEngine
    rng = np.random.default_rng(11)
    x3 = rng.standard_normal((3, NPIX))
    print(f"[gpu] share nside={NSIDE} J={J}", flush=True)
    owner = NativeEngine(NSIDE, J, batch=3, precision="fp64")
    one = owner.share(batch=1)
    subset = owner.share(batch=2, families=["S1", "S2"])
    pair = owner.share(batch=1, cross=True)
    with NativeEngine(NSIDE, J, batch=1, precision="fp64") as ref1, \
            NativeEngine(NSIDE, J, batch=1, cross=True, precision="fp64") as refx:
        v1, _ = one.forward(x3[:1])
        r1, _ = ref1.forward(x3[:1])
        assert rel(v1, r1) < 1e-14
        cot1 = rng.standard_normal(v1.shape) + 1j * rng.standard_normal(v1.shape)
        assert rel(one.vjp(cot1), ref1.vjp(cot1)) < 1e-13

        # A plan of some families: the full plan's rows of those families, in order.
        sel = [i for i, c in enumerate(owner.plan) if c["family"] in ("S1", "S2")]
        assert [dict(c, index=0) for c in subset.plan] == [dict(owner.plan[i], index=0) for i in sel]
        vs, _ = subset.forward(x3[1:])
        vo, _ = owner.forward(x3)
        assert rel(vs, vo[1:, sel]) < 1e-14
        cot = rng.standard_normal(vo.shape) + 1j * rng.standard_normal(vo.shape)
        cot[:, [i for i in range(owner.ncoeff) if i not in sel]] = 0.0
        go = owner.vjp(cot)
        vs, _ = subset.forward(x3[1:])
        assert rel(subset.vjp(cot[1:, sel]), go[1:]) < 1e-13

        # Pairs on the same fields (two field rows per pair).
        vx, _ = pair.forward(x3[:2][None])
        rx, _ = refx.forward(x3[:2][None])
        assert rel(vx, rx) < 1e-14

    # Turns: after another engine's forward, a VJP is refused until the next own forward.
    owner.forward(x3)
    one.forward(x3[:1])
    with pytest.raises(RuntimeError, match="forward sweep of this evaluator"):
        owner.vjp(cot)
    owner.forward(x3)
    assert rel(owner.vjp(cot), go) < 1e-14

    # More maps than the fields hold.
    with pytest.raises(RuntimeError, match="more maps than the shared fields"):
        owner.share(batch=4)
    with pytest.raises(RuntimeError, match="more maps than the shared fields"):
        owner.share(batch=2, cross=True)

    # The owner may go first.
    owner.close()
    v1b, _ = one.forward(x3[:1])
    assert np.array_equal(v1b, v1)
    for e in (one, subset, pair):
        e.close()


def test_shared_family_subset_matches_the_oracle():
    """A shared engine with a plan of some families (S3 and S4 only, batch 1 of the owner's 2)
    against the oracle's complete plan restricted to those rows, at the reference's device bars
    (statistics 1e-13, gradient 1e-12 relative to the largest entry; crates/stl-gpu/tests/
    stats.rs:248-275): values, and the gradient of a cotangent on those rows only."""
    import oracle as O
    from stl_port import NativeEngine
    rng = np.random.default_rng(12)
    x2 = rng.standard_normal((2, NPIX))
    print(f"[gpu] share families S3,S4 nside={NSIDE} J={J} vs oracle", flush=True)
    with NativeEngine(NSIDE, J, batch=2, precision="fp64") as owner:
        sub = owner.share(batch=1, families=["S3", "S4"])
        sel = np.array([i for i, c in enumerate(owner.plan) if c["family"] in ("S3", "S4")])
        owner.forward(x2)  # the owner's forward first: the shared engine then takes its turn
        vs, _ = sub.forward(x2[1:])
        cot = rng.standard_normal(vs.shape) + 1j * rng.standard_normal(vs.shape)
        gs = sub.vjp(cot)
        sub.close()
    ref = O.Engine(NSIDE, J)
    ref.load(x2[1])
    full = ref.forward()
    assert vs.shape == (1, sel.size)
    assert rel(vs[0], full[0, sel]) < 1e-13
    cfull = np.zeros_like(full)
    cfull[0, sel] = cot[0]
    gref = ref.gradient(cfull)
    assert rel(gs, gref) < 1e-12
