"""How much the one unpinned detail of the stepper can move a trajectory.

The oracle's runge_kutta_4 step rounds its coefficients as Boost.Odeint
1.69's runge_kutta4 does (explicit_generic_rk: the float products
(1/6)*dt, (1/3)*dt; oracle/pa_oracle.cpp, step_frozen).  Boost is not in
the reference tree, so that rounding is restated from odeint's published
source, not checked against it.  The plausible alternative is
runge_kutta4_classic's dt/6, dt/3 (one ulp apart at dt = 0.01 and 0.005).
These tests trace the two workloads of the tracer with both forms on the
CPU and bound the share of trajectories that the alternative would move
past the path's 1e-5 relative bar (or to another step count): that share
is the parity at risk if the restatement were wrong.  Measured here:
ABC 64^3, 4,096 seeds, 256 steps: 6 trajectories (0.15%), largest 2.8e-3;
one rank's share of ABC 512^3, 262,144 seeds, 1,000 steps, |u| < 1e-6:
own block 46 (0.018%), union of 7 blocks 123 (0.047%)."""
import numpy as np
import pytest

from pars_amd import workloads as W

f32 = np.float32


def _compare(alt, alt_steps, ref, ref_steps):
    rel = (np.abs(alt[:, :3].astype(np.float64) - ref[:, :3].astype(np.float64)) /
           np.maximum(np.abs(ref[:, :3].astype(np.float64)), 1.0)).max(1)
    moved = (alt_steps != ref_steps) | (rel > 1e-5)
    changed = np.any(alt.view(np.uint32) != ref.view(np.uint32), axis=1)
    return dict(changed=float(changed.mean()), outside=int(moved.sum()), max_rel=float(rel.max()))


def test_coefficient_forms_differ(orc):
    """The two roundings are different numbers at the configured step sizes,
    and the oracle's two steppers do differ on some inputs, so the
    measurement below compares two distinct steppers."""
    rng = np.random.default_rng(3)
    for dt in (0.01, 0.005):
        assert f32(f32(1) / f32(6)) * f32(dt) != f32(dt) / f32(6)
        differ = 0
        for _ in range(500):  # positions near 0, where one ulp of a coefficient shows
            p = np.append(rng.uniform(0, 1e-3, 3), 0).astype(np.float32)
            v = rng.uniform(-2, 2, 3).astype(np.float32)
            a = orc.step(orc.RK4, p, v, dt)
            b = orc.step(orc.RK4_DT_OVER_6, p, v, dt)
            differ += a.tobytes() != b.tobytes()
        assert differ > 0


def test_rounding_risk_single_block(orc):
    """ABC 64^3 FP32, 4,096 seeds (stride 4), R = 258, dt 0.01."""
    n = 64
    h = 2 * np.pi / (n - 1)
    hf = f32(h)
    info = W.BlockInfo(0, (0, 0, 0), (0, 0, 0), (n, n, n))
    size = [f32(n) * hf] * 3
    F = orc.Field(W.abc_block_f64(info, h).astype(np.float32), (0, 0, 0), size, [hf] * 3)
    seeds = orc.seed_generate((0, 0, 0), size, [f32(4) * hf] * 3, 258, 0)
    a, sa = orc.endpoints(seeds, F, dt=0.01, integrator=orc.RK4)
    b, sb = orc.endpoints(seeds, F, dt=0.01, integrator=orc.RK4_DT_OVER_6)
    r = _compare(b, sb, a, sa)
    print("single block:", r)
    assert r["changed"] > 0.5           # most trajectories change bits
    assert 0 < r["outside"] <= 12       # a handful leave the 1e-5 bar (measured 6 of 4,096)
    assert r["max_rel"] < 1e-2


@pytest.mark.parametrize("union", [False, True], ids=["own-block", "union-of-7"])
def test_rounding_risk_share(orc, union):
    """One rank's share: 7 x 129^3 ABC blocks, 262,144 seeds (stride 2),
    R = 1002, dt 0.005, |u| < 1e-6."""
    sh = W.share(512, 64, 21)
    blocks = [sh.local] + list(sh.neighbors)
    fs = [orc.Field(W.abc_block_f64(b, sh.h).astype(np.float32), b.offset, b.size, b.spacing) for b in blocks]
    seeds = orc.seed_generate(sh.local.offset, sh.local.size, sh.local.spacing * f32(2), 1002, 21)
    kw = dict(neighbors=fs[1:], neighbor_rank=sh.neighbor_ranks, face_masks=sh.face_masks, dt=0.005, union=union,
              stop_rule=orc.STOP_NORM, stop_eps=1e-6)
    a, sa = orc.endpoints(seeds, fs[0], integrator=orc.RK4, **kw)
    b, sb = orc.endpoints(seeds, fs[0], integrator=orc.RK4_DT_OVER_6, **kw)
    r = _compare(b, sb, a, sa)
    print("share (%s):" % ("union" if union else "own block"), r)
    assert r["changed"] > 0.1
    assert r["outside"] <= 0.001 * len(seeds)  # measured 46 (own block) and 123 (union) of 262,144
