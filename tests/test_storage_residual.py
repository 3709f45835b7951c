"""Measurement behind a design decision: how much storing each field block
as per-brick FP32 bases plus 16-bit residuals would cut the 16-bit storage
error of the headline workload, on the CPU (the oracle traces the decoded
values; the same restatement gives the GPU's BF16 storage error exactly).

One rank's share of ABC 512^3 (7 x 129^3 blocks), every 32nd of 262,144
seeds, 1,000 frozen RK4 steps, dt 0.005, union of 7, |u| < 1e-6.  Endpoint
error against the FP32 field, max over the seeds with the same step count,
relative to the block extent:

  bf16 2.0e-2, fp16 7.4e-3;  bf16 + 8^3-brick bases 2.6e-3, 4^3 4.2e-3;
  fp16 + 8^3-brick bases 1.0e-3, 4^3 5.1e-4.

The residual form shrinks the storage error (RMS 6x for BF16, 12x for FP16
at 4^3), but only FP16 residuals reach 1e-3 of the extent, and the 16-bit
forms do not gain speed on a latency-bound trace; the form is therefore not
built (the test keeps the measurement reproducible)."""
import numpy as np
import pytest

from pars_amd import workloads as W

f32 = np.float32


def _bf16(a):
    u = a.astype(np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16  # round to nearest even
    return r.astype(np.uint32).view(np.float32)


def _f16(a):
    return a.astype(np.float16).astype(np.float32)


def _residual(a, q, brick):
    out = np.empty_like(a)
    X, Y, Z, _ = a.shape
    for i in range(0, X, brick):
        for j in range(0, Y, brick):
            for k in range(0, Z, brick):
                blk = a[i:i + brick, j:j + brick, k:k + brick]
                base = blk.reshape(-1, 3).mean(0).astype(np.float32)
                out[i:i + brick, j:j + brick, k:k + brick] = (base + q(blk - base)).astype(np.float32)
    return out


@pytest.fixture(scope="module")
def share_run(orc):
    sh = W.share(512, 64, 21)
    blocks = [sh.local] + list(sh.neighbors)
    raw = [W.abc_block_f64(b, sh.h).astype(np.float32) for b in blocks]
    seeds = orc.seed_generate(sh.local.offset, sh.local.size, sh.local.spacing * f32(2), 1002, 21)[::32].copy()

    def ends(arrs):
        fs = [orc.Field(a, b.offset, b.size, b.spacing) for a, b in zip(arrs, blocks)]
        return orc.endpoints(seeds, fs[0], neighbors=fs[1:], neighbor_rank=sh.neighbor_ranks,
                             face_masks=sh.face_masks, dt=0.005, integrator=orc.RK4, union=True,
                             stop_rule=orc.STOP_NORM, stop_eps=1e-6)

    e32, s32 = ends(raw)
    extent = float(sh.local.size[0])

    def err(arrs):
        e, s = ends(arrs)
        same = s == s32
        d = np.linalg.norm(e[same, :3].astype(np.float64) - e32[same, :3], axis=1)
        return d.max() / extent, np.sqrt((d ** 2).mean())

    return raw, err


def test_residual_storage_errors(share_run):
    raw, err = share_run
    plain_bf, rms_bf = err([_bf16(a) for a in raw])
    plain_f16, rms_f16 = err([_f16(a) for a in raw])
    assert 1.5e-2 < plain_bf < 2.5e-2 and 5e-3 < plain_f16 < 1e-2  # the GPU forms' storage errors
    res_bf8, rms_bf8 = err([_residual(a, _bf16, 8) for a in raw])
    res_f16_4, rms_f16_4 = err([_residual(a, _f16, 4) for a in raw])
    assert rms_bf8 < rms_bf / 4 and rms_f16_4 < rms_f16 / 8  # the residual form cuts the typical error
    assert res_bf8 > 1e-3  # but BF16 residuals stay above 1e-3 of the extent
    assert res_f16_4 < 1e-3  # FP16 residuals on 4^3 bricks get below it
