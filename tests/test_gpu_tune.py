"""Launch configurations and the autotuner.  Every configuration of the fused
kernel set (lane-walking or one-cell kernels, cells per lane, z-line
exchange, workgroup order, CT fusion of none, two or all three sweeps) must
give the same bits, and autotuning must
leave the state exactly where it was (it saves and restores each member's
device state around the timed candidates)."""
import pytest

import sams_amd.lare as L
from lare_testlib import bit_diff

pytestmark = pytest.mark.gpu

CONFIGS = [
    dict(walk=1, march_k=8, zx_k=8, zline=1, order=1, ct_fuse=1),
    dict(walk=1, march_k=16, zx_k=16, zline=0, order=0, ct_fuse=1),
    dict(walk=1, march_k=8, zx_k=16, zline=1, order=0, ct_fuse=0),
    dict(walk=2, march_k=16, zx_k=8, zline=1, order=1, ct_fuse=1),
    dict(walk=1, march_k=16, zx_k=8, zline=1, order=1, ct_fuse=2),
    dict(walk=1, march_k=8, zx_k=16, zline=0, order=0, ct_fuse=2),
]


def _same(a, b):
    for f in L.STATE_FIELDS:
        d = bit_diff(a.download(f), b.download(f))
        if d:
            return f"{f}: {d}"
    return None


def test_every_configuration_gives_the_same_bits():
    p = L.problem("bw", n=26, ny=9, nz=7, seed=11, noise=0.02, eta=1e-3)
    ref = L.Lare.from_problem(p)
    dts = ref.steps(5)
    for cfg in CONFIGS:
        g = L.Lare.from_problem(p)
        g.set_tune(**cfg)
        got = g.get_tune()
        assert all(got[k] == v for k, v in cfg.items()), (cfg, got)
        assert [d.hex() for d in g.steps(5)] == [d.hex() for d in dts], cfg
        assert _same(g, ref) is None, cfg


def test_autotune_keeps_the_state_and_continues_identically():
    p = L.problem("ot", n=24, ny=20, nz=18, seed=5, amp=0.04)
    tuned = L.Lare.from_problem(p)
    plain = L.Lare.from_problem(p)
    for g in (tuned, plain):
        g.steps(2)
    t = tuned.autotune(steps=2)
    assert t["ms_per_step"] > 0.0
    assert t["walk"] in (0, 1, 2) and t["march_k"] in (8, 16) and t["zx_k"] in (8, 16)
    assert tuned.get_tune()["walk"] == t["walk"]
    assert _same(tuned, plain) is None, "state changed by autotuning"
    assert tuned.get_dt().hex() == plain.get_dt().hex()
    assert [d.hex() for d in tuned.steps(4)] == [d.hex() for d in plain.steps(4)]
    assert _same(tuned, plain) is None


def test_batch_autotune_keeps_every_member():
    ps = [L.problem("ot", n=16, ny=12, nz=10, seed=2, amp=0.03),
          L.problem("bw", n=30, ny=6, nz=5, seed=3, noise=0.01, eta=1e-3),
          L.problem("ot", n=9, ny=11, nz=13, seed=4, amp=0.05)]
    mem = [L.Lare.from_problem(p) for p in ps]
    alone = [L.Lare.from_problem(p) for p in ps]
    b = L.Batch(mem)
    b.steps(2)
    b.sync()
    t = b.autotune(steps=2)
    assert t["ms_per_step"] > 0.0 and b.get_tune()["order"] == t["order"]
    b.steps(3)
    b.sync()
    for i, g in enumerate(alone):
        g.steps(5)
        assert _same(mem[i], g) is None, i


def test_invalid_configuration_is_an_error():
    g = L.Lare.from_problem(L.problem("ot", n=8))
    with pytest.raises(RuntimeError, match="invalid configuration"):
        g.set_tune(march_k=5)
    with pytest.raises(RuntimeError, match="invalid configuration"):
        g.set_tune(ct_fuse=3)
    g.set_kernel_set(0)
    with pytest.raises(RuntimeError, match="kernel set 1"):
        g.autotune()
