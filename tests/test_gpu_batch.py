"""Batched persistent forward launch, FDK plans and hipGraph replay: results equal the per-problem plain path."""
import numpy as np
import pytest

from oracle import oracle as O
from virtualcbct_amd import io, phantoms

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as T
    if not T.cuda.is_available():
        pytest.skip("no GPU")
    return T


def problems(torch):
    """Three phantoms of uneven slice count, angle count and energy bins."""
    from virtualcbct_amd.projector import Projector
    out = []
    rng = np.random.default_rng(21)
    for (N, Nz, nA, H, W, NE, seed) in ((64, 12, 20, 72, 64, 9, 1), (64, 20, 33, 72, 64, 5, 2), (48, 7, 12, 40, 56, 14, 3)):
        vol = phantoms.ellipsoid_phantom(N, Nz, 0.3, 0.4, seed, n_ell=6, n_labels=4)
        ang = np.deg2rad(-180.0 + 360.0 / nA * np.arange(nA) + 1e-4)
        cs, sn, xc, yc = io.trig_table(W, H, ang, np.full(nA, 0.05), np.zeros(nA))
        P = Projector(N, 0.3, Nz, 0.4, H, W, nA, cs, sn, xc, yc)
        P.set_volume(vol)
        atten = rng.uniform(0.15, 0.6, (6, NE)).astype(np.float32)
        pdf = rng.uniform(0.1, 1.0, NE).astype(np.float32)
        P.set_spectrum(atten, pdf)
        g = O.Geometry(N=N, s=0.3, Nz=Nz, t=0.4, H=H, W=W, cos_t=cs, sin_t=sn, xc=xc, yc=yc)
        out.append((P, vol, atten, pdf, g))
    return out


def test_batch_forward_equals_plain(torch):
    from virtualcbct_amd.batch import Batch, autotune_log
    probs = problems(torch)
    plain = [P.forward(1000.0).clone() for P, *_ in probs]
    b = Batch()
    outs = []
    for P, *_ in probs:
        o = torch.full(P.shape_proj, -1.0, device="cuda")
        b.add_forward(P, o, 1000.0)
        outs.append(o)
    b.run()
    torch.cuda.synchronize()
    for o, p in zip(outs, plain):
        assert torch.equal(o, p)
    assert b.blocks >= 0  # 0 = flat grid chosen, > 0 = persistent grid size
    print(autotune_log())
    # one of them against the oracle
    P, vol, atten, pdf, g = probs[2]
    ref = O.forward_collapsed(g, vol, atten, pdf, 1000.0)
    assert np.max(np.abs(outs[2].cpu().numpy() - ref) / ref) < 1e-4


def test_batch_graph_replays_match_eager(torch):
    from virtualcbct_amd.batch import Batch
    from virtualcbct_amd.fdk import FdkPlan
    probs = problems(torch)
    b = Batch()
    outs, recs, eager = [], [], []
    for i, (P, vol, *_rest) in enumerate(probs):
        o = torch.empty(P.shape_proj, device="cuda")
        b.add_forward(P, o, 1000.0)
        g = P.g
        ang = np.deg2rad(-180.0 + 360.0 / g.n_angles * np.arange(g.n_angles) + 1e-4)
        plan = FdkPlan(g.proj_height, g.proj_width, g.n_angles, 32, 6, -1, 0.3, 0.5, 1000, ang,
                       np.zeros(g.n_angles), np.zeros(g.n_angles))
        filt = torch.empty(P.shape_proj, device="cuda")
        rec = torch.empty((6, 32, 32), device="cuda")
        b.add_fdk(plan, o, filt, rec)
        outs.append(o)
        recs.append(rec)
        # eager reference of the same chain
        pe = P.forward(1000.0)
        eager.append((pe, plan.execute(pe)))
    torch.cuda.synchronize()
    b.capture()
    for rep in range(3):
        for o, r in zip(outs, recs):
            o.fill_(0.0)
            r.fill_(123.0)
        torch.cuda.synchronize()
        b.replay()
        torch.cuda.synchronize()
        for (o, r), (pe, re) in zip(zip(outs, recs), eager):
            assert torch.equal(o, pe), f"replay {rep}: projections differ"
            assert torch.equal(r, re), f"replay {rep}: reconstruction differs"
        # eager work between replays
        _ = probs[0][0].forward(1000.0)
        torch.cuda.synchronize()


def test_fdk_plan_equals_one_shot(torch):
    from virtualcbct_amd import fdk
    rng = np.random.default_rng(4)
    n, H, W = 24, 48, 64
    proj = torch.from_numpy(rng.uniform(0.1, 1.0, (n, H, W)).astype(np.float32)).cuda()
    ang = np.deg2rad(np.arange(n) * 15.0 - 170.0001)
    xs = np.full(n, 0.4)
    ys = np.full(n, -0.2)
    plan = fdk.FdkPlan(H, W, n, 40, 9, -2, 0.45, 0.4, 1000, ang, xs, ys)
    assert plan.slices_per_thread in (2, 4, 8)
    a = plan.execute(proj)
    filt = fdk.filter_sinogram_device(proj, xs, ys)
    b = torch.zeros((9, 40, 40), device="cuda")
    fdk.reconstruct_image_device(b, filt, -2, 0.45, 0.4, 1000, ang, xs)
    torch.cuda.synchronize()
    assert torch.allclose(a, b, rtol=1e-6, atol=1e-4)
    ref = O.reconstruct_image(40, 9, -2, 0.45, 0.4, O.filter_sinogram(proj.cpu().numpy(), xs, ys), 1000, ang, xs)
    assert np.abs(a.cpu().numpy() - ref).max() <= 1e-4 * np.abs(ref).max() + 1e-3


def test_every_grid_form_equals_plain_mixed_voxel_types(torch):
    """u16 volumes and FP32 maps in one batch; the flat grid and persistent grids (including one far smaller
    than the work, so that waves drain their own XCD queue and steal from the others), with columns or angles
    divided among the XCDs, and the flat grid's 4-wave block shapes and orders, all reproduce the plain
    per-problem launches bit for bit (the problems' angle counts, 20, 33 and 12, are not multiples of 8)."""
    from virtualcbct_amd.batch import Batch
    probs = problems(torch)
    # a fourth problem with FP32 maps (the differentiable path's voxel type)
    P0, vol0, atten0, pdf0, g0 = probs[0]
    from virtualcbct_amd.projector import Projector
    Pf = Projector(g0.N, g0.s, g0.Nz, g0.t, g0.H, g0.W, g0.n_ang, g0.cos_t, g0.sin_t, g0.xc, g0.yc)
    maps = torch.from_numpy(np.ascontiguousarray(vol0.transpose(1, 0, 2, 3)).astype(np.float32) / 1000.0).cuda()
    Pf.set_volume(maps)
    Pf.set_spectrum(atten0, pdf0)
    allp = [p[0] for p in probs[:2]] + [Pf] + [probs[2][0]]  # interleaved types: the batch groups them
    plain = [P.forward(1000.0).clone() for P in allp]
    cases = [(grid, (1, 4, 0, split)) for grid in (0, 24, 512, 1536) for split in (0, 1)]
    cases += [(0, shape) for shape in ((2, 2, 0, 0), (2, 2, 1, 1), (4, 1, 0, 1), (4, 1, 1, 0), (1, 4, 1, 1))]
    cases += [(-2, (1, 4, 0, 0))]  # the sequence of plain launches
    for grid, shape in cases:
        print("grid", grid, "tile", shape, flush=True)
        b = Batch()
        b.set_grid(grid)
        b.set_tile(*shape)
        outs = []
        for P in allp:
            o = torch.full(P.shape_proj, -1.0, device="cuda")
            b.add_forward(P, o, 1000.0)
            outs.append(o)
        b.run()
        torch.cuda.synchronize()
        assert b.blocks == grid and b.tile == shape
        for i, (o, p) in enumerate(zip(outs, plain)):
            assert torch.equal(o, p), (grid, shape, i)
        b.close()
    b = Batch()
    for bad in ((1, 2, 0, 0), (2, 4, 0, 0), (1, 4, 0, 2)):
        with pytest.raises(Exception, match="tile shape"):
            b.set_tile(*bad)
    b.close()
    Pf.close()


def test_fdk_plan_untuned_equals_one_shot(torch, monkeypatch):
    """With autotuning off the plan backprojects with 4 slices per thread, as the one-shot call does: the
    results are bit-identical, and within the FDK bar of the FP64 oracle."""
    from virtualcbct_amd import fdk
    monkeypatch.setenv("VCBCT_NO_AUTOTUNE", "1")
    rng = np.random.default_rng(5)
    n, H, W = 20, 44, 60  # a geometry no other test tunes (the tuner caches per geometry)
    proj = torch.from_numpy(rng.uniform(0.1, 1.0, (n, H, W)).astype(np.float32)).cuda()
    ang = np.deg2rad(np.arange(n) * 18.0 - 170.0001)
    xs = np.full(n, -0.3)
    ys = np.full(n, 0.1)
    plan = fdk.FdkPlan(H, W, n, 36, 7, -1, 0.45, 0.4, 1000, ang, xs, ys)
    assert plan.slices_per_thread == 4
    a = plan.execute(proj)
    filt = fdk.filter_sinogram_device(proj, xs, ys)
    b = torch.zeros((7, 36, 36), device="cuda")
    fdk.reconstruct_image_device(b, filt, -1, 0.45, 0.4, 1000, ang, xs)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    ref = O.reconstruct_image(36, 7, -1, 0.45, 0.4, O.filter_sinogram(proj.cpu().numpy(), xs, ys), 1000, ang, xs)
    assert np.abs(a.cpu().numpy() - ref).max() <= 1e-4 * np.abs(ref).max() + 1e-3
    plan.close()


@pytest.mark.parametrize("side", [1.0, -1.0])
def test_fdk_plan_half_fan(torch, side):
    """A half-fan plan weights and pads the projections inside its filter's load (2w (-log p) for a real column,
    0 for a padding one): its filtered panel equals the one-shot weighting + filter (-log of the FP32 p^(2w)) to
    FP32 rounding, its reconstruction the one-shot half-fan FDK and the oracle's chain (FP64 filter and
    backprojection of the oracle's weighted, padded panel); captured in a batch graph, every replay equals the
    eager result.  Both shift signs: the padding after the real columns, and before them (the filter's load
    then reads the caller's column k - P for panel column k)."""
    from virtualcbct_amd import fdk
    from virtualcbct_amd.batch import Batch
    rng = np.random.default_rng(6)
    n, H, W = 24, 40, 64
    proj = torch.from_numpy(rng.uniform(0.1, 1.0, (n, H, W)).astype(np.float32)).cuda()
    ang = np.deg2rad(np.arange(n) * 15.0 - 170.0001)
    xs = side * (7.5 + 0.05 * rng.standard_normal(n))  # ~12 of 32 half-panel pixels off centre (0.625 cm per pixel)
    ys = np.full(n, -0.2)
    plan = fdk.FdkPlan(H, W, n, 40, 9, -2, 0.45, 0.4, 1000, ang, xs, ys, half_fan=True)
    assert plan.width == fdk.half_fan_width(W, xs) > W
    filt = plan.filter(proj)
    pw, xo, yo, pit = fdk.offset_weight_device(proj, xs, ys)
    one_f = fdk.filter_sinogram_device(pw, xo, yo, pit)
    ef = float((filt - one_f).abs().max()) / float(one_f.abs().max())
    print(f"half-fan plan filter vs one-shot weighting + filter: {ef:.3e} of max")
    assert ef < 1e-5
    rec = torch.zeros(plan.shape_recon, device="cuda")
    plan.backproject(filt, rec)
    one = fdk.fdk_half_fan_device(proj, -2, 0.45, 0.4, 40, 9, 1000, ang, xs, ys)
    torch.cuda.synchronize()
    assert torch.allclose(rec, one, rtol=1e-5, atol=1e-3)
    pw_o, xo_o, yo_o = O.offset_weight(proj.cpu().numpy(), xs, ys)
    ref = O.reconstruct_image(40, 9, -2, 0.45, 0.4, O.filter_sinogram(pw_o, xo_o, yo_o, pitch_iso=pit), 1000, ang,
                              xo_o, pitch_iso=pit)
    err = np.abs(rec.cpu().numpy() - ref).max()
    print(f"half-fan plan vs oracle chain: {err:.3e} (max |ref| {np.abs(ref).max():.3e})")
    assert err <= 1e-4 * np.abs(ref).max() + 1e-3
    b = Batch()
    filt2 = torch.empty(plan.shape_filtered, device="cuda")
    rec2 = torch.empty(plan.shape_recon, device="cuda")
    b.add_fdk(plan, proj, filt2, rec2)
    torch.cuda.synchronize()
    b.capture()
    for rep in range(3):
        filt2.fill_(0.0)
        rec2.fill_(123.0)
        torch.cuda.synchronize()
        b.replay()
        torch.cuda.synchronize()
        assert torch.equal(filt2, filt), f"replay {rep}: filtered panel differs"
        assert torch.equal(rec2, rec), f"replay {rep}: reconstruction differs"
        _ = plan.execute(proj)  # eager work between replays
        torch.cuda.synchronize()
    b.close()


@pytest.mark.parametrize("W,shift_px,L", [(300, 70.0, 1024), (400, 115.0, 1280)])
def test_fdk_plan_half_fan_filter_lengths(torch, W, shift_px, L):
    """The half-fan plan's fused filter on the transform lengths compiled in: 300 columns padded to 440 (a
    1024-point transform) and 400 columns padded to 630 (1280 = 5 x 256, the 115 mm offset's panel).  Equal to the
    one-shot weighting + filter to FP32 rounding, and to the oracle's filter of the oracle's weighted panel."""
    from virtualcbct_amd import fdk
    rng = np.random.default_rng(W)
    n, H = 5, 6
    proj = torch.from_numpy(rng.uniform(0.1, 1.0, (n, H, W)).astype(np.float32)).cuda()
    ang = np.deg2rad(np.arange(n) * 40.0 - 90.0)
    xs = (shift_px + 0.5 * rng.standard_normal(n)) * 40.0 / W  # the reference's unit: W / 40 columns
    ys = np.full(n, 0.3)
    plan = fdk.FdkPlan(H, W, n, 16, 2, 0, 0.45, 0.4, 1000, ang, xs, ys, half_fan=True)
    Wp = plan.width
    Lp = fdk.filter_fft_length(Wp)
    print(f"W {W}: padded to {Wp} columns, transform length {Lp}")
    assert Wp == fdk.half_fan_width(W, xs) and Lp == L
    filt = plan.filter(proj)
    pw, xo, yo, pit = fdk.offset_weight_device(proj, xs, ys)
    one_f = fdk.filter_sinogram_device(pw, xo, yo, pit)
    ef = float((filt - one_f).abs().max()) / float(one_f.abs().max())
    pw_o, xo_o, yo_o = O.offset_weight(proj.cpu().numpy(), xs, ys)
    ref = O.filter_sinogram(pw_o, xo_o, yo_o, pitch_iso=pit, acc64=True)
    eo = float(np.abs(filt.cpu().numpy() - ref).max() / np.abs(ref).max())
    print(f"  vs one-shot weighting + filter {ef:.3e}, vs the oracle's FP64 chain {eo:.3e} of max")
    plan.close()
    assert ef < 1e-5 and eo < 5e-5


@pytest.mark.parametrize("R,Z,scale,n,H,W", [
    (40, 9, 0.4, 24, 48, 64),   # 6.4 cm blocks: footprints wider than the LDS tile, taps read from global memory
    (37, 5, 0.05, 30, 33, 50),  # 0.8 cm blocks: footprints staged; odd sizes, partial blocks
    (64, 12, 0.1, 36, 64, 96),  # footprints staged; some columns near the panel edges
])
def test_fdk_plan_staged_backprojector_equals_unstaged(torch, monkeypatch, R, Z, scale, n, H, W):
    """The LDS-staged backprojector reads the same taps as the unstaged one (from its tile, or from global
    memory where a tap falls outside it), so the reconstructions are equal bit for bit, for any slices per
    thread the tuner picks (each slice's sum runs over the angles in the same order)."""
    from virtualcbct_amd import fdk
    rng = np.random.default_rng(R * 100 + n)
    proj = torch.from_numpy(rng.uniform(0.1, 1.0, (n, H, W)).astype(np.float32)).cuda()
    ang = np.deg2rad(np.arange(n) * (360.0 / n) - 170.0001)
    xs = rng.uniform(-0.5, 0.5, n)
    ys = np.full(n, 0.1)
    recs = {}
    for lds in ("0", "1"):
        monkeypatch.setenv("VCBCT_BP_LDS", lds)
        plan = fdk.FdkPlan(H, W, n, R, Z, -1, 0.3, scale, 1000, ang, xs, ys)
        assert plan.bp_staged == (lds == "1")
        filt = plan.filter(proj)
        rec = torch.zeros(plan.shape_recon, device="cuda")
        plan.backproject(filt, rec)
        torch.cuda.synchronize()
        recs[lds] = rec
        plan.close()
    assert float(recs["0"].abs().max()) > 0
    assert torch.equal(recs["0"], recs["1"])
