"""Dual-energy elemental decomposition (several problems optimised together by Adam on the adjoint kernel):
the first gradient equals the FP64 oracle adjoint summed over the two spectra, problems optimised together
follow the same trajectory as each optimised alone, and the loss falls."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as T
    if not T.cuda.is_available():
        pytest.skip("no GPU")
    return T


def _geom(Nz):
    ang = np.deg2rad(np.array([-170.0001, -110.0001, -50.00001, 9.99999, 69.99999, 130.0]))
    cs, sn, xc, yc = O.trig_table_proj(18, 18, ang, np.full(6, 0.02), np.zeros(6))
    return O.Geometry(N=20, s=0.7, Nz=Nz, t=0.9, H=18, W=18, cos_t=cs, sin_t=sn, xc=xc, yc=yc)


def _spectra(rng):
    out = []
    for NE in (6, 9):  # two spectra of different bin counts (the decomposition's low and high kV)
        out.append((rng.uniform(0.1, 0.6, (6, NE)).astype(np.float32), rng.uniform(0.2, 1.0, NE).astype(np.float32)))
    return out


def _problem(torch, Nz, rng, spectra):
    from virtualcbct_amd.projector import Projector
    g = _geom(Nz)
    truth = rng.uniform(0.05, 1.0, (6, Nz, g.N, g.N)).astype(np.float32)
    projs, targets = [], []
    for atten, pdf in spectra:
        P = Projector(g.N, g.s, g.Nz, g.t, g.H, g.W, g.n_ang, g.cos_t, g.sin_t, g.xc, g.yc, g.x_reg, g.y_reg,
                      g.z_reg, g.det_pitch, g.det_pitch_iso, g.sad, g.det_pixel_num, g.nmat, 0)
        P.set_spectrum(atten, pdf)
        t = O.logproj(g, np.ascontiguousarray(truth.transpose(1, 0, 2, 3)), atten, pdf, 1.0)
        projs.append(P)
        targets.append(torch.from_numpy(t.astype(np.float32)).cuda())
    return {"projectors": projs, "targets": targets, "shape": truth.shape, "xfactor": 1.0, "g": g, "truth": truth}


def test_decompose_gradient_and_batching(torch):
    from virtualcbct_amd.decompose import decompose
    from virtualcbct_amd.autograd import poly_logproj
    rng = np.random.default_rng(21)
    spectra = _spectra(rng)
    probs = [_problem(torch, Nz, rng, spectra) for Nz in (3, 5)]
    init = [torch.full(p["shape"], 0.4, device="cuda") for p in probs]

    # first gradient of problem 1 vs the FP64 oracle adjoint summed over the spectra
    pr = probs[1]
    m = init[1].clone().requires_grad_(True)
    for P, t in zip(pr["projectors"], pr["targets"]):
        ((poly_logproj(m, P, 1.0) - t) ** 2).sum().backward()
    g = pr["g"]
    x0 = np.ascontiguousarray(init[1].cpu().numpy().transpose(1, 0, 2, 3))
    go = sum(O.adjoint(g, x0, a, p, 1.0, t.cpu().numpy().astype(np.float64))[0] for (a, p), t in
             zip(spectra, pr["targets"])).transpose(1, 0, 2, 3)
    err = np.abs(m.grad.cpu().numpy() - go).max() / np.abs(go).max()
    print(f"decomposition gradient vs FP64 oracle: {err:.3e}")
    assert err < 1e-4

    fit_b, hist_b = decompose(probs, iters=6, lr=0.02, init=init)
    assert hist_b[-1] < 0.8 * hist_b[0], hist_b
    # optimising the problems together follows the same trajectory as optimising each alone (up to the
    # summation order of the FP32 atomic scatter; cross-talk between problems would move whole maps)
    for i, pr in enumerate(probs):
        fit_1, _ = decompose([pr], iters=6, lr=0.02, init=[init[i]])
        d = (fit_1[0] - fit_b[i]).abs()
        print(f"problem {i}: together vs alone max |diff| {d.max().item():.2e}")
        assert d.max().item() < 1e-3, (i, d.max().item())
        assert torch.quantile(d.flatten(), 0.99).item() < 1e-5
    for pr in probs:
        for P in pr["projectors"]:
            P.close()


def test_multi_spectrum_loss_grad_matches_oracle(torch, tmp_path):
    """One traversal and one scatter for both spectra (vcbct_projector_multi_loss_grad) against the FP64 oracle
    adjoint summed over the spectra, for generic tables (six channels) and for the workloads' photoelectric +
    Compton tables at 80 and 120 kV (a joint basis of two channels)."""
    from virtualcbct_amd import phantoms
    rng = np.random.default_rng(41)
    g = _geom(5)
    truth = rng.uniform(0.05, 1.0, (6, 5, g.N, g.N)).astype(np.float32)
    x = rng.uniform(0.05, 1.0, truth.shape).astype(np.float32)
    xo = np.ascontiguousarray(x.transpose(1, 0, 2, 3))
    for spectra, want in ((_spectra(rng), 6),
                          ([phantoms.synthetic_spectrum(str(tmp_path), kv) for kv in (80, 120)], 2)):
        from virtualcbct_amd.projector import Projector
        P = Projector(g.N, g.s, g.Nz, g.t, g.H, g.W, g.n_ang, g.cos_t, g.sin_t, g.xc, g.yc, g.x_reg, g.y_reg,
                      g.z_reg, g.det_pitch, g.det_pitch_iso, g.sad, g.det_pixel_num, g.nmat, 0)
        P.set_spectra(spectra)
        assert P.multi_channels == want
        tg = [O.logproj(g, np.ascontiguousarray(truth.transpose(1, 0, 2, 3)), a, p, 1.0) for a, p in spectra]
        T = torch.from_numpy(np.stack(tg).astype(np.float32)).cuda()
        P.set_volume(torch.from_numpy(x).cuda())
        grad, loss = P.multi_loss_grad(T, 1.0)
        go, lo = 0.0, 0.0
        for (a, p), t in zip(spectra, tg):
            gs, ls = O.adjoint(g, xo, a, p, 1.0, t.astype(np.float32).astype(np.float64))
            go, lo = go + gs, lo + ls
        go = go.transpose(1, 0, 2, 3)
        err = float(np.abs(grad.cpu().numpy() - go).max() / np.abs(go).max())
        print(f"{want} channels: multi-spectrum gradient vs FP64 oracle {err:.3e}; loss {float(loss.item()):.6e} "
              f"vs {lo:.6e}")
        assert err < 1e-4
        assert float(loss.item()) == pytest.approx(lo, rel=1e-5)
        P.close()


def test_device_adam_matches_torch_adam(torch):
    """vcbct_adam_step (step count on the device, non-negative clamp) against torch.optim.Adam + clamp_."""
    from virtualcbct_amd._lib import check, lib
    import ctypes as C
    rng = np.random.default_rng(43)
    x0 = rng.uniform(0.0, 1.0, 10007).astype(np.float32)
    xa = torch.from_numpy(x0).cuda()
    xb = xa.clone().requires_grad_(True)
    m, v = torch.zeros_like(xa), torch.zeros_like(xa)
    t = torch.zeros(1, dtype=torch.int32, device="cuda")
    opt = torch.optim.Adam([xb], lr=0.03)
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for k in range(5):
        gr = torch.from_numpy(rng.normal(0, 1, x0.shape).astype(np.float32)).cuda()
        check(lib().vcbct_adam_step(xa.numel(), C.c_void_p(xa.data_ptr()), C.c_void_p(gr.data_ptr()),
                                    C.c_void_p(m.data_ptr()), C.c_void_p(v.data_ptr()), C.c_void_p(t.data_ptr()),
                                    0.03, 0.9, 0.999, 1e-8, 0.0, st))
        xb.grad = gr.clone()
        opt.step()
        with torch.no_grad():
            xb.clamp_(min=0.0)
        d = float((xa - xb.detach()).abs().max())
        assert d < 1e-6, (k, d)
    assert int(t.item()) == 5


def test_fused_decomposition_graph_replays_and_checkpoint(torch, tmp_path):
    """The fused decomposition replayed from its hipGraph equals the same iterations run eagerly, replay after
    replay (with the program's own eager work and synchronisations between them); a run checkpointed and
    resumed ends where the uninterrupted run ends; and the fused iterations follow the autograd path's."""
    from virtualcbct_amd.decompose import FusedDecomposition, decompose, decompose_autograd
    rng = np.random.default_rng(47)
    spectra = _spectra(rng)
    probs = [_problem(torch, Nz, rng, spectra) for Nz in (3, 5)]
    init = [torch.full(p["shape"], 0.4, device="cuda") for p in probs]
    eager = FusedDecomposition(probs, lr=0.02, init=init, graph=False)
    graph = FusedDecomposition(probs, lr=0.02, init=init, graph=True)
    for k in range(5):
        eager.step()
        junk = torch.randn(1000, device="cuda").sum()  # eager work between replays
        graph.step()
        torch.cuda.synchronize()
        for a, b in zip(eager.maps(), graph.maps()):
            d = float((a - b).abs().max())
            assert d < 1e-5, (k, d)
        del junk
    assert graph.graph is not None
    he, hg = eager.losses(), graph.losses()
    np.testing.assert_allclose(hg, he, rtol=1e-6)
    assert he[-1] < he[0]
    # checkpoint / resume
    ck = str(tmp_path / "ck.npz")
    full, hf = decompose(probs, iters=6, lr=0.02, init=init)
    _, h1 = decompose(probs, iters=3, lr=0.02, init=init, checkpoint=ck)
    part, h2 = decompose(probs, iters=6, lr=0.02, init=init, checkpoint=ck, resume=True)
    assert len(h2) == 6 and np.allclose(h2[:3], h1, rtol=0)
    for a, b in zip(full, part):
        assert float((a - b).abs().max()) < 1e-5
    np.testing.assert_allclose(h2, hf, rtol=1e-6)
    # the fused path (device Adam) follows the autograd path (torch Adam)
    ref, hr = decompose_autograd(probs, iters=6, lr=0.02, init=init)
    for a, b in zip(full, ref):
        assert float((a - b).abs().max()) < 1e-4
    np.testing.assert_allclose(hf, hr, rtol=1e-5)
    for p in probs:
        for P in p["projectors"]:
            P.close()
