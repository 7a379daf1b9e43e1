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
