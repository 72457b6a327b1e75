"""GPU parity of the SH colour head (pf_sh_fwd / pf_sh_bwd, through
`color_fn.SHColor`) against the FP64 evaluation `synth.sh_eval`, whose basis
is pinned against scipy's real spherical harmonics in test_oracle_encoding.py.

Site counts straddle the kernels' 128-site workgroups (1, 129, 300, 1000) and
texel counts cover 1..8. Forward: within 2e-5 abs (16 FP32 products of O(1)).
Backward: grad_sh[s, k, c] = sum_t Y_k(d_t) g[s, t, c] over texels whose
colour channel is positive (Y_0 = 1), within 1e-5 abs per entry; texels whose
pre-clamp colour is within 1e-5 of zero are given zero incoming gradient, so
an FP32-vs-FP64 disagreement on the clamp cannot decide the comparison.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [(1, 1), (129, 3), (300, 5), (1000, 8)]


@pytest.mark.parametrize("n,T", CASES)
def test_sh_head_matches_fp64(n, T):
    import torch
    from powerfoam_amd import synth
    from powerfoam_amd.camera import look_at
    from powerfoam_amd.color_fn import SHColor
    rng = np.random.default_rng(100 * n + T)
    sh = rng.normal(0.0, 0.3, (n, 16, 3)).astype(np.float32)
    sh[:, 0] += 0.3
    ts = rng.normal(0.0, 1.0, (n, T, 3)).astype(np.float32)
    cam = look_at((0.3, -0.2, 4.0), (0.0, 0.0, 0.0), 64, 48, 60.0)
    eye = cam.eye.detach().cpu().numpy().astype(np.float32)
    d = (ts - eye[None, None]).astype(np.float64)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    y = np.concatenate([np.ones((n, T, 1)), synth.sh_basis(d)], -1)       # [n,T,16]
    lin = np.einsum("ntk,nkc->ntc", y, sh.astype(np.float64))
    want = np.maximum(lin, 0.0)

    print(f"sh head n={n} T={T}", flush=True)
    sh_t = torch.tensor(sh, device="cuda", requires_grad=True)
    rgb = SHColor().forward(torch.tensor(ts, device="cuda"), cam, sh_t)
    got = rgb.detach().cpu().numpy()
    assert got.shape == (n, T, 3)
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-5)
    np.testing.assert_allclose(want, synth.sh_eval(sh.astype(np.float64)[:, None], d), rtol=0, atol=1e-12)

    g = rng.normal(size=(n, T, 3))
    g[np.abs(lin) < 1e-5] = 0.0
    (rgb * torch.tensor(g, dtype=torch.float32, device="cuda")).sum().backward()
    want_g = np.einsum("ntk,ntc->nkc", y, g * (lin > 0.0))
    np.testing.assert_allclose(sh_t.grad.cpu().numpy(), want_g, rtol=0, atol=1e-5)


def test_sh_head_unaligned_arrays_match_aligned():
    """Arrays not 16-byte aligned take the kernels' 4-byte copy path: same
    colours and coefficient gradients as aligned copies of the same data."""
    import torch
    from powerfoam_amd.camera import look_at
    from powerfoam_amd.color_fn import SHColor
    rng = np.random.default_rng(7)
    n, T = 300, 8
    sh = rng.normal(0.0, 0.3, (n, 16, 3)).astype(np.float32)
    ts = rng.normal(0.0, 1.0, (n, T, 3)).astype(np.float32)
    cam = look_at((0.3, -0.2, 4.0), (0.0, 0.0, 0.0), 64, 48, 60.0)
    g = torch.tensor(rng.normal(size=(n, T, 3)).astype(np.float32), device="cuda")
    outs = []
    for off in (0, 1):
        flat_sh = torch.zeros(n * 48 + off, device="cuda")
        flat_sh[off:] = torch.tensor(sh.ravel(), device="cuda")
        flat_ts = torch.zeros(n * T * 3 + off, device="cuda")
        flat_ts[off:] = torch.tensor(ts.ravel(), device="cuda")
        sh_t = flat_sh[off:].view(n, 16, 3).detach().requires_grad_(True)
        ts_t = flat_ts[off:].view(n, T, 3)
        assert (ts_t.data_ptr() % 16 != 0) == (off == 1)
        print(f"sh head offset {off}", flush=True)
        rgb = SHColor().forward(ts_t, cam, sh_t)
        (rgb * g).sum().backward()
        outs.append((rgb.detach().cpu().numpy(), sh_t.grad.cpu().numpy()))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])
