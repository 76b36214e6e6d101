"""Parity of every drop-in operator (HIP, through the C ABI) with the FP32
oracle on seeded random inputs.  The kernels keep the reference's rounding
order, so most outputs are bit-identical; the exceptions are noted per test
(expf/powf of the GPU math library, float atomics)."""
import numpy as np
import pytest
import torch

from singleventriclesegmentation_amd import opticalFlow as OF
from singleventriclesegmentation_amd.flow_utils import apply_gaussian_blur3d, gaussian_weights3d

pytestmark = pytest.mark.gpu

SHAPES = [(9, 10, 11), (16, 16, 16), (5, 7, 33)]
BCS = [0, 1, 2]
IPS = [0, 1, 2]


def g(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)


def h(t):
    return t.detach().cpu().numpy()


def mesh(shape, L=None):
    NZ, NY, NX = shape
    if L is None:
        L = (NZ - 1, NY - 1, NX - 1)
    return OF.MeshInfo3D(NZ, NY, NX, *L)


def hs(m):
    return (m.gethZ(), m.gethY(), m.gethX())


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("bc", BCS)
def test_nabla_forward_backward(O, dev, shape, bc):
    r = np.random.default_rng(1)
    m = mesh(shape, L=(2.5, 3.0, 7.0))
    op = OF.Nabla3D_CD(m, bc)
    b = r.standard_normal(shape).astype(np.float32)
    v = r.standard_normal(shape + (3,)).astype(np.float32)
    q = r.standard_normal(shape + (3, 3)).astype(np.float32)
    assert np.array_equal(h(op.forward(g(b, dev))), O.nabla_fwd(b, hs(m), bc))
    assert np.array_equal(h(op.forwardVectorField(g(v, dev))), O.nabla_fwd(v, hs(m), bc))
    assert np.array_equal(h(op.backward(g(v, dev))), O.nabla_bwd(v, hs(m), bc))
    assert np.array_equal(h(op.backwardVectorField(g(q, dev))), O.nabla_bwd(q, hs(m), bc))


@pytest.mark.parametrize("shape", SHAPES)
def test_aniso(O, dev, shape):
    r = np.random.default_rng(2)
    m = mesh(shape)
    G = r.standard_normal(shape + (3,)).astype(np.float32) * 0.1
    G[0, 0, 0] = 0.0  # the |grad| == 0 branch
    G[1, 1, 1] = (0.0, 0.0, 0.7)  # n_z > 0.5 branch
    op = OF.AnisotropicNabla3D(m, 5.0, 0.5)
    s, n, t1, t2 = (h(t) for t in op.computeTangentVecs(g(G, dev)))
    so, no, t1o, t2o = O.aniso_tangents(G, 5.0, 0.5)
    # s uses the shared faithful expf/powf (svs_math.h) on both sides: identical bits.
    assert np.array_equal(s, so)
    assert np.array_equal(n, no) and np.array_equal(t1, t1o) and np.array_equal(t2, t2o)
    b = r.standard_normal(shape + (3, 3)).astype(np.float32)
    args = [g(a, dev) for a in (so, no, t1o, t2o)]
    out = h(op.forwardVectorField(g(b, dev), *args))
    assert np.array_equal(out, O.aniso_apply(b, so, no, t1o, t2o))
    assert np.array_equal(h(op.backwardVectorField(g(b, dev), *args)), out)


@pytest.mark.parametrize("shape", SHAPES)
def test_prox(O, dev, shape):
    r = np.random.default_rng(3)
    m = mesh(shape)
    primal = r.standard_normal(shape + (3,)).astype(np.float32)
    gg = r.standard_normal(shape + (3,)).astype(np.float32) * 0.2
    gg[0, 0, 0] = 0.0  # g2 <= 1e-10 branch
    rho = r.standard_normal(shape).astype(np.float32) * 0.05
    out = h(OF.TVL1OF3D_proxPrimal(g(primal, dev), 0.37, 1.0, g(rho, dev), g(gg, dev), m))
    assert np.array_equal(out, O.prox_primal(primal, np.float32(0.37), np.float32(1.0), rho, gg))
    dual = r.standard_normal(shape + (3, 3)).astype(np.float32) * 0.03
    out = h(OF.TVL1OF3D_proxDual(g(dual, dev), 0.5, 0.025, m))
    assert np.array_equal(out, O.prox_dual(dual, np.float32(0.025)))


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("ip", IPS)
def test_warp(O, dev, shape, bc, ip):
    r = np.random.default_rng(4 + bc + 3 * ip)
    m = mesh(shape)
    u = r.standard_normal(shape).astype(np.float32)
    phi = (r.standard_normal(shape + (3,)) * 1.5).astype(np.float32)
    phi[0, 0, 0] = (-1e-9, -3e-8, 0.0)  # w + 1 rounds up to 2 in the local coordinate
    op = OF.Warping3D(m, ip, bc)
    assert np.array_equal(h(op.forward(g(u, dev), g(phi, dev))), O.warp(u, phi, hs(m), ip, bc))
    v = r.standard_normal(shape + (3,)).astype(np.float32)
    assert np.array_equal(h(op.forwardVectorField(g(v, dev), g(phi, dev))), O.warp(v, phi, hs(m), ip, bc))


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("ip", IPS)
def test_warp_with_nan_inputs(O, dev, ip, bc):
    """NaN in the image and in the flow: the same NaN positions and the same
    bits elsewhere as the oracle (a NaN coordinate floors to index 0, as the
    reference's CUDA conversion gives)."""
    r = np.random.default_rng(30 + ip + 3 * bc)
    shape = (9, 10, 11)
    m = mesh(shape)
    u = r.standard_normal(shape).astype(np.float32)
    u[4, 5, 6] = np.nan
    phi = (r.standard_normal(shape + (3,)) * 1.5).astype(np.float32)
    phi[2, 3, 4, :] = np.nan
    phi[6, 1, 8, 1] = np.nan
    op = OF.Warping3D(m, ip, bc)
    assert np.array_equal(h(op.forward(g(u, dev), g(phi, dev))), O.warp(u, phi, hs(m), ip, bc), equal_nan=True)
    cnn = OF.WarpingCNN3D(m, ip, bc)
    got = h(cnn.forward(g(u[None, None], dev), g(phi[None], dev)))[0, 0]
    assert np.array_equal(got, O.warp(u, phi, hs(m), ip, bc), equal_nan=True)


@pytest.mark.parametrize("ip", IPS)
def test_warp_cnn(O, dev, ip):
    r = np.random.default_rng(7)
    shape = (12, 13, 14)
    m = mesh(shape)
    u = r.standard_normal((2, 3) + shape).astype(np.float32)
    phi = (r.standard_normal((2,) + shape + (3,)) * 2).astype(np.float32)
    out = h(OF.WarpingCNN3D(m, ip, 1).forward(g(u, dev), g(phi, dev)))
    assert np.array_equal(out, O.warp_cnn(u, phi, hs(m), ip, 1))


@pytest.mark.parametrize("ip", [1, 2])
@pytest.mark.parametrize("bc", BCS)
def test_warp_backward(O, dev, ip, bc):
    r = np.random.default_rng(8)
    shape = (10, 11, 12)
    m = mesh(shape)
    u = r.standard_normal(shape).astype(np.float32)
    phi = (r.standard_normal(shape + (3,)) * 1.5).astype(np.float32)
    go = r.standard_normal(shape).astype(np.float32)
    gu, gphi = (h(t) for t in OF.Warping3D(m, ip, bc).backward(g(u, dev), g(phi, dev), g(go, dev)))
    guo, gphio = O.warp_backward(u, phi, go, hs(m), ip, bc)
    # grad_phi is accumulated per voxel in the reference's order: identical.
    assert np.array_equal(gphi, gphio)
    # grad_u sums 64 scattered contributions with atomics: order-dependent low bits.
    np.testing.assert_allclose(gu, guo, rtol=0, atol=2e-5 * np.abs(guo).max())


def test_warp_cnn_backward(O, dev):
    r = np.random.default_rng(9)
    shape = (8, 9, 10)
    m = mesh(shape)
    u = r.standard_normal((2, 2) + shape).astype(np.float32)
    phi = (r.standard_normal((2,) + shape + (3,)) * 1.5).astype(np.float32)
    go = r.standard_normal((2, 2) + shape).astype(np.float32)
    gu, gphi = (h(t) for t in OF.WarpingCNN3D(m, 2, 1).backward(g(u, dev), g(phi, dev), g(go, dev)))
    guo, gphio = O.warp_cnn_backward(u, phi, go, hs(m), 2, 1)
    assert np.array_equal(gphi, gphio)
    np.testing.assert_allclose(gu, guo, rtol=0, atol=2e-5 * np.abs(guo).max())


@pytest.mark.parametrize("ip", IPS)
@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("src,dst", [((20, 21, 22), (10, 11, 11)), ((10, 11, 11), (20, 21, 22))])
def test_prolongation(O, dev, ip, bc, src, dst):
    r = np.random.default_rng(10)
    ms, md = mesh(src), mesh(dst)
    Ls = (ms.getLZ(), ms.getLY(), ms.getLX())
    Ld = (md.getLZ(), md.getLY(), md.getLX())
    op = OF.Prolongation3D(ms, md, ip, bc)
    for trailing, fn in (((), op.forward), ((3,), op.forwardVectorField), ((3, 3), op.forwardMatrixField)):
        u = r.standard_normal(src + trailing).astype(np.float32)
        assert np.array_equal(h(fn(g(u, dev))), O.prolongate(u, Ls, dst, Ld, ip, bc))


@pytest.mark.parametrize("sigma", [1.0, 1.5])
def test_blur(O, dev, sigma):
    r = np.random.default_rng(11)
    vol = r.random((17, 18, 19)).astype(np.float32)
    out = h(apply_gaussian_blur3d(g(vol, dev), sigma))
    assert np.array_equal(out, O.blur(vol, gaussian_weights3d(sigma)))


def test_invalid_inputs_raise(dev):
    m = mesh((8, 8, 8))
    op = OF.Nabla3D_CD(m, 1)
    with pytest.raises(RuntimeError):
        op.forward(torch.zeros(8, 8, 8))  # CPU tensor
    with pytest.raises(RuntimeError):
        op.forward(torch.zeros(8, 8, 8, device=dev, dtype=torch.float64))
    with pytest.raises(RuntimeError):
        op.forward(torch.zeros(8, 8, 16, device=dev)[:, :, ::2])  # non-contiguous
    with pytest.raises(RuntimeError):
        op.forwardVectorField(torch.zeros(8, 8, 8, device=dev))  # wrong rank
    with pytest.raises(RuntimeError):
        OF.Warping3D(m, 2, 1).forward(torch.zeros(8, 8, 8, device=dev), torch.zeros(8, 8, 7, 3, device=dev))


def test_gpu_reproduces_golden_ops(dev):
    import os
    g = dict(np.load(os.path.join(os.path.dirname(__file__), "golden", "ops_10x11x12.npz"), allow_pickle=False))
    m = mesh(g["u"].shape)
    assert np.array_equal(h(OF.Warping3D(m, 2, 1).forward(g_(g["u"], dev), g_(g["phi"], dev))), g["warp_cubic_mirror"])
    assert np.array_equal(h(OF.Warping3D(m, 1, 2).forwardVectorField(g_(g["v"], dev), g_(g["phi"], dev))),
                          g["warp_vf_linear_reflect"])
    op = OF.Nabla3D_CD(m, 1)
    assert np.array_equal(h(op.forwardVectorField(g_(g["v"], dev))), g["nabla_fwd_vf_mirror"])
    assert np.array_equal(h(op.backwardVectorField(g_(g["q"], dev))), g["nabla_bwd_vf_mirror"])
    an = OF.AnisotropicNabla3D(m, 5.0, 0.5)
    outs = [h(t) for t in an.computeTangentVecs(g_(g["G"], dev))]
    for a, k in zip(outs, ("aniso_s", "aniso_n", "aniso_t1", "aniso_t2")):
        assert np.array_equal(a, g[k])
    args = [g_(g[k], dev) for k in ("aniso_s", "aniso_n", "aniso_t1", "aniso_t2")]
    assert np.array_equal(h(an.forwardVectorField(g_(g["q"], dev), *args)), g["aniso_apply"])
    assert np.array_equal(h(OF.TVL1OF3D_proxPrimal(g_(g["v"], dev), 0.37, 1.0, g_(g["rho"], dev), g_(g["G"], dev), m)),
                          g["prox_primal"])
    assert np.array_equal(h(OF.TVL1OF3D_proxDual(g_(g["q"] * np.float32(0.03), dev), 0.5, 0.025, m)), g["prox_dual"])
    pm = OF.Prolongation3D(OF.MeshInfo3D(10, 11, 12, 9, 10, 11), OF.MeshInfo3D(5, 6, 6, 4, 5, 5), 2, 1)
    assert np.array_equal(h(pm.forwardMatrixField(g_(g["q"], dev))), g["prolong_mf"])


def g_(a, dev):
    return g(a, dev)


@pytest.mark.parametrize("src,dst,T,pad", [((9, 40, 36), (20, 24, 22), 7, 10), ((12, 160, 144), (80, 80, 80), 25, 35)])
def test_preprocessing_resize_matches_oracle(O, dev, src, dst, T, pad):
    """svs/preprocessing/prolongation.py:69-85 with conf/preprocessing.yaml (LINEAR,
    MIRROR, numDofs): the 4D image through forwardVectorField, the masks through
    forward + round, zero frames up to pad_time."""
    from singleventriclesegmentation_amd.preprocessing import prolongate_patient
    r = np.random.default_rng(sum(src) + T)
    img = r.random(src + (T,)).astype(np.float32)
    m_dia = (r.random(src) > 0.6).astype(np.float32)
    m_sys = (r.random(src) > 0.4).astype(np.float32)
    out, a, b = prolongate_patient(g(img, dev), g(m_dia, dev), g(m_sys, dev), shape_zyx=dst, pad_time=pad)
    Ls = tuple(n - 1 for n in src)
    Ld = tuple(n - 1 for n in dst)
    assert tuple(out.shape) == dst + (pad,)
    assert np.array_equal(h(out)[..., :T], O.prolongate(img, Ls, dst, Ld, ip=1, bc=1))
    assert not h(out)[..., T:].any()
    assert np.array_equal(h(a), np.round(O.prolongate(m_dia, Ls, dst, Ld, ip=1, bc=1)))
    assert np.array_equal(h(b), np.round(O.prolongate(m_sys, Ls, dst, Ld, ip=1, bc=1)))
