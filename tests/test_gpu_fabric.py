"""GPU parity of the fabric path (getFSE, getDEM, anisotropic_tensor) through the C ABI against the
FP64 oracle, on the reference's DEM tables (tests/golden/dem_tables.npz)."""
from pathlib import Path

import numpy as np
import pytest

from dimspers_amd import _lib
from dimspers_amd.inputs import make_inputs, scaled_workload
from oracle import fabric_oracle as fo

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).parent / "golden" / "dem_tables.npz"
W_CUTOFF = 0.03
SHAPE_NODES = np.geomspace(1.0, 60.0, 14)
SHAPE_R1 = 1.05 * np.log10(SHAPE_NODES) - 0.02
SHAPE_R2 = 2.1 * np.exp(-(SHAPE_NODES - 1.0) / 6.0) - 0.05


def _dem(name="DEM_1em3_vol30_"):
    d = np.load(GOLDEN)
    return d[name + "C"], d[name + "ax"]


def _ctx(nr, nt, w=3):
    inp = make_inputs(scaled_workload(w, nr, nt))
    print(f"[case] fabric mesh {nr}x{nt}", flush=True)
    return inp, _lib.Context.from_mesh(inp.mesh)


def _random_F(rng, nel, nip):
    """Deformation gradients with every branch: generic, strongly stretched (a1/a2 > 1e3),
    huge (normalised), diagonal and identity points."""
    F = rng.normal(size=(nel, nip, 2, 2)) + 1.5 * np.eye(2)
    flat = F.reshape(-1, 2, 2)
    n = flat.shape[0]
    idx = rng.permutation(n)
    k = n // 10
    a = rng.uniform(0, 2 * np.pi, k)
    c, s = np.cos(a), np.sin(a)
    R = np.stack([np.stack([c, -s], -1), np.stack([s, c], -1)], -2)
    D = np.zeros((k, 2, 2))
    D[:, 0, 0], D[:, 1, 1] = 60.0, 1.0 / 60.0              # a1/a2 = 3600
    flat[idx[:k]] = R @ D
    flat[idx[k:2 * k]] *= 1e40
    flat[idx[2 * k:2 * k + 5]] = np.eye(2)
    flat[idx[2 * k + 5:2 * k + 10]] = np.diag([2.0, 0.5])
    return F


def _points(F):
    """nel x nip x 2 x 2 -> points column-major (q outer, e inner) as the oracle takes them."""
    return np.transpose(F, (1, 0, 2, 3)).reshape(-1, 2, 2)


def _unpoints(P, nel, nip, last):
    return np.transpose(P.reshape((nip, nel) + last), (1, 0) + tuple(range(2, 2 + len(last))))


def _close(got, want, tol):
    scale = np.maximum(np.abs(want), 1.0)
    return np.max(np.abs(got - want) / scale) <= tol


@pytest.mark.parametrize("mode", [_lib.FSE_NO_HEALING, _lib.FSE_FABRIC_DESTRUCTION, _lib.FSE_ANNEALING_DESTRUCTION])
@pytest.mark.parametrize("nip", [6, 7])
def test_fse_matches_oracle(mode, nip):
    inp, ctx = _ctx(8, 64)
    nel = inp.mesh.nel
    rng = np.random.default_rng(10 + mode)
    F = _random_F(rng, nel, nip)
    got, Fg = ctx.fse(F, mode=mode, eps=1e3)
    want, Fw = fo.get_fse(_points(F), mode, eps=1e3)
    want = _unpoints(want, nel, nip, (6,))
    Fw = _unpoints(Fw, nel, nip, (2, 2))
    # axes absolute (unit vectors), semi-axes relative
    err = max(np.max(np.abs(got[..., :4] - want[..., :4])),
              np.max(np.abs(got[..., 4:] - want[..., 4:]) / np.maximum(np.abs(want[..., 4:]), 1e-300)))
    print(f"[result] fse mode {mode} nip {nip}: max rel err {err:.2e}, destroyed "
          f"{int(np.sum(np.all(want == fo.ISOTROPIC_FSE, axis=2)))}")
    assert err < 1e-13
    assert np.array_equal(Fg, Fw) if mode != _lib.FSE_ANNEALING_DESTRUCTION else _close(Fg, Fw, 1e-15)
    ctx.close()


@pytest.mark.parametrize("shape", ["table", "points"])
@pytest.mark.parametrize("nip", [6, 7])
def test_anisotropic_tensor_matches_oracle(shape, nip):
    inp, ctx = _ctx(32, 256, 4)
    nel = inp.mesh.nel
    T, ax = _dem()
    dem = fo.get_dem(T, ax)
    ctx.set_dem(T, ax, W_CUTOFF)
    rng = np.random.default_rng(20 + nip)
    F = _random_F(rng, nel, nip)
    fse, _ = ctx.fse(F, mode=_lib.FSE_FABRIC_DESTRUCTION, eps=1e3)      # stays on the device
    pts = np.transpose(fse, (1, 0, 2)).reshape(-1, 6)
    if shape == "table":
        got = ctx.anisotropic_tensor(nip=nip, shape_table=(SHAPE_NODES, SHAPE_R1, SHAPE_R2))
        r1, r2 = fo.shape_from_table(pts[:, 4], SHAPE_NODES, SHAPE_R1, SHAPE_R2)
    else:
        r1p, r2p = rng.uniform(-0.2, 1.9, (nel, nip)), rng.uniform(-0.2, 2.2, (nel, nip))
        got = ctx.anisotropic_tensor(fse=fse, shape_points=(r1p, r2p))
        r1, r2 = r1p.ravel(order="F"), r2p.ravel(order="F")
    want = fo.anisotropic_tensor(pts, nel, dem, W_CUTOFF, r1, r2)
    rows = fo.dem_index(dem, r1, r2)
    errs = [np.max(np.abs(g - w)) / np.abs(w).max() for g, w in zip(got, want)]
    print(f"[result] anisotropic tensor {shape} nip {nip}: max rel err {max(errs):.2e}, "
          f"{np.unique(rows).size} table entries used")
    assert max(errs) < 1e-14
    assert np.unique(rows).size >= (8 if shape == "table" else 30)
    ctx.close()


def test_install_into_assembly_and_stress():
    """The installed tensor assembles and post-processes exactly as the same arrays given by the host."""
    inp, ctx = _ctx(8, 64)
    m = inp.mesh
    T, ax = _dem("DEM_1em3_vol20_")
    ctx.set_dem(T, ax, W_CUTOFF)
    rng = np.random.default_rng(5)
    fse, _ = ctx.fse(_random_F(rng, m.nel, 7), mode=_lib.FSE_FABRIC_DESTRUCTION)
    ctx.set_fields(inp.g, inp.rho, inp.eta_ip, None)                     # isotropic until installed
    eta7 = np.asfortranarray(np.exp(rng.normal(size=(m.nel, 7))))
    ctx.set_stress_fields(eta7, None, nip=7)
    got = ctx.anisotropic_tensor(shape_table=(SHAPE_NODES, SHAPE_R1, SHAPE_R2),
                                 install=_lib.INSTALL_ASSEMBLY | _lib.INSTALL_STRESS)
    ctx.assemble(_lib.ASM_FP64)
    K_inst = ctx.export_csc(_lib.MAT_K)
    U = rng.normal(size=m.nU)
    s_inst = ctx.stress(U)
    ctx.close()
    ctx2 = _lib.Context.from_mesh(m)
    ctx2.set_fields(inp.g, inp.rho, inp.eta_ip, got)
    ctx2.assemble(_lib.ASM_FP64)
    K_host = ctx2.export_csc(_lib.MAT_K)
    ctx2.set_stress_fields(eta7, got, nip=7)
    s_host = ctx2.stress(U)
    ctx2.close()
    assert K_inst.nnz == K_host.nnz
    assert abs(K_inst - K_host).max() <= 1e-15 * abs(K_host).max()
    for k in _lib.STRESS_FIELDS:
        assert np.max(np.abs(s_inst[k] - s_host[k])) <= 1e-15 * np.abs(s_host[k]).max(), k


def test_fabric_error_paths():
    inp, ctx = _ctx(4, 16)
    nel = inp.mesh.nel
    T, ax = _dem()
    shp = (SHAPE_NODES, SHAPE_R1, SHAPE_R2)
    with pytest.raises(_lib.StkError) as e:
        ctx.anisotropic_tensor(nip=6, shape_table=shp)                   # no DEM table
    assert e.value.code == _lib.STK_E_STATE
    ctx.set_dem(T, ax, W_CUTOFF)
    with pytest.raises(_lib.StkError) as e:
        ctx.anisotropic_tensor(nip=6, shape_table=shp)                   # no device ellipses
    assert e.value.code == _lib.STK_E_STATE
    ctx.fse(np.tile(np.eye(2), (nel, 6, 1, 1)), mode=_lib.FSE_NO_HEALING)
    with pytest.raises(_lib.StkError) as e:
        ctx.anisotropic_tensor(nip=7, shape_table=shp)                   # ellipses have 6 points
    assert e.value.code == _lib.STK_E_STATE
    with pytest.raises(_lib.StkError) as e:
        ctx.anisotropic_tensor(nip=6, shape_table=(SHAPE_NODES[::-1], SHAPE_R1, SHAPE_R2))
    assert e.value.code == _lib.STK_E_ARG
    with pytest.raises(_lib.StkError) as e:
        ctx.anisotropic_tensor(nip=6, shape_table=shp, install=_lib.INSTALL_ASSEMBLY)   # no fields
    assert e.value.code == _lib.STK_E_STATE
    with pytest.raises(_lib.StkError) as e:
        ctx.fse(np.tile(np.eye(2), (nel, 6, 1, 1)), mode=_lib.FSE_FABRIC_DESTRUCTION, eps=0.0)
    assert e.value.code == _lib.STK_E_ARG
    with pytest.raises(_lib.StkError) as e:
        ctx.set_dem(T[:10], ax[:10], W_CUTOFF)                           # one block of equal a1
    assert e.value.code == _lib.STK_E_SHAPE
    bad = ax.copy()
    bad[3, 2] = 0.0
    with pytest.raises(_lib.StkError) as e:
        ctx.set_dem(T, bad, W_CUTOFF)                                    # log10 of a zero semi-axis
    assert e.value.code == _lib.STK_E_ARG
    # identity F everywhere: unit ellipses on StaticArrays' diagonal branch
    ell, _ = ctx.fse(np.tile(np.eye(2), (nel, 6, 1, 1)), mode=_lib.FSE_NO_HEALING)
    assert np.array_equal(ell.reshape(-1, 6), np.tile([0.0, 1.0, 1.0, 0.0, 1.0, 1.0], (nel * 6, 1)))
    ctx.close()


def test_fabric_at_full_size():
    """256 x 2048 (7.3M points): ellipses of F = R(alpha) diag(s, 1/s) R(beta)^T hold a1 = s and the
    axis alpha at every point; the tensor of a sample of 4096 elements matches the oracle.  The
    closed-form eigenvector of an axis at angle d from the vertical carries an error of about
    eps / d, at most about sqrt(eps), hence 1e-7 on the directions of 7.3M random axes."""
    inp, ctx = _ctx(256, 2048, 4)
    nel = inp.mesh.nel
    rng = np.random.default_rng(7)
    alpha = rng.uniform(0, 2 * np.pi, (nel, 7))
    beta = rng.uniform(0, 2 * np.pi, (nel, 7))
    s = np.exp(rng.uniform(0.05, 3.0, (nel, 7)))
    ca, sa, cb, sb = np.cos(alpha), np.sin(alpha), np.cos(beta), np.sin(beta)
    F = np.empty((nel, 7, 2, 2))
    # R(alpha) diag(s, 1/s) R(beta)^T written out
    F[..., 0, 0] = ca * s * cb + sa / s * sb
    F[..., 0, 1] = ca * s * sb - sa / s * cb
    F[..., 1, 0] = sa * s * cb - ca / s * sb
    F[..., 1, 1] = sa * s * sb + ca / s * cb
    del ca, sa, cb, sb
    ell, _ = ctx.fse(F, mode=_lib.FSE_FABRIC_DESTRUCTION, eps=1e3)
    assert np.max(np.abs(ell[..., 4] / s - 1)) < 1e-12
    assert np.max(np.abs(ell[..., 0] * np.sin(alpha) - ell[..., 2] * np.cos(alpha))) < 1e-7
    T, ax = _dem()
    ctx.set_dem(T, ax, W_CUTOFF)
    got = ctx.anisotropic_tensor(nip=7, shape_table=(SHAPE_NODES, SHAPE_R1, SHAPE_R2))
    sample = np.sort(rng.choice(nel, 4096, replace=False))
    pts = np.transpose(ell[sample, :6], (1, 0, 2)).reshape(-1, 6)
    r1, r2 = fo.shape_from_table(pts[:, 4], SHAPE_NODES, SHAPE_R1, SHAPE_R2)
    want = fo.anisotropic_tensor(pts, sample.size, fo.get_dem(T, ax), W_CUTOFF, r1, r2)
    err = max(np.max(np.abs(g[sample] - w)) / np.abs(w).max() for g, w in zip(got, want))
    print(f"[result] fabric 256x2048: a1 and axis closed form ok, tensor sample max rel err {err:.2e}")
    assert err < 1e-14
    ctx.close()


def test_reference_style_chain():
    """getFSE -> anisotropic_tensor with the reference's names (dimspers_amd.fabric) gives the
    oracle's tensor; FSE = None reads the device ellipses."""
    from dimspers_amd import fabric
    inp, ctx = _ctx(8, 64)
    nel = inp.mesh.nel
    T, ax = _dem("DEM_1em3_vol10_")
    D = fabric.getDEM(T, ax, 0.1, W_CUTOFF, SHAPE_NODES, SHAPE_R1, SHAPE_R2)
    FSE = fabric.FiniteStrain.new(nel, nip=7)
    rng = np.random.default_rng(8)
    F = _random_F(rng, nel, 7)
    FSE, F1 = fabric.getFSE(ctx, F, FSE)
    C = fabric.anisotropic_tensor(ctx, None, D)
    pts = np.transpose(FSE.fse, (1, 0, 2)).reshape(-1, 6)
    r1, r2 = fo.shape_from_table(pts[:, 4], SHAPE_NODES, SHAPE_R1, SHAPE_R2)
    want = fo.anisotropic_tensor(pts, nel, fo.get_dem(T, ax), W_CUTOFF, r1, r2)
    for g, w in zip(C.components(), want):
        assert np.max(np.abs(g - w)) <= 1e-14 * np.abs(w).max()
    _, Fw = fo.get_fse(_points(F), fo.FABRIC_DESTRUCTION, eps=1e3)
    assert np.array_equal(F1, _unpoints(Fw, nel, 7, (2, 2)))
    ctx.close()


def test_solve_with_the_installed_dem_tensor():
    """End to end at 64 x 512: simple-shear F -> ellipses -> DEM tensor installed on the device ->
    assembly -> mixed solve to the 1e-8 full-system residual, against the plain FP64 solve."""
    inp, ctx = _ctx(64, 512, 4)
    m = inp.mesh
    T, ax = _dem()
    ctx.set_dem(T, ax, W_CUTOFF)
    # F = R(phi) [[1, g], [0, 1]] R(phi)^T with the shear g and direction phi varying over the mesh
    thc = np.asarray(m.theta)[np.asarray(m.e2n)[:3]].mean(axis=0)
    g = 2.0 + 2.0 * np.sin(3 * thc)[:, None] * np.ones((1, 7))
    phi = (thc[:, None] + 0.3 * np.arange(7)[None, :])
    c, s = np.cos(phi), np.sin(phi)
    F = np.empty((m.nel, 7, 2, 2))
    F[..., 0, 0] = 1 - g * c * s
    F[..., 0, 1] = g * c * c
    F[..., 1, 0] = -g * s * s
    F[..., 1, 1] = 1 + g * c * s
    ctx.fse(F, mode=_lib.FSE_FABRIC_DESTRUCTION)
    ctx.set_fields(inp.g, inp.rho, inp.eta_ip, None)
    C = ctx.anisotropic_tensor(shape_table=(SHAPE_NODES, SHAPE_R1, SHAPE_R2), install=_lib.INSTALL_ASSEMBLY)
    assert np.ptp(C[2]) > 0.1 and np.abs(C[4]).max() > 0.01      # anisotropic, rotated
    ctx.assemble(_lib.ASM_FP64 | _lib.ASM_FP32)
    rep = ctx.solve(_lib.default_opts(method=_lib.SOLVE_MIXED, stop=_lib.STOP_FULL, rtol=1e-8))
    U1, P1 = ctx.solution()
    rep0 = ctx.solve(_lib.default_opts(method=_lib.SOLVE_PLAIN, stop=_lib.STOP_FULL, rtol=1e-8))
    U0, P0 = ctx.solution()
    eu = np.linalg.norm(U1 - U0) / np.linalg.norm(U0)
    ep = np.linalg.norm((P1 - P1.mean()) - (P0 - P0.mean())) / np.linalg.norm(P0 - P0.mean())
    print(f"[result] DEM tensor solve 64x512: mixed {rep.solve_ms:.1f} ms {rep.outer_iterations} outer, "
          f"plain {rep0.solve_ms:.1f} ms {rep0.outer_iterations} outer; errU {eu:.1e} errP {ep:.1e}")
    assert rep.converged and rep0.converged
    assert rep.full_residual < 1e-8 and rep0.full_residual < 1e-8
    assert eu < 1e-6 and ep < 1e-6
    ctx.close()
