"""GPU finite-volume assembly against the assembly oracle: the device mesh
equals the numpy layout bit for bit, metrics and the assembled system
(diag, off-diagonals, b, boundary flags) are bit-identical to the oracle's
sequential face loops, and the AMG on the device-assembled system reproduces
the oracle's hierarchy and solution exactly."""
import numpy as np
import pytest

import hexmesh
import oracle_lib as O

pytestmark = pytest.mark.gpu

D, NM = 0, 1
BC_T = [0, D, D, NM, NM, NM, NM]
BC_V = [0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0]


def _dev_fields(mesh, sigma, block, seed):
    import torch
    nt = mesh.n_total
    k = torch.empty(nt, dtype=torch.float64, device="cuda:0")
    s = torch.empty(nt, dtype=torch.float64, device="cuda:0")
    x = torch.zeros(nt, dtype=torch.float64, device="cuda:0")
    mesh.lognormal_field(k, sigma, block, seed)
    mesh.uniform_field(s, 0.0, 1.0, seed + 1)
    return k, s, x


@pytest.mark.parametrize("N,jit,seed", [(4, 0.0, 0), (6, 0.2, 7)])
def test_device_mesh_matches_layout(N, jit, seed):
    from fvm_amd.thermal import Mesh
    m = Mesh.hex(N, jit, seed)
    a = m.arrays()
    ref = hexmesh.hex_mesh(N, jit, seed)
    assert np.array_equal(a["nodes"], ref["nodes"])
    assert np.array_equal(a["face_nodes"], ref["face_nodes"])
    assert np.array_equal(a["face_cells"], ref["face_cells"])
    assert np.array_equal(a["group_ptr"], ref["group_ptr"])


@pytest.mark.parametrize("N,jit,seed", [(4, 0.0, 0), (7, 0.2, 3)])
def test_metrics_bit_identical(N, jit, seed):
    from fvm_amd.thermal import Mesh
    m = Mesh.hex(N, jit, seed)
    g = m.metrics()
    o = O.fv_metrics(hexmesh.hex_mesh(N, jit, seed))
    for key in ("fA", "fAm", "fc", "cc", "vol"):
        assert np.array_equal(g[key], o[key]), key


@pytest.mark.parametrize("N,jit,sigma,block,seed", [(5, 0.2, 1.5, 1, 11), (8, 0.2, 2.5, 4, 12)])
def test_thermal_system_bit_identical(N, jit, sigma, block, seed):
    from fvm_amd.thermal import Mesh, assemble
    m = Mesh.hex(N, jit, seed)
    k, s, x = _dev_fields(m, sigma, block, seed)
    S = assemble(m, k, s, x, BC_T, BC_V).to_host()
    o = O.fv_thermal(hexmesh.hex_mesh(N, jit, seed), k.cpu().numpy(), s.cpu().numpy(), x.cpu().numpy(),
                     BC_T, BC_V)
    for key in ("row", "col", "diag", "off", "b", "is_boundary"):
        assert np.array_equal(S[key], o[key]), key
    # lognormal blocks: constant inside each block
    kk = k.cpu().numpy()[: N ** 3].reshape(N, N, N)
    if block > 1:
        assert np.all(kk[:block, :block, :block] == kk[0, 0, 0])


def test_mesh_from_host_arrays_gives_same_system():
    from fvm_amd.thermal import Mesh, assemble
    N, jit, seed = 5, 0.15, 21
    ref = hexmesh.hex_mesh(N, jit, seed)
    m1 = Mesh.hex(N, jit, seed)
    m2 = Mesh.from_arrays(ref["nodes"], ref["face_nodes"], ref["face_cells"], ref["n_cells"], ref["n_ghost"],
                          ref["group_ptr"])
    k, s, x = _dev_fields(m1, 1.0, 1, seed)
    a = assemble(m1, k, s, x, BC_T, BC_V).to_host()
    b = assemble(m2, k, s, x, BC_T, BC_V).to_host()
    for key in ("row", "col", "diag", "off", "b", "is_boundary"):
        assert np.array_equal(a[key], b[key]), key


def test_amg_on_device_assembled_system_exact():
    import torch
    import fvm_amd
    from fvm_amd.thermal import Mesh, assemble
    N = 12
    m = Mesh.hex(N, 0.2, 5)
    k, s, x = _dev_fields(m, 1.5, 1, 5)
    S = assemble(m, k, s, x, BC_T, BC_V)
    h = S.to_host()
    amg = fvm_amd.AMG()
    amg.verbosity = 0
    amg.nMaxIterations = 200
    from fvm_amd._lib import check, lib
    check(lib().fvm_amg_setup_system(amg._h, S._h))
    b = torch.from_numpy(h["b"]).cuda()
    xd = torch.zeros(h["n_total"], dtype=torch.float64, device="cuda:0")
    r0, rf, cyc = amg.solve_device(b, xd)
    oa = O.OracleAMG(n_max_iterations=200)
    oa.setup(h["n_self"], h["n_total"], h["row"], h["col"], h["diag"], h["off"], h["is_boundary"])
    ores = oa.solve(h["b"])
    assert cyc == ores["cycles"]
    assert np.array_equal(xd.cpu().numpy(), ores["x"])
    assert amg.num_levels() == oa.num_levels()
    for l in range(1, oa.num_levels()):
        assert np.array_equal(amg.level(l)["diag"], oa.level(l)["diag"])


def _fib(i, n=256):
    z = 1.0 - (2 * i + 1) / n
    phi = i * np.pi * (3.0 - np.sqrt(5.0))
    r = np.sqrt(1.0 - z * z)
    return np.array([r * np.cos(phi), r * np.sin(phi), z])


@pytest.mark.parametrize("N,jit,seed,i,tau", [(4, 0.0, 0, 0, 0.5), (6, 0.2, 3, 77, 3.0), (7, 0.2, 9, 200, 9.5)])
def test_phonon_system_bit_identical(N, jit, seed, i, tau):
    import torch
    from fvm_amd.thermal import Mesh, assemble_phonon
    m = Mesh.hex(N, jit, seed)
    ref = hexmesh.hex_mesh(N, jit, seed)
    nc = ref["n_cells"]
    rng = np.random.default_rng(seed)
    e0 = rng.uniform(0.5, 1.5, nc)
    x = np.zeros(m.n_total)
    fc = ref["face_cells"].reshape(-1, 2)
    gp = ref["group_ptr"]
    x[fc[gp[1]:gp[2], 1]] = 1.0
    x[:nc] = rng.uniform(0, 0.1, nc)
    te0 = torch.tensor(e0, dtype=torch.float64, device="cuda:0")
    tx = torch.tensor(x, dtype=torch.float64, device="cuda:0")
    vg = _fib(i)
    S = assemble_phonon(m, vg, tau, te0, tx).to_host()
    o = O.fv_phonon(ref, vg, tau, e0, x)
    for key in ("row", "col", "diag", "off", "b", "is_boundary"):
        assert np.array_equal(S[key], o[key]), key


@pytest.mark.parametrize("N,jit,seed", [(5, 0.2, 31), (8, 0.15, 32)])
def test_gradient_bit_identical(N, jit, seed):
    import torch
    from fvm_amd.thermal import Mesh
    m = Mesh.hex(N, jit, seed)
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, m.n_total)
    g = m.gradient(torch.tensor(x, dtype=torch.float64, device="cuda:0")).cpu().numpy()
    o = O.fv_gradient(hexmesh.hex_mesh(N, jit, seed), x)
    assert np.array_equal(g, o)


@pytest.mark.parametrize("N,jit,seed", [(5, 0.2, 41), (9, 0.2, 42)])
def test_thermal_system_with_nonuniform_field_bit_identical(N, jit, seed):
    """Secondary (non-orthogonal) flux from the least-squares gradient of a
    non-uniform current field, as in every outer iteration after the first."""
    import torch
    from fvm_amd.thermal import Mesh, assemble
    m = Mesh.hex(N, jit, seed)
    k, s, _ = _dev_fields(m, 1.5, 1, seed)
    rng = np.random.default_rng(seed)
    xh = rng.normal(0, 1, m.n_total)
    x = torch.tensor(xh, dtype=torch.float64, device="cuda:0")
    S = assemble(m, k, s, x, BC_T, BC_V).to_host()
    o = O.fv_thermal(hexmesh.hex_mesh(N, jit, seed), k.cpu().numpy(), s.cpu().numpy(), xh, BC_T, BC_V)
    for key in ("row", "col", "diag", "off", "b", "is_boundary"):
        assert np.array_equal(S[key], o[key]), key


@pytest.mark.parametrize("central", [False, True])
def test_convection_diffusion_system_bit_identical(central):
    """ConvectionDiscretization with a caller-given face flux, a continuity
    residual and outflow specified-temperature faces (extrapolation)."""
    import torch
    from fvm_amd.thermal import Mesh, assemble
    N, jit, seed = 7, 0.2, 51
    m = Mesh.hex(N, jit, seed)
    ref = hexmesh.hex_mesh(N, jit, seed)
    k, s, _ = _dev_fields(m, 1.0, 1, seed)
    rng = np.random.default_rng(seed)
    xh = rng.normal(0, 1, m.n_total)
    fA = O.fv_metrics(ref)["fA"].reshape(-1, 3)
    Fh = fA @ np.array([1.3, -0.4, 0.6]) + rng.normal(0, 0.01, m.n_faces)
    crh = rng.normal(0, 0.05, m.n_cells)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda:0")
    S = assemble(m, k, s, t(xh), BC_T, BC_V, flux=t(Fh), continuity_residual=t(crh), central=central).to_host()
    o = O.fv_thermal(ref, k.cpu().numpy(), s.cpu().numpy(), xh, BC_T, BC_V, conv_flux=Fh, cont_res=crh,
                     central=central)
    for key in ("row", "col", "diag", "off", "b", "is_boundary"):
        assert np.array_equal(S[key], o[key]), key
    # both x walls have outflow faces somewhere: some of their ghosts extrapolate
    assert S["is_boundary"].sum() > 4 * N * N


def test_neumann_heat_flux_system_bit_identical():
    """Specified heat fluxes q != 0 on the y and z walls (GenericBCS.h:123-150)."""
    import torch
    from fvm_amd.thermal import Mesh, assemble
    N, jit, seed = 6, 0.2, 61
    m = Mesh.hex(N, jit, seed)
    k, s, _ = _dev_fields(m, 1.0, 1, seed)
    rng = np.random.default_rng(seed)
    xh = rng.normal(0, 1, m.n_total)
    bt = [0, D, NM, NM, NM, NM, NM]
    bv = [0.0, 1.0, 0.7, -1.2, 2.5, 0.3, -0.4]
    x = torch.tensor(xh, dtype=torch.float64, device="cuda:0")
    S = assemble(m, k, s, x, bt, bv).to_host()
    o = O.fv_thermal(hexmesh.hex_mesh(N, jit, seed), k.cpu().numpy(), s.cpu().numpy(), xh, bt, bv)
    for key in ("row", "col", "diag", "off", "b", "is_boundary"):
        assert np.array_equal(S[key], o[key]), key
