"""Host-side logic without a GPU: record packing + the kernel's closed form
(numpy model) against the oracle, validation errors, and the C ABI exports."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from conftest import unit_square, rel
from kernel_model import model_apply
from oracle import sipg as S
from ddg_amd import mesh as pmesh
from ddg_amd.operator import pack
from ddg_amd.reference import ref_element

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6])
def test_packed_closed_form_matches_oracle(N):
    pm, om = unit_square(4, 3, N, jitter=0.1, seed=N)
    rng = np.random.default_rng(N)
    tab = ref_element(N).kernel_tables()
    for bc in (pmesh.all_dirichlet(pm), pmesh.dirichlet_x_neumann_y(pm)):
        for kw in ({}, {"kappa": np.exp(0.5 * rng.standard_normal(pm.K)),
                        "tau_scale": 200 * (1 + rng.random(pm.K))}):
            rec, nbr, _ = pack(pm, bc, **kw)
            u = rng.standard_normal((pm.Np, pm.K))
            ud = rng.standard_normal((pm.Np, pm.K))
            y = model_apply(N, tab, rec, nbr, u.T.copy(), udir=ud.T.copy(), alpha=2.5).T
            yo = S.sipg_apply_2d(om, u, bc_types=bc, u_dir=ud, **kw) + 2.5 * S.mass_apply(om, u)
            assert rel(y, yo) < 1e-12


def test_probe_mode_gives_diagonal_blocks():
    pm, om = unit_square(3, 3, 3)
    bc = pmesh.dirichlet_x_neumann_y(pm)
    rec, nbr, _ = pack(pm, bc)
    tab = ref_element(3).kernel_tables()
    blocks = np.stack([model_apply(3, tab, rec, nbr, np.tile(np.eye(pm.Np)[j], (pm.K, 1)),
                                   probe=True) for j in range(pm.Np)], axis=2)
    ref = S.extract_block_diag_matfree(S.flat_apply(om, bc), om)
    assert rel(blocks, ref) < 1e-12


def test_asymmetric_interior_tag_rejected():
    pm, _ = unit_square(2, 2, 2)
    bc = pmesh.all_dirichlet(pm)
    k = int(np.argmax((pm.EToE != np.arange(pm.K)[:, None]).any(1)))
    f = int(np.argmax(pm.EToE[k] != k))
    bc[k, f] = 1
    with pytest.raises(ValueError):
        pack(pm, bc)


def test_bad_bc_shape_rejected():
    pm, _ = unit_square(2, 2, 2)
    with pytest.raises(ValueError):
        pack(pm, np.zeros((pm.K, 2), dtype=np.int32))


def test_boundary_face_tagged_interior_is_no_coupling():
    # the reference gives such a face zero jump and cancelling averages
    pm, om = unit_square(3, 3, 2)
    bc = pmesh.all_dirichlet(pm)
    bc[pmesh.boundary_faces(pm) & (pmesh.face_normals(pm)[1] > 0.5)] = 0
    rec, nbr, _ = pack(pm, bc)
    u = np.random.default_rng(0).standard_normal((pm.Np, pm.K))
    y = model_apply(2, ref_element(2).kernel_tables(), rec, nbr, u.T.copy()).T
    assert rel(y, S.sipg_apply_2d(om, u, bc_types=bc)) < 1e-12


def _header_functions():
    txt = (ROOT / "include" / "ddgx.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(ddgx_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_header_symbol():
    lib = ROOT / "ddg_amd" / "libddgx.so"
    if not lib.exists():
        pytest.fail("libddgx.so not built: run `python -m ddg_amd.build`")
    L = ctypes.CDLL(str(lib))
    names = _header_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(L, n), n
    from ddg_amd import _lib
    assert set(names) == set(_lib.EXPORTS)


def test_error_codes_map_to_python_exceptions():
    """Every DDGX_ERR_* code of the header raises its Python exception."""
    import re
    from ddg_amd import _lib
    text = (ROOT / "include" / "ddgx.h").read_text()
    codes = {m.group(1): int(m.group(2)) for m in re.finditer(r"DDGX_(OK|ERR_\w+) = (\d+)", text)}
    assert codes == {"OK": 0, "ERR_INVALID": 1, "ERR_UNSUPPORTED": 2, "ERR_HIP": 3,
                     "ERR_NOMEM": 4, "ERR_BUSY": 5}
    _lib.check(0)
    kinds = {1: ValueError, 2: NotImplementedError, 3: _lib.DdgxError, 4: _lib.DdgxError,
             5: _lib.DdgxBusyError}
    for rc, kind in kinds.items():
        with pytest.raises(kind):
            _lib.check(rc, "call")


def test_table_size_matches_python():
    from ddg_amd import _lib
    L = _lib.lib()
    from ddg_amd.reference import mfma_fragments
    for N in range(1, 7):
        ref = ref_element(N)
        assert L.ddgx_table_size(N) == ref.kernel_tables().size + mfma_fragments(ref, "f64").size


@pytest.mark.parametrize("N", [1, 3, 4, 6])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_mfma_fragments_reproduce_apply(N, dtype):
    """Lane-level emulation of the matrix-core data flow (fragment layout,
    k permutation, output layout) against the closed-form model."""
    from kernel_model import model_apply_mfma
    from ddg_amd.reference import mfma_fragments, mfma_node_map
    pm, _ = unit_square(3, 3, N, jitter=0.1, seed=4)
    bc = pmesh.dirichlet_x_neumann_y(pm)
    rec, nbr, _ = pack(pm, bc)
    ref = ref_element(N)
    tab = ref.kernel_tables()
    u = np.random.default_rng(N).standard_normal((pm.K, pm.Np))
    y = model_apply_mfma(N, tab, mfma_fragments(ref, dtype), rec, nbr, u, dtype,
                         mfma_node_map(ref, dtype))
    assert rel(y, model_apply(N, tab, rec, nbr, u)) < 1e-12


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_mfma_node_map_face_traces(N, dtype):
    """Every node sits at exactly one accumulator position, and for N <= 4 the
    face-major trace writes of the kernel (register f, face node p = g or
    g + 1, plus the three vertices' second faces) cover each of the 3*Nfp
    face slots once, with the node Fmask puts there."""
    from ddg_amd.reference import mfma_node_map
    ref = ref_element(N)
    nm = mfma_node_map(ref, dtype)
    held = sorted(int(n) for n in nm.reshape(-1) if n < ref.Np)
    assert held == list(range(ref.Np))
    if nm.shape[0] != 1:
        return
    Nfp, Fm = ref.Nfp, ref.Fmask
    writes = {}
    for f in range(3):
        for g in range(N):
            writes.setdefault((f, g if f < 2 else g + 1), []).append(nm[0, g, f])
    writes.setdefault((2, 0), []).append(nm[0, 0, 0])
    writes.setdefault((0, N), []).append(nm[0, 0, 1])
    writes.setdefault((1, N), []).append(nm[0, N - 1, 2])
    assert sorted(writes) == [(f, p) for f in range(3) for p in range(Nfp)]
    for (f, p), nodes in writes.items():
        assert nodes == [Fm[p, f]]


def test_forest_hierarchy_matches_oracle():
    """The vectorised refinement (ddg_amd.mesh.refine_uniform) numbers
    vertices and elements exactly as the loop restatement of the reference's
    refinement forest does."""
    from ddg_amd import mesh as pmesh
    from ddg_amd.multigrid import build_h_meshes
    from oracle import multigrid as omg
    VX, VY, EToV = pmesh.uniform_rectangle_mesh_2d(0.0, 1.0, 0.0, 2.0, 3, 2)
    base = pmesh.build_mesh_2d(VX, VY, EToV, 1)
    ours = build_h_meshes(base, 3)
    ref = omg.forest_hierarchy(VX, VY, EToV, 3)
    assert len(ours) == len(ref) == 4
    for (a, b, c), (x, y, z) in zip(ours, ref):
        np.testing.assert_array_equal(a, x)
        np.testing.assert_array_equal(b, y)
        np.testing.assert_array_equal(c, z)


def test_transfer_matrices_match_oracle():
    from ddg_amd.reference import h_child_interp_matrices, p_interp_matrix
    from oracle import mesh as om
    from oracle import multigrid as omg
    VX, VY, EToV = om.uniform_rectangle_mesh_2d(0.0, 1.0, 0.0, 1.0, 1, 1)
    ms = {n: om.build_mesh_2d(VX, VY, EToV, n) for n in range(1, 7)}
    for lo, hi in [(1, 2), (2, 4), (1, 4), (3, 6), (5, 6)]:
        np.testing.assert_allclose(p_interp_matrix(lo, hi), omg.p_interp_matrix(ms[lo], ms[hi]),
                                   atol=1e-13)
    for n in range(1, 7):
        np.testing.assert_allclose(h_child_interp_matrices(n),
                                   omg.h_child_interp_matrices(ms[n]), atol=1e-13)


# Names the reference's own modules import from ddg.dg2d.sipg / multigrid /
# incompressible_ns (incompressible_ns.py:55-61, 1071-1073; multigrid.py:43-46;
# coupled_ns.py:133-138; scalar_transport.py:44-46; ins_metrics.py:33;
# velocity_reconstruction.py:211; test_matfree_preconditioner.py:29-33;
# applications/ns_rising_bubble_2d.py:53-56, ns_flow_past_cylinder_2d.py:59-62).
_CALLER_NAMES = {
    "sipg": ["BC_DIRICHLET", "BC_INTERIOR", "BC_NEUMANN", "IterativeConfig",
             "_face_1d_mass_matrices", "_local_mass", "_face_param_for_face", "_bc_face_arrays",
             "_color_elements", "_extract_block_diag_dense", "_block_jacobi_apply",
             "_block_diag_mass", "extract_block_diag_matfree", "sipg_bilinear_form",
             "sipg_apply_2d", "sipg_matrix_2d", "factor_sipg_2d", "sipg_dirichlet_load_2d",
             "sipg_neumann_load_2d", "sipg_helmholtz_matrix_2d", "factor_sipg_helmholtz_2d",
             "solve_sipg_poisson_2d", "solve_sipg_helmholtz_2d",
             "make_sipg_block_jacobi_preconditioner",
             "make_helmholtz_block_jacobi_preconditioner"],
    "multigrid": ["make_sipg_mg_preconditioner", "make_sipg_hpmg_preconditioner",
                  "_operator_apply", "_p_interp_matrix", "_h_child_interp_matrices",
                  "_build_h_meshes"],
    "incompressible_ns": ["_pressure_solve", "_viscous_solve"],
}


def test_drop_in_exports_the_names_callers_import():
    import importlib
    for mod, names in _CALLER_NAMES.items():
        m = importlib.import_module(f"ddg_amd.{mod}")
        missing = [n for n in names if not hasattr(m, n)]
        assert not missing, f"ddg_amd.{mod} lacks {missing}"


def test_host_side_helpers_match_oracle():
    from ddg_amd import sipg as D
    pm, om = unit_square(3, 2, 4, jitter=0.1, seed=4)
    bc = pmesh.dirichlet_x_neumann_y(pm)
    assert rel(D._local_mass(pm), S.local_mass(om)) < 1e-13
    assert rel(D._face_1d_mass_matrices(pm), S.face_1d_mass_matrices(om)) < 1e-13
    for f in range(3):
        Fm = om.Fmask[:, f]
        want = om.r[Fm] if f < 2 else om.s[Fm]
        np.testing.assert_allclose(D._face_param_for_face(pm, f), want, rtol=0, atol=1e-15)
    got = D._bc_face_arrays(pm, bc)
    want = S._bc_face_arrays(om, bc)
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)
    col, nc = D._color_elements(pm)
    colo, nco = S.color_elements(om)
    assert nc == nco <= 4 and np.array_equal(col, colo)
    A = S.sipg_matrix_2d(om, bc_types=bc, alpha=3.0)
    blocks = D._extract_block_diag_dense(A, pm.Np, pm.K)
    np.testing.assert_array_equal(blocks, S.extract_block_diag_dense(A, pm.Np, pm.K))
    r = np.random.default_rng(0).standard_normal(pm.K * pm.Np)
    assert rel(D._block_jacobi_apply(blocks, pm.Np, pm.K)(r), S.block_jacobi_apply(blocks)(r)) \
        < 1e-13
    Mb = D._block_diag_mass(pm, D._local_mass(pm))
    assert rel(Mb, np.kron(np.diag(om.J[0]), S.local_mass(om))) < 1e-13
