"""Reference call semantics at the drop-in boundary (giem2g_interfaces.F90):

- a background (frequency) change after the conductivity was set derives the
  per-cell scalings from the new background, as the reference does at every
  apply (apply_ie_operator_module.F90:66-68);
- giem2g_set_anomaly_conductivity aliases the caller's array (:199): an
  in-place change is seen by the next giem2g_apply_ie_operator;
- giem2g_calc_ie_kernel takes the geometry from its anomaly argument and
  derives Lnumber from it (:158-166);
- the counters of the reference's split (dot products, FFT, contraction),
  with profiling on;
- device entry points reject tensors of the wrong dtype or size on the host.
Each result is compared with the oracle at the FP64 bar (1e-12).
"""
import ctypes as C
import dataclasses

import numpy as np
import pytest

from tests.conftest import rel

pytestmark = pytest.mark.gpu


def oracle_model(po, m):
    return po.Model(m.csigb(), m.thick, m.omega, m.Nx, m.Ny, m.Nz, m.dx, m.dy, m.z)


def rand_vec(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def oracle_apply(po, m, sa, v):
    om = oracle_model(po, m)
    Gs = om.tensor_fft(om.tensor_realspace())
    return om.apply(Gs, sa, v)


def test_background_change_after_sigma_rederives_scalings(oracle):
    from giem2g_amd import models, operator as gop
    m = models.small(16, 16, 4, seed=11, freq=1.0)
    op = gop.operator_for(m)
    sa = m.csiga()  # sigma_a - i w eps0 of the first frequency, kept by the operator
    m2 = dataclasses.replace(m, freq=7.0)
    op.set_background(m2.csigb(), m2.thick, m2.omega)
    op.build_tensor()
    v = rand_vec(3 * m.ncell, 1)
    assert rel(op.apply_host(v), oracle_apply(oracle, m2, sa, v)) < 1e-12
    op.close()


def _reference_operator(lib, L, m, z=None, dx=None):
    z = np.ascontiguousarray(m.z if z is None else z)
    dz = np.ascontiguousarray(np.diff(z))
    an = L.Anomaly(m.Nx, m.Ny, m.Nz, m.dx if dx is None else dx, m.dy, z.ctypes.data, dz.ctypes.data)
    data = L.Data()
    lib.giem2g_calc_data_sizes(an, C.byref(data))
    lib.giem2g_prepare_ie_kernel(an, data)
    return data.giem2g_tensor, an, (z, dz)


def test_reference_abi_aliases_conductivity(oracle):
    from giem2g_amd import lib as L, models
    lib = L.load()
    m = models.small(16, 8, 5, seed=12, freq=3.0)
    h, an, keep = _reference_operator(lib, L, m)
    cs = np.ascontiguousarray(m.csigb())
    th = np.ascontiguousarray(np.array(m.thick, np.float64))
    bkg = L.Background(m.Nl, cs.ctypes.data, th.ctypes.data)
    lib.giem2g_calc_ie_kernel(h, bkg, an, m.omega)
    lib.giem2g_calc_fft_of_ie_kernel(h)
    sa = m.csiga()
    lib.giem2g_set_anomaly_conductivity(h, sa.ctypes.data)
    v = rand_vec(3 * m.ncell, 2)
    out = np.empty_like(v)
    lib.giem2g_apply_ie_operator(h, v.ctypes.data, out.ctypes.data)
    assert rel(out, oracle_apply(oracle, m, sa.copy(), v)) < 1e-12
    # change the caller's array in place: the next apply must see it
    sa[:] = sa[::-1] * (0.5 + 0.25j)
    lib.giem2g_apply_ie_operator(h, v.ctypes.data, out.ctypes.data)
    assert rel(out, oracle_apply(oracle, m, sa.copy(), v)) < 1e-12
    lib.giem2g_delete_ie_operator(h)


def test_calc_ie_kernel_takes_geometry_from_anomaly(oracle):
    """Prepared with one geometry, built with another anomaly (other dx and
    layer borders crossing the 1 km interface): the tensor and the operator
    are the oracle's for the anomaly given to calc_ie_kernel."""
    from giem2g_amd import lib as L, models
    lib = L.load()
    m = models.small(16, 16, 6, seed=13, freq=2.0)
    h, _, keep = _reference_operator(lib, L, m)
    m2 = dataclasses.replace(m, dx=150.0, z=940.0 + np.arange(m.Nz + 1) * 20.0)
    z2 = np.ascontiguousarray(m2.z)
    dz2 = np.ascontiguousarray(np.diff(z2))
    an2 = L.Anomaly(m2.Nx, m2.Ny, m2.Nz, m2.dx, m2.dy, z2.ctypes.data, dz2.ctypes.data)
    cs = np.ascontiguousarray(m2.csigb())
    th = np.ascontiguousarray(np.array(m2.thick, np.float64))
    bkg = L.Background(m2.Nl, cs.ctypes.data, th.ctypes.data)
    lib.giem2g_calc_ie_kernel(h, bkg, an2, m2.omega)
    om2 = oracle_model(oracle, m2)
    Gr = om2.tensor_realspace()
    G = np.empty_like(Gr)
    L.check(lib.giem2g_op_get_tensor(h, G.ctypes.data))
    assert rel(G, Gr) < 1e-10
    lib.giem2g_calc_fft_of_ie_kernel(h)
    sa = m2.csiga()
    lib.giem2g_set_anomaly_conductivity(h, sa.ctypes.data)
    v = rand_vec(3 * m.ncell, 3)
    out = np.empty_like(v)
    lib.giem2g_apply_ie_operator(h, v.ctypes.data, out.ctypes.data)
    assert rel(out, om2.apply(om2.tensor_fft(Gr), sa, v)) < 1e-12
    # an anomaly of other dimensions is refused (logged, operator unchanged)
    msgs = []
    cb = L.LOGGER(lambda s: msgs.append(s.decode()))
    lib.giem2g_set_logger(cb)
    bad = L.Anomaly(m.Nx, m.Ny, m.Nz + 1, m.dx, m.dy, z2.ctypes.data, dz2.ctypes.data)
    lib.giem2g_calc_ie_kernel(h, bkg, bad, m2.omega)
    lib.giem2g_set_logger(L.LOGGER())
    assert msgs and "dimensions differ" in msgs[0]
    lib.giem2g_delete_ie_operator(h)


def test_profiled_solve_counters_and_stats(oracle):
    import torch
    from giem2g_amd import lib as L, models, operator as gop
    m = models.c0()
    op = gop.operator_for(m)
    op.reset_counters()
    op.set_profiling(True)
    E = torch.from_numpy(op.plane_wave(1)).cuda()
    _, info = op.solve(E, outer_m=10, inner_m=10, max_restart=20, tol=1e-8)
    c = op.counters()
    assert c["profiling"] == 1
    assert c["matvecs"] == info["matvecs_fp64"] == c["contr_launches_fp64"] == c["contr_rhs_fp64"]
    assert c["t_contr_fp64"] > 0 and c["t_fwd"] > 0 and c["t_bwd"] > 0 and c["t_dot"] > 0
    assert abs(c["t_apply"] - (c["t_fwd"] + c["t_contr"] + c["t_bwd"])) < 1e-9
    assert c["gs_passes"] >= info["outer_iterations"] and c["dotprods"] > c["gs_passes"]
    # every pass reads its basis columns at least once
    assert c["krylov_bytes"] > 16 * 3 * m.ncell * c["gs_passes"]
    msgs = []
    cb = L.LOGGER(lambda s: msgs.append(s.decode()))
    op.lib.giem2g_set_logger(cb)
    op.lib.giem2g_print_stats(op.h)
    op.lib.giem2g_set_logger(L.LOGGER())
    assert "Time for dot products" in msgs[0]
    assert "Number of dot products: %d" % c["dotprods"] in msgs[0]
    assert op.counters()["matvecs"] == 0  # print_stats resets
    op.set_profiling(False)
    op.close()


def test_device_entry_points_check_tensors():
    import torch
    from giem2g_amd import models, operator as gop
    m = models.small(8, 8, 8, seed=2)
    op = gop.operator_for(m)
    v = torch.zeros(3 * m.ncell, dtype=torch.complex128, device="cuda")
    with pytest.raises(ValueError, match="complex128"):
        op.apply_dev(v.to(torch.complex64), torch.empty_like(v))
    with pytest.raises(ValueError, match="elements"):
        op.apply_dev(v[:-1], torch.empty_like(v))
    with pytest.raises(ValueError, match="complex64"):
        op.apply_dev_f32(v, torch.empty_like(v))
    vb = torch.zeros(2, 3 * m.ncell, dtype=torch.complex128, device="cuda")
    with pytest.raises(ValueError, match="complex128"):
        op.apply_batch_dev(vb, torch.empty_like(vb, dtype=torch.complex64))
    with pytest.raises(ValueError, match="elements"):
        op.apply_pair_dev(v, v, v, v[:-3])
    op.close()
