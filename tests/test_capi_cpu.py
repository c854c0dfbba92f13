"""The C-ABI library without a GPU: it loads, exports every symbol the public
header declares, and its host-side logic matches the reference."""
import ctypes

import pytest

from cbns3d_cart_mpi_amd import capi as C
from oracle import oracle as O


def test_library_exports_every_header_symbol():
    L = C.lib()
    names = C.header_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
        assert n in C._SIGS, f"{n} has no binding"


def test_params_derive_matches_reference_restatement():
    for N, Mach, Re, Pr, gamma, T_inf in [(64, 0.1, 1600.0, 0.71, 1.4, 255.556), (32, 0.5, 100.0, 0.72, 1.3, 300.0),
                                          (48, 0.3, 3200.0, 2.5, 1.67, 110.5)]:
        p = C.make_params(N, Mach=Mach, Re=Re, Pr=Pr, gamma=gamma, T_inf=T_inf)
        q = O.make_params(N, Mach=Mach, Re=Re, Pr=Pr, gamma=gamma, T_inf=T_inf)
        for f in ("Mach2", "gM2", "Re_inv", "p_inf", "gam1", "gam1_inv", "Pr_inv", "C_T_inf", "C_dt_v"):
            assert getattr(p, f) == getattr(q, f), f


def test_strerror_covers_statuses():
    for s in range(0, -9, -1):
        assert C.strerror(s) != "unknown status"
    assert C.strerror(-99) == "unknown status"


def test_null_arguments_are_rejected():
    L = C.lib()
    assert L.cbns_params_derive(None) == -1
    assert L.cbns_destroy(None) == -1
    assert L.cbns_step(None, None, None) == -1
    assert L.cbns_face_flux_eval(0, 5, 1.4, 1, None, None, None) == -1
    for f in ("cbns_batch_set_band", "cbns_batch_set_overlap", "cbns_batch_set_workgroups"):
        assert getattr(L, f)(None, 32) == -1, f
    assert L.cbns_batch_sync(None) == -1
    assert L.cbns_batch_destroy(None) == -1
    # output and diagnostics beside the step
    assert L.cbns_snapshot_begin(None) == -1
    assert L.cbns_snapshot_wait(None, None) == -1
    assert L.cbns_snapshot_block_state(None, 0, None) == -1
    assert L.cbns_snapshot_block_primitive(None, 0, 0, None) == -1
    assert L.cbns_spectra(None, 8, None, None, None) == -1
    assert L.cbns_get_block_mu(None, 0, None) == -1
    assert L.cbns_set_block_mu(None, 0, None) == -1


def test_batch_create_without_gpu_reports_no_device():
    if C.device_count() > 0:
        pytest.skip("a gfx950 device is present")
    h = ctypes.c_void_p()
    assert C.lib().cbns_batch_create(1 << 30, 4, ctypes.byref(h)) == -4
    assert C.lib().cbns_batch_create(1 << 30, 0, ctypes.byref(h)) == -1


def test_create_without_gpu_reports_no_device():
    if C.device_count() > 0:
        pytest.skip("a gfx950 device is present")
    p = C.make_params(16)
    h = ctypes.c_void_p()
    d = (ctypes.c_int * 3)(1, 1, 1)
    assert C.lib().cbns_create(ctypes.byref(p), d, 0, ctypes.byref(h)) == -4


def test_create_rejects_blocks_thinner_than_the_halo():
    p = C.make_params((16, 16, 16))
    h = ctypes.c_void_p()
    d = (ctypes.c_int * 3)(5, 1, 1)  # 3-cell blocks
    assert C.lib().cbns_create(ctypes.byref(p), d, 0, ctypes.byref(h)) == -1


def test_create_rejects_blocks_past_the_32_bit_field_bound():
    """A block with more than 2^29 ghosted cells is refused before any device
    work (its fields would pass 4 GB); the same grid over more blocks is not
    refused for its size (without a GPU it then reports the missing device)."""
    p = C.make_params((1024, 1024, 1024))
    h = ctypes.c_void_p()
    one = (ctypes.c_int * 3)(1, 1, 1)
    assert C.lib().cbns_create(ctypes.byref(p), one, 0, ctypes.byref(h)) == -1
    q = C.make_params((800, 800, 800))  # 808^3 < 2^29
    if C.device_count() == 0:
        assert C.lib().cbns_create(ctypes.byref(q), one, 0, ctypes.byref(h)) == -4
        eight = (ctypes.c_int * 3)(2, 2, 2)
        assert C.lib().cbns_create(ctypes.byref(p), eight, 0, ctypes.byref(h)) == -4
