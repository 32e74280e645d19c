"""Argument checks of the Python binding that run before any launch (no GPU):
propagators whose size does not match the lattice, outputs of the wrong
shape or dtype, too-small site-scaled buffers, mixed devices and host
tensors are rejected with ValueError/TypeError instead of reaching a kernel
that would read or write past a buffer."""

import numpy as np
import pytest
import torch

from contract_amd import pycontract as pc

DIMS = (4, 4, 4, 8)
V = 512
TERMS = [(0, 1.0), (5, 1.0)]


def prop(v=V, dtype=torch.complex128):
    return torch.zeros((v, 4, 4, 3, 3), dtype=dtype)


def test_project_rejects_propagator_of_other_volume():
    with pytest.raises(ValueError, match="holds 256 sites"):
        pc.baryonTwoPointProject(prop(), prop(256), prop(), TERMS, 5, 5, [0], [(0, 0, 0)], DIMS)
    with pytest.raises(ValueError, match="holds 512 sites, dims"):
        pc.baryonTwoPointProject(prop(), prop(), prop(), TERMS, 5, 5, [0], [(0, 0, 0)], (4, 4, 4, 16))


def test_project_rejects_bad_output():
    ps = [prop(), prop(), prop()]
    with pytest.raises(ValueError, match="shape"):
        pc.baryonTwoPointProject(*ps, TERMS, 5, 5, [0, 1], [(0, 0, 0)], DIMS,
                                 out=torch.zeros((1, 1, 8), dtype=torch.complex128))
    with pytest.raises(ValueError, match="complex128"):
        pc.baryonTwoPointProject(*ps, TERMS, 5, 5, [0], [(0, 0, 0)], DIMS,
                                 out=torch.zeros((1, 1, 8), dtype=torch.complex64))
    with pytest.raises(ValueError, match="contiguous"):
        pc.baryonTwoPointProject(*ps, TERMS, 5, 5, [0], [(0, 0, 0)], DIMS,
                                 out=torch.zeros((1, 1, 16), dtype=torch.complex128)[:, :, ::2])


def test_project_rejects_host_tensors():
    with pytest.raises(ValueError, match="GPU"):
        pc.baryonTwoPointProject(prop(), prop(), prop(), TERMS, 5, 5, [0], [(0, 0, 0)], DIMS)


def test_scaled_propagator_size_is_checked():
    short = pc.ScaledPropagator(torch.zeros(V * 1152, dtype=torch.uint8), V)  # exponents missing
    ok = pc.ScaledPropagator(torch.zeros(V * 1153, dtype=torch.uint8), V)
    with pytest.raises(ValueError, match="needs"):
        pc.baryonTwoPointProject(short, ok, ok, TERMS, 5, 5, [0], [(0, 0, 0)], DIMS)
    wrong = pc.ScaledPropagator(torch.zeros(256 * 1153, dtype=torch.uint8), 256)
    with pytest.raises(ValueError, match="holds 256 sites"):
        pc.baryonTwoPointProject(ok, wrong, ok, TERMS, 5, 5, [0], [(0, 0, 0)], DIMS)
    with pytest.raises(TypeError):
        pc.baryonTwoPointProject(ok, prop(dtype=torch.complex64), ok, TERMS, 5, 5, [0], [(0, 0, 0)], DIMS)


def test_batch_checks_every_request_and_output():
    good = (prop(), prop(), prop(), DIMS)
    bad = (prop(), prop(), prop(1024), DIMS)
    with pytest.raises(ValueError, match="request 1"):
        pc.baryonTwoPointProjectBatch([good, bad], TERMS, 5, 5, [0], [(0, 0, 0)])
    with pytest.raises(ValueError, match="2 outputs for 1 requests"):
        pc.baryonTwoPointProjectBatch([good], TERMS, 5, 5, [0], [(0, 0, 0)],
                                      outs=[torch.zeros((1, 1, 8), dtype=torch.complex128)] * 2)
    with pytest.raises(ValueError, match=r"outs\[0\]"):
        pc.baryonTwoPointProjectBatch([good], TERMS, 5, 5, [0], [(0, 0, 0)],
                                      outs=[torch.zeros((1, 2, 8), dtype=torch.complex128)])


def test_meson_grid_checks_requests_and_outputs():
    with pytest.raises(ValueError, match="request 0"):
        pc.mesonGridProject([(prop(), prop(256), DIMS)], [(0, 0, 0)])
    with pytest.raises(ValueError, match=r"outs\[0\] has shape"):
        pc.mesonGridProject([(prop(), prop(), DIMS)], [(0, 0, 0)],
                            outs=[torch.zeros((16, 16, 2, 8), dtype=torch.complex128)])
    with pytest.raises(ValueError, match="GPU"):
        pc.mesonGridProject([(prop(), prop(), DIMS)], [(0, 0, 0)])


def test_drop_in_helpers_check_volumes_and_dtype():
    with pytest.raises(ValueError, match="different site counts"):
        pc.mesonTwoPoint(prop(), prop(256), 5, 5)
    with pytest.raises(ValueError, match="different site counts"):
        pc.baryonTwoPoint(prop(), prop(), prop(508), 0, 5, 5, 8)
    with pytest.raises(TypeError, match="complex128"):
        pc.baryonDiquark(prop(), prop(dtype=torch.complex64), 5, 5)
    with pytest.raises(ValueError, match="GPU"):
        pc.baryonTwoPoint(prop(), prop(), prop(), 0, 5, 5, 8)


def test_diquark_binding_checks_forms_and_sizes():
    ok = pc.ScaledPropagator(torch.zeros(V * 1153, dtype=torch.uint8), V)
    with pytest.raises(TypeError, match="one storage form"):
        pc.baryonDiquarkInto(prop(), ok, ok, 5, 5)
    small = pc.ScaledPropagator(torch.zeros((V - 4) * 1153, dtype=torch.uint8), V - 4)
    with pytest.raises(ValueError, match="different site counts"):
        pc.baryonDiquarkInto(small, ok, ok, 5, 5)
    with pytest.raises(ValueError, match="GPU"):  # sizes pass, host tensors are refused
        pc.baryonDiquarkInto(ok, ok, ok, 5, 5)
    with pytest.raises(TypeError, match="complex128"):
        pc.baryonDiquarkInto(prop(), prop(dtype=torch.complex64), prop(), 5, 5)


def test_raw_abi_checks_tensor_spans():
    with pytest.raises(ValueError, match="propag_j holds"):
        pc.meson_two_point(torch.zeros(V, dtype=torch.complex128), prop(), prop(256), V, 5, 5)
    with pytest.raises(ValueError, match="correl holds"):
        pc.baryon_two_point(torch.zeros(V // 2, dtype=torch.complex128), prop(), prop(), prop(), 0, V, 5, 5, 8)
    with pytest.raises(ValueError, match=r"correl\[3\] holds"):
        pc.meson_all_sink_two_point([torch.zeros(V if g != 3 else 8, dtype=torch.complex128) for g in range(16)],
                                    prop(), prop(), V, 5)
    with pytest.raises(ValueError, match="GPU tensor"):
        pc.baryon_diquark(prop(), prop(), prop(), V, 5, 5)


def test_scaled_conversion_output_is_checked():
    with pytest.raises(ValueError, match="out holds"):
        pc.toScaledFP32(prop(), out=torch.zeros(V * 1152, dtype=torch.uint8))


def test_baryon_plan_setting_round_trips():
    assert pc.get_baryon_plan() == (8, 128)
    with pc.baryon_plan((16, 256)):
        assert pc.get_baryon_plan() == (16, 256)
        with pc.baryon_plan((0, 64)):  # zero takes the default
            assert pc.get_baryon_plan() == (8, 64)
        assert pc.get_baryon_plan() == (16, 256)
    assert pc.get_baryon_plan() == (8, 128)
    with pytest.raises(ValueError, match="items_per_cu"):
        pc.set_baryon_plan((1000, 128))
    with pytest.raises(ValueError, match="min_item_sites"):
        pc.set_baryon_plan((8, 2))
    assert pc.get_baryon_plan() == (8, 128)


def test_workspace_bytes_follow_the_plan():
    small = pc.project_workspace_bytes((24, 24, 24, 64), 27)
    with pc.baryon_plan((16, 4)):
        assert pc.project_workspace_bytes((24, 24, 24, 64), 27) > small
    assert pc.project_workspace_bytes((24, 24, 24, 64), 27) == small
    assert np.isfinite(small) and small > 0
