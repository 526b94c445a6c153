"""Measurement yardsticks of the library run and order their data.

svm_phase_boundary_probe times a dependency between two phases as a kernel
boundary in a HIP graph and as a grid barrier in one persistent launch; in
both forms every workgroup must read exactly what its neighbour wrote in the
phase before.
"""
import pytest

import curvedspacevertexmodel_amd as svm

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("workgroups,mib", [(256, 1), (512, 8)])
def test_phase_boundary_probe_orders_the_data(workgroups, mib):
    launch_us, barrier_us, wrong = svm.phase_boundary_probe(workgroups, 20, mib << 20)
    assert wrong == 0
    assert launch_us > 0 and barrier_us > 0


def test_phase_boundary_probe_refuses_bad_arguments():
    with pytest.raises(svm.SVMError):
        svm.phase_boundary_probe(0, 20)
    with pytest.raises(svm.SVMError):
        svm.phase_boundary_probe(1 << 20, 20)        # more workgroups than the device keeps resident
