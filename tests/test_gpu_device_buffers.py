"""State through caller-owned device buffers (torch tensors on the GPU): the
device copies equal the host copies bit for bit, a state set from a device
buffer is the one the next steps start from, and pointers that are not device
memory are refused before any kernel touches them.  (The extent check sees
the device allocation, not a tensor inside torch's caching allocator, so a
short tensor is the caller's error and is not tested here.)"""
import numpy as np
import pytest
import torch

import svmtest
import curvedspacevertexmodel_amd as svm

pytestmark = pytest.mark.gpu


def make(n=800, seed=1250):
    ctx = svm.Context([dict(n_cells=n, a0=1.0, p0=3.8, kr=1.0, t1_threshold=0.1, seed=seed)])
    pts = svm.generate_cell_points(n, seed)
    ctx.build_from_cell_points(0, pts[svm.spatial_order(pts)])
    ctx.set_directors(0, svm.generate_directors(ctx.positions(0), seed))
    ctx.set_active(0, dt=0.01, mu=1.0, v0=0.2, tau=1.0)
    return ctx


def test_device_copies_equal_host_copies():
    ctx = make()
    ctx.active_steps(7)
    ctx.compute_forces()
    nv = ctx.positions(0).shape[0]
    for getter, host in (("get_positions_device", ctx.positions), ("get_directors_device", ctx.directors),
                         ("get_forces_device", ctx.forces), ("get_velocities_device", ctx.velocities)):
        buf = torch.empty((nv, 3), dtype=torch.float64, device="cuda")
        getattr(ctx, getter)(0, buf)
        ctx.synchronize()
        np.testing.assert_array_equal(buf.cpu().numpy(), host(0), err_msg=getter)


def test_state_set_from_device_buffers_drives_the_steps():
    a, b = make(), make()
    a.active_steps(5)
    b.active_steps(5)
    nv = a.positions(0).shape[0]
    # rotate sphere b's state on the device by a rigid rotation, and a's on the host
    th = 0.3
    Q = np.array([[np.cos(th), -np.sin(th), 0.0], [np.sin(th), np.cos(th), 0.0], [0.0, 0.0, 1.0]])
    pos = torch.empty((nv, 3), dtype=torch.float64, device="cuda")
    dirs = torch.empty((nv, 3), dtype=torch.float64, device="cuda")
    b.get_positions_device(0, pos)
    b.get_directors_device(0, dirs)
    b.synchronize()
    Qt = torch.tensor(Q, dtype=torch.float64, device="cuda")
    pos_r = (pos @ Qt.T).contiguous()
    dirs_r = (dirs @ Qt.T).contiguous()
    torch.cuda.synchronize()
    b.set_positions_device(0, pos_r)
    b.set_directors_device(0, dirs_r)
    a.set_positions(0, pos_r.cpu().numpy())
    a.set_directors(0, dirs_r.cpu().numpy())
    np.testing.assert_array_equal(b.positions(0), a.positions(0))
    a.active_steps(6)
    b.active_steps(6)
    np.testing.assert_array_equal(b.positions(0), a.positions(0))
    np.testing.assert_array_equal(b.directors(0), a.directors(0))


def test_bad_buffers_are_refused():
    ctx = make(300, 1251)
    nv = ctx.positions(0).shape[0]
    host = np.zeros((nv, 3))
    with pytest.raises(svm.SVMError) as e:
        ctx.get_positions_device(0, host.ctypes.data)     # host memory
    assert e.value.code == 1
    with pytest.raises(svm.SVMError) as e:
        ctx.set_positions_device(0, 0)
    assert e.value.code == 1
    # the state is untouched by the refused calls
    before = ctx.positions(0)
    ctx.active_steps(1)
    assert np.isfinite(ctx.positions(0)).all() and not np.array_equal(before, ctx.positions(0))
