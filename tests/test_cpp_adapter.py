"""The C++ adapter (include/svm.hpp) and a reference-style driver over it
(tests/cpp/driver.cpp): it compiles against the C ABI, fails loudly without a
GPU, and on the GPU gives the same numbers as the same sequence of calls made
through the Python binding."""
import os
import shutil
import subprocess

import pytest

import svmtest

ROOT = svmtest.ROOT
LIBDIR = os.path.join(ROOT, "curvedspacevertexmodel_amd")


def build_driver(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libsvm.so")):
        pytest.skip("libsvm.so is not built")
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("no g++")
    exe = str(tmp_path / "driver")
    subprocess.run([gxx, "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "driver.cpp"), "-L", LIBDIR, "-lsvm",
                    f"-Wl,-rpath,{LIBDIR}", "-o", exe], check=True, capture_output=True, text=True)
    return exe


def test_driver_compiles_and_fails_loudly_without_gpu(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    exe = build_driver(tmp_path)
    r = subprocess.run([exe, "64", "1", "3.8", "5"], capture_output=True, text=True)
    assert r.returncode == 1
    assert "svm_create failed" in r.stderr


@pytest.mark.gpu
def test_driver_matches_python_binding(tmp_path):
    import curvedspacevertexmodel_amd as svm
    exe = build_driver(tmp_path)
    n, seed, p0, steps = 512, 1137, 3.8, 100
    r = subprocess.run([exe, str(n), str(seed), str(p0), str(steps)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = dict(line.split() for line in r.stdout.strip().splitlines())
    ctx = svm.Context([dict(n_cells=n, a0=1.0, p0=p0, kr=1.0, t1_threshold=0.1, seed=seed)])
    ctx.build_from_cell_points(0, svm.generate_cell_points(n, seed))
    ctx.set_directors(0, svm.generate_directors(ctx.positions(0), seed))

    def energy():
        ctx.compute_geometry()
        return ctx.energies()[0][0]

    ctx.set_preferred_parameters(0, 1.0, 1.0)
    ctx.set_brownian(0, dt=0.01, mu=1.0, T=0.0)
    ctx.brownian_steps(200)
    assert float(out["energy_after_stabilise"]) == energy()
    ctx.set_preferred_parameters(0, 1.0, p0)
    ctx.set_active(0, dt=0.01, mu=1.0, v0=0.1, tau=1.0)
    ctx.active_steps(steps)
    assert float(out["energy_after_active"]) == energy()
    ctx.fire_set_parameters(0, dt=0.001, alpha_start=0.99, dt_max=0.05, dt_inc=1.1, dt_dec=0.95, alpha_dec=0.9,
                            n_min=4, force_cutoff=1e-10, alpha_min=0.0, stop_rms=1, max_iterations=20000)
    ctx.fire_minimize(100)
    st = ctx.fire_state(0)
    assert int(out["fire_iterations"]) == st["iterations"]
    assert float(out["energy_after_fire"]) == energy()
    cs = 0.0
    for x, y, z in ctx.positions(0).tolist():   # the driver's summation order
        cs += x + 2.0 * y + 3.0 * z
    assert float(out["position_checksum"]) == cs
    vctx = svm.Context([dict(n_cells=n, a0=1.0, p0=p0, kr=1.0, t1_threshold=0.1, seed=seed)])
    vctx.build_voronoi_from_cell_points(0, svm.generate_cell_points(n, seed))
    vctx.set_brownian(0, dt=0.01, mu=1.0, T=0.0)
    vctx.voronoi_brownian_steps(5)
    vctx.compute_geometry()
    assert float(out["voronoi_energy"]) == vctx.energies()[0][0]
    cs = 0.0
    for x, y, z in vctx.generators(0).tolist():
        cs += x + 2.0 * y + 3.0 * z
    assert float(out["voronoi_checksum"]) == cs
    counts, idx, vec, dist = svm.neighbor_list(ctx.positions(0), 1.0)
    assert int(out["neighbor_pairs"]) == counts.sum() and int(out["neighbor_nmax"]) == max(4, idx.shape[1])
    cs = 0.0
    for i in range(len(counts)):                 # the driver's summation order
        for k in range(counts[i]):
            cs += float(idx[i, k]) + float(dist[i, k])
    assert float(out["neighbor_checksum"]) == cs
