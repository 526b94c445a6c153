"""Trajectory output with dynamical analysis, in the shape of the reference's
runTrajectory driver (runTrajectory.cpp:128-169): at log-spaced timesteps the
cell positions and the energy are written to one NetCDF file, and the
geodesic mean squared displacement and the overlap of the cell positions
against the start (computed on the GPU, every sphere at once) to two more.

    logSpacedIntegers   inc/simpleUtilities.h:5-45, the save schedule
    run_trajectory      the driver loop, for every sphere of a context
"""
import math
import os

import numpy as np

import curvedspacevertexmodel_amd as svm


class logSpacedIntegers:
    """Save points round(10^(exponent k)), k = 1, 2, ..., each once
    (inc/simpleUtilities.h:5-45)."""

    def __init__(self, firstSave=0, exponent=0.05):
        self.nextSave = firstSave
        self.exponent = exponent
        self.base = math.pow(10.0, exponent)
        self.logSaveIdx = 0
        if firstSave != 0:
            cur = 0
            while cur < firstSave:
                self.logSaveIdx += 1
                cur = int(_round_half_away(math.pow(self.base, self.logSaveIdx)))

    def update(self):
        self.logSaveIdx += 1
        cur = int(_round_half_away(math.pow(self.base, self.logSaveIdx)))
        while cur == self.nextSave:
            self.logSaveIdx += 1
            cur = int(_round_half_away(math.pow(self.base, self.logSaveIdx)))
        self.nextSave = cur


def _round_half_away(x):
    """C's round(): halves away from zero (Python's round() goes to even)."""
    return math.floor(x + 0.5) if x >= 0 else math.ceil(x - 0.5)


def run_trajectory(ctx, updater, steps, dt, out_dir, tag="traj", exponent=0.05, fraction=1.0, cutoff=0.5):
    """Advance every sphere `steps` timesteps of `updater` (0 Brownian, 1
    active), saving at the log-spaced steps: per sphere s,
    <tag>_cellPositions_s<s>.nc (vector = cell positions, value = energy),
    <tag>_msd_s<s>.nc and <tag>_overlap_s<s>.nc (vector = [t, msd] and
    [t, overlap], value = the step).  Returns the save steps and the
    per-save (msd, overlap) arrays [saves][spheres]."""
    os.makedirs(out_dir, exist_ok=True)
    n = ctx.n
    ctx.compute_geometry()
    ctx.dyn_set_reference()
    files = []
    for s in range(n):
        base = os.path.join(out_dir, f"{tag}_%s_s{s}.nc")
        files.append((svm.VectorValueNetCDF(base % "cellPositions", 3 * ctx.n_cells[s]),
                      svm.VectorValueNetCDF(base % "msd", 2), svm.VectorValueNetCDF(base % "overlap", 2)))
    lsi = logSpacedIntegers(0, exponent)
    lsi.update()
    saves, msds, overlaps = [], [], []
    try:
        for ii in range(steps):
            if ii == lsi.nextSave:
                lsi.update()
                ctx.compute_geometry()
                pe, ke = ctx.energies()        # sim->computeEnergy(): kinetic + model energy
                msd, ov = ctx.dyn_measure(fraction, cutoff)
                for s in range(n):
                    fp, fm, fo = files[s]
                    fp.writeState(ctx.cell_positions(s).ravel(), float(ke[s] + pe[s]))
                    fm.writeState([ii * dt, msd[s]], ii)
                    fo.writeState([ii * dt, ov[s]], ii)
                saves.append(ii)
                msds.append(msd)
                overlaps.append(ov)
            if updater == 0:
                ctx.brownian_steps(1)
            else:
                ctx.active_steps(1)
    finally:
        for trio in files:
            for f in trio:
                f.close()
    return np.array(saves), np.array(msds), np.array(overlaps)
