"""Sedona's HDF5 model files (grid_2D_cyln::read_model_file,
src/grid/grid_2D_cyln.cpp:43-230) into the grid arrays a Problem takes
(`Problem(..., model=read_model_2d(path))`), read with pubsed_amd.hdf5.

2D cylindrical (varpi, z) model: `comp` [nx][nz][n_elems] gives the grid
shape and mass fractions, `Z` and `A` the elements, `time` the model time;
the zone edges come from `x_out` / `z_out` when both exist, else from `dr`
(uniform widths) with the lower corner `rmin` or, without it, varpi from 0
and z centred on 0 (:89-145); `rho`, `temp` per zone; zone volumes
pi (r1^2 - r0^2) dz with the reference's pi = 3.14159 (:217-222); zones
numbered with z fastest (:218).  Velocities are homologous (v = x / t) as
under `hydro_module = "homologous"` (grid_general.cpp:19-23), so `vx`,
`vz` are not read; `erad` is the initial radiation field, used only when
particles are initialised from it.
"""
import numpy as np

from .hdf5 import H5File
from . import workloads as W


def read_model_2d(path):
    f = H5File(path)
    comp = np.asarray(f["comp"], np.float64)
    nx, nzz, ne = comp.shape
    t = float(np.asarray(f["time"]).ravel()[0])
    if "x_out" in f and "z_out" in f:
        xe = np.asarray(f["x_out"], np.float64)
        ze = np.asarray(f["z_out"], np.float64)
        if "rmin" in f:
            x_min, z_min = (float(v) for v in np.asarray(f["rmin"], np.float64)[:2])
        else:
            x_min, z_min = 0.0, ze[0] - (ze[1] - ze[0])
    elif "dr" in f:
        dr = np.asarray(f["dr"], np.float64)
        if "rmin" in f:
            x_min, z_min = (float(v) for v in np.asarray(f["rmin"], np.float64)[:2])
        else:
            x_min, z_min = 0.0, -dr[1] * nzz / 2.0
        xe = x_min + dr[0] * np.arange(1, nx + 1)
        ze = z_min + dr[1] * np.arange(1, nzz + 1)
    else:
        raise ValueError("%s: neither dr nor x_out / z_out define the grid" % path)
    x0 = np.concatenate([[x_min], xe[:-1]])
    dz = np.diff(np.concatenate([[z_min], ze]))
    vol = (W.PI_REF * (xe * xe - x0 * x0))[:, None] * dz[None, :]
    return dict(kind=2, nz=nx * nzz, nx=nx, nzz=nzz, t=t, x_out=xe, z_out=ze, x_min=x_min, z_min=z_min,
                vol=vol.ravel(), rho=np.asarray(f["rho"], np.float64).ravel(),
                T=np.asarray(f["temp"], np.float64).ravel(), X=comp.reshape(nx * nzz, ne),
                elem_Z=np.asarray(f["Z"]).astype(np.int32), elem_A=np.asarray(f["A"]).astype(np.int32),
                e_rad=np.asarray(f["erad"], np.float64).ravel() if "erad" in f else None)
