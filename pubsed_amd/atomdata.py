"""Sedona's atomic data files (`data/*.hdf5`) into the flattened arrays of
`sgpu_atoms` (include/sedona_gpu.h), read with pubsed_amd.hdf5.

AtomicData::read_atomic_data (src/opacity/AtomicData.cpp:155-174) picks the
reader by the file's `version` dataset: absent (1) is the "old style" that
every file under the reference's data/ uses, read as
read_atomic_data_oldstyle (:459-750) does:

* per element group `Z/`: attributes n_ions, n_levels, n_lines; datasets
  ion_chi, ion_ground, level_i, level_g, level_E, line_l, line_u, line_A;
* ion stages beyond max_ion_stage are dropped and a bare stage (chi 99999,
  one level of g 1) is appended in their place; levels are kept up to the
  first whose stage is dropped, lines up to the first whose lower level is
  dropped; statistical weights of 0 become 1; integer datasets stored as
  floats are truncated, as the HDF5 library's conversion does;
* line data derived as AtomicData.cpp:620-640 (nu = dE ev_to_ergs / h,
  B_ul, B_lu, f_lu, the grid bin), here by workloads.level_lines;
* photoionisation cross-sections: the hydrogenic approximation
  (:690-745), which the opacity kernels evaluate in place.

An element absent from the file gets the reference's placeholder (one ion
stage of chi 999999 with one level of g 1, :486-502).  Version-2 ("new
style") files (:180-452) are rejected: no file of that kind exists to test
a reader against.  Fuzz lines are not in these files (they come from a
separate fuzz file); the arrays for them are empty.
"""
import numpy as np

from .hdf5 import H5File
from . import workloads as W


class AtomDataError(ValueError):
    pass


def _int(a):
    """H5LTread_dataset_int of a dataset stored as integers or floats."""
    a = np.asarray(a)
    if a.dtype.kind == "f":
        return np.trunc(a).astype(np.int64)
    return a.astype(np.int64)


def read_element(f, Z, max_ion_stage=9999):
    """One element's (ion_chi, ion stage of each level, level E, level g,
    lines (ll, lu, A)) as read_atomic_data_oldstyle keeps them, level and
    line indices relative to the element."""
    g = str(int(Z))
    if g not in f:
        return (np.array([999999.0]), np.zeros(1, np.int64), np.zeros(1), np.ones(1, np.int64),
                (np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0)))
    at = f.attrs(g)
    if "n_ions" not in at or "n_levels" not in at:
        raise AtomDataError("element %s: n_ions / n_levels attributes missing" % g)
    n_tot_ions = int(at["n_ions"])
    n_ions = n_tot_ions
    add_last = n_ions > max_ion_stage
    if add_last:
        n_ions = max_ion_stage
    chi = np.asarray(f[g + "/ion_chi"], np.float64)[:n_ions].copy()
    ground = _int(f[g + "/ion_ground"])[:n_ions]
    n_tot_levels = int(at["n_levels"])
    lev_i = _int(f[g + "/level_i"])[:n_tot_levels]
    keep = np.nonzero(lev_i > n_ions - 1)[0]
    n_levels = int(keep[0]) if len(keep) else n_tot_levels
    lev_i = lev_i[:n_levels]
    lev_g = _int(f[g + "/level_g"])[:n_levels]
    lev_g = np.where(lev_g == 0, 1, lev_g)
    lev_E = np.asarray(f[g + "/level_E"], np.float64)[:n_levels].copy()
    n_tot_lines = int(at["n_lines"]) if "n_lines" in at else 0
    if n_tot_lines > 0:
        ll = _int(f[g + "/line_l"])[:n_tot_lines]
        drop = np.nonzero(ll >= n_levels)[0]
        n_lines = int(drop[0]) if len(drop) else n_tot_lines
        ll = ll[:n_lines]
        lu = _int(f[g + "/line_u"])[:n_lines]
        A = np.asarray(f[g + "/line_A"], np.float64)[:n_lines].copy()
    else:
        ll = lu = np.zeros(0, np.int64)
        A = np.zeros(0)
    if add_last:
        chi = np.concatenate([chi, [99999.0]])
        lev_i = np.concatenate([lev_i, [n_ions]])
        lev_g = np.concatenate([lev_g, [1]])
        lev_E = np.concatenate([lev_E, [0.0]])
        n_ions += 1
    # the kernels take each stage's ground level as its first level
    for s in range(n_ions):
        first = np.nonzero(lev_i == s)[0]
        want = ground[s] if s < len(ground) else (first[0] if len(first) else -1)
        if len(first) == 0 or first[0] != want or np.any(np.diff(lev_i) < 0):
            raise AtomDataError("element %s: levels must be ordered by ion stage with each stage's "
                                "ground level first" % g)
    return chi, lev_i, lev_E, lev_g, (ll, lu, A)


def read_atomic_data(path, elem_Z, nu_grid, max_ion_stage=9999):
    """The atoms of elem_Z (in that order) from a Sedona atomic data file, in
    the form of workloads.synthetic_atoms (with no fuzz lines)."""
    f = H5File(path)
    if "version" in f and int(np.asarray(f["version"]).ravel()[0]) != 1:
        raise AtomDataError("%s: only version-1 (old style) atomic data files are read" % path)
    lev_E, lev_g, lev_ion, ion_chi = [], [], [], []
    lev_off, ion_off = [0], [0]
    lines = []
    for a, Z in enumerate(elem_Z):
        chi, li, E, gw, (ll, lu, A) = read_element(f, Z, max_ion_stage)
        base = lev_off[-1]
        lev_E += list(E)
        lev_g += list(gw)
        lev_ion += list(li)
        ion_chi += list(chi)
        lev_off.append(len(lev_E))
        ion_off.append(len(ion_chi))
        for l, u, Av in zip(ll, lu, A):
            if E[u] - E[l] != 0:           # AtomicData.cpp:624 (no frequency for dE = 0)
                lines.append((a, base + int(l), base + int(u), float(Av)))
    n_atoms = len(elem_Z)
    lev_E = np.array(lev_E, np.float64)
    lev_g = np.array(lev_g, np.int32)
    lev_ion = np.array(lev_ion, np.int32)
    lev_off = np.array(lev_off, np.int32)
    pairs = (np.array([x[1] for x in lines], np.int32), np.array([x[2] for x in lines], np.int32),
             np.array([x[0] for x in lines], np.int32))
    A = np.array([x[3] for x in lines], np.float64)
    bl = W.level_lines(nu_grid, lev_off, lev_E, lev_g, lev_ion, None, A=A, pairs=pairs)
    nn = nu_grid["n"]
    return dict(**bl, n_atoms=n_atoms, lev_off=lev_off, ion_off=np.array(ion_off, np.int32),
                lev_E=lev_E, lev_g=lev_g, lev_ion=lev_ion, ion_chi=np.array(ion_chi, np.float64),
                atom_zero_eps=np.zeros(n_atoms, np.int32),
                max_levels=int(np.max(np.diff(lev_off))),
                line_nu=np.zeros(0), line_gf=np.zeros(0), line_El=np.zeros(0),
                line_ion=np.zeros(0, np.int32), line_bin=np.zeros(0, np.int32),
                line_bin_off=np.zeros(n_atoms * (nn + 1), np.int64), line_atom=np.zeros(0, np.int32))
