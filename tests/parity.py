"""Edgel-by-edgel comparison of a GPU edgel list with the FP64 oracle.

Edgels are matched by their interp-grid pixel (row * 2W + col).  The stated
tolerances (per matched edgel) are

    position   |dx|, |dy| <= 1e-3 px
    angle      |d theta| <= 1e-3 rad   (NaN must match NaN)
    strength   |ds| <= 1e-4 * |s|

An edgel present in only one list is accepted when the oracle's comparison
margin at that pixel is below EPS_MAG (gradient-magnitude units, threshold and
peak test) or EPS_SUB (|s*| against sqrt(2), interp px); such edgels are counted
and reported, not hidden.
"""
import numpy as np

TOL_XY = 1e-3
TOL_THETA = 1e-3
TOL_S_REL = 1e-4
EPS_MAG = 2e-3
EPS_SUB = 2e-3


def compare(ref, gpu_edges, gpu_index):
    """ref: oracle.detect(..., margins=True) result; gpu_*: numpy arrays."""
    g = np.asarray(gpu_edges, np.float64)
    gi = np.asarray(gpu_index, np.int64)
    oi = ref["eidx"]
    rep = {"n_oracle": int(len(oi)), "n_gpu": int(len(gi))}
    rep["gpu_sorted"] = bool(np.all(np.diff(gi) > 0)) if len(gi) > 1 else True
    common, ia, ib = np.intersect1d(oi, gi, return_indices=True)
    only_o = np.setdiff1d(oi, gi)
    only_g = np.setdiff1d(gi, oi)
    mm = ref["mmag"].reshape(-1)
    ms = ref["msub"].reshape(-1)

    def explained(ix):
        return (mm[ix] < EPS_MAG) | (ms[ix] < EPS_SUB)

    rep["only_oracle"] = int(len(only_o))
    rep["only_gpu"] = int(len(only_g))
    rep["unexplained_set"] = int((~explained(only_o)).sum() + (~explained(only_g)).sum()) if (
        len(only_o) + len(only_g)) else 0
    a = ref["edges"][ia]
    b = g[ib]
    dxy = np.abs(a[:, :2] - b[:, :2]).max(1) if len(a) else np.zeros(0)
    th_a, th_b = a[:, 2], b[:, 2]
    nan_mismatch = np.isnan(th_a) != np.isnan(th_b)
    dth = np.abs(np.where(np.isnan(th_a) | np.isnan(th_b), 0.0, th_a - th_b))
    dth = np.minimum(dth, np.abs(dth - 2 * np.pi))   # atan2 branch cut at +-pi
    ds = np.abs(a[:, 3] - b[:, 3]) / np.maximum(np.abs(a[:, 3]), 1e-30)
    bad = (dxy > TOL_XY) | (dth > TOL_THETA) | (ds > TOL_S_REL) | nan_mismatch
    rep["matched"] = int(len(common))
    rep["max_dxy"] = float(dxy.max()) if len(dxy) else 0.0
    rep["max_dtheta"] = float(dth.max()) if len(dth) else 0.0
    rep["max_ds_rel"] = float(ds.max()) if len(ds) else 0.0
    rep["p999_dxy"] = float(np.quantile(dxy, 0.999)) if len(dxy) else 0.0
    rep["out_of_tol"] = int(bad.sum())
    rep["out_of_tol_explained"] = int((bad & explained(common)).sum())
    rep["theta_nan"] = int(np.isnan(th_a).sum())
    return rep


def assert_parity(rep, max_out_of_tol_frac=0.0):
    assert rep["gpu_sorted"], rep
    assert rep["unexplained_set"] == 0, rep
    unexpl = rep["out_of_tol"] - rep["out_of_tol_explained"]
    assert unexpl <= max_out_of_tol_frac * max(rep["matched"], 1), rep


def compare_subpix(ref, gpu_edges, gpu_index):
    """Subpixel correction path: GPU list vs oracle.subpix() result, matched
    by map pixel (i * W + j).  Same tolerances as the detection path.  The
    edgel set is decided in FP64 on both sides (orientation estimate, NMS,
    isolation distances), so it must be identical."""
    g = np.asarray(gpu_edges, np.float64)
    gi = np.asarray(gpu_index, np.int64)
    oi = ref["idx"]
    rep = {"n_oracle": int(len(oi)), "n_gpu": int(len(gi))}
    rep["gpu_sorted"] = bool(np.all(np.diff(gi) > 0)) if len(gi) > 1 else True
    common, ia, ib = np.intersect1d(oi, gi, return_indices=True)
    rep["only_oracle"] = int(len(oi) - len(common))
    rep["only_gpu"] = int(len(gi) - len(common))
    a, b = ref["edges"][ia], g[ib]
    dxy = np.abs(a[:, :2] - b[:, :2]).max(1) if len(a) else np.zeros(0)
    dth = np.abs(np.angle(np.exp(1j * (a[:, 2] - b[:, 2])))) if len(a) else np.zeros(0)
    ds = np.abs(a[:, 3] - b[:, 3]) / np.maximum(np.abs(a[:, 3]), 1e-30) if len(a) else np.zeros(0)
    rep["max_dxy"] = float(dxy.max()) if len(dxy) else 0.0
    rep["max_dtheta"] = float(dth.max()) if len(dth) else 0.0
    rep["max_ds_rel"] = float(ds.max()) if len(ds) else 0.0
    rep["n_bad"] = int(((dxy > TOL_XY) | (dth > TOL_THETA) | (ds > TOL_S_REL)).sum())
    return rep


def assert_subpix_parity(rep):
    assert rep["gpu_sorted"], rep
    assert rep["only_oracle"] == 0 and rep["only_gpu"] == 0, rep
    assert rep["n_bad"] == 0, rep
