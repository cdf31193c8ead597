"""Comparison of two renders of the same image: the accuracy report of the
mixed-precision form against the FP64 render.

pixel_errors   per-pixel relative Stokes-I error over the pixels the FP64
               render lights (I != 0), plus the pixels lit in only one of them
frequency_errors  per-ray relative error of the observed frequency (the
               redshifted emission frequency) over the rays both renders bring
               to the camera
spectrum_l1    relative L1 distance of two spectra
ring_diameter  diameter of the black-hole shadow: the 4-connected region of
               unlit pixels (I == 0) that holds the image centre, reported as
               the diameter of the circle of the same area, in M
"""
import numpy as np


def pixel_errors(I, I_ref):
    I = np.asarray(I, dtype=np.float64)
    I_ref = np.asarray(I_ref, dtype=np.float64)
    lit = I_ref != 0.0
    rel = np.abs(I[lit] - I_ref[lit]) / np.abs(I_ref[lit])
    out = {
        "lit_pixels": int(lit.sum()),
        "lit_only_one": int(np.sum(lit != (I != 0.0))),
        "median_rel_err": float(np.median(rel)) if rel.size else 0.0,
        "max_rel_err": float(rel.max()) if rel.size else 0.0,
        "frac_rel_err_gt_1e-4": float(np.mean(rel > 1e-4)) if rel.size else 0.0,
        "image_l1_rel_err": float(np.abs(I - I_ref).sum() / max(np.abs(I_ref).sum(), 1e-300)),
    }
    return out


def frequency_errors(freq, freq_ref, arrived, arrived_ref):
    """Per-ray relative error of the observed frequency (emitted frequency
    times the redshift factor E_obs / E_emit) over the rays that reach the
    camera in both renders (one ray per pixel: the per-pixel error)."""
    f = np.asarray(freq, dtype=np.float64)
    g = np.asarray(freq_ref, dtype=np.float64)
    both = np.asarray(arrived) & np.asarray(arrived_ref) & (g != 0.0)
    rel = np.abs(f[both] - g[both]) / np.abs(g[both])
    return {
        "rays": int(both.sum()),
        "arrived_only_one": int(np.sum(np.asarray(arrived) != np.asarray(arrived_ref))),
        "median_rel_err": float(np.median(rel)) if rel.size else 0.0,
        "max_rel_err": float(rel.max()) if rel.size else 0.0,
        "frac_rel_err_gt_1e-4": float(np.mean(rel > 1e-4)) if rel.size else 0.0,
    }


def spectrum_l1(spec, spec_ref):
    """sum |S - S_ref| / sum |S_ref| of two spectra on the same bins."""
    s = np.asarray(spec, dtype=np.float64)
    r = np.asarray(spec_ref, dtype=np.float64)
    return float(np.abs(s - r).sum() / max(np.abs(r).sum(), 1e-300))


def ring_diameter(I, width, height=None):
    """Equal-area diameter (in M) of the unlit region around the image centre;
    NaN when the centre pixel is lit."""
    from scipy import ndimage
    I = np.asarray(I, dtype=np.float64)
    ny, nx = I.shape
    height = width if height is None else height
    labels, _ = ndimage.label(I == 0.0)
    c = labels[ny // 2, nx // 2]
    if c == 0:
        return float("nan")
    area = np.count_nonzero(labels == c) * (width / nx) * (height / ny)
    return float(2.0 * np.sqrt(area / np.pi))
