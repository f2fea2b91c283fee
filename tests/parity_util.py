"""Comparison helpers shared by the parity tests.

Tolerances (FP64 mode): a CSR value may differ from the oracle's by at most
RTOL * (largest |K| of its row of the oracle); a residual entry by at most
RTOL * (largest |f| of its step).  Row/step scaling is used because entries that sum
element contributions of opposite sign cancel, and the element contractions round
differently (the engine contracts on MFMA tiles, the oracle in the reference's loop
order; both add the element contributions of a row in element index order).

Mixed mode (FP32-MFMA element contractions, FP64 material update and accumulation) is
held to RTOL_MIXED on the same scaled measures."""
import numpy as np

RTOL_FP64 = 1e-12
RTOL_MIXED = 1e-6


def row_scale(K, rp):
    n = len(rp) - 1
    out = np.zeros(n)
    absK = np.abs(K)
    nz = np.diff(rp) > 0
    if K.size:
        out[nz] = np.maximum.reduceat(absK, rp[:-1][nz])
    return out


def csr_err(Ke, Ko, rp):
    """max over entries of |Ke-Ko| / row scale (rows with zero scale use 1)."""
    if Ko.size == 0:
        return 0.0
    sc = row_scale(Ko, rp)
    sc[sc == 0] = 1.0
    per_entry_scale = np.repeat(sc, np.diff(rp))
    return float(np.max(np.abs(Ke - Ko) / per_entry_scale))


def vec_err(fe, fo):
    sc = np.max(np.abs(fo)) if fo.size else 1.0
    sc = sc if sc > 0 else 1.0
    return float(np.max(np.abs(fe - fo)) / sc) if fo.size else 0.0


def rel_norm_err(xe, xo):
    n = np.linalg.norm(xo)
    return float(np.linalg.norm(xe - xo) / (n if n > 0 else 1.0))
