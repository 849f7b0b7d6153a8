"""Writes the committed fixtures of tests/golden.

mnist500.npz: the reference's own MATLAB data files converted to numpy
(code/matlab/complete_matlab/trainX.mat -> X [500x784], initial2.mat -> Y0
[500x2], labls.mat -> labels).  Read with scipy.io.loadmat (data only).

mnist500_oracle.npz: outputs of the FP64 oracle on that input (k=90, u=30):
kNN ids/distances, the perplexity-equalised P, P_sym, the per-iteration terms
at Y0 and at a spread Y (with the reference's algebraic Z and with the direct
Z), and 100 iterations of the loop (direct Z).  These are
regression vectors for the GPU parity tests; the oracle itself is pinned by
tests/test_oracle.py.

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF = "/root/reference/code/matlab/complete_matlab"


def convert_mat():
    import scipy.io as sio

    X = sio.loadmat(os.path.join(REF, "trainX.mat"))["train_X"].astype(np.float32)
    Y0 = sio.loadmat(os.path.join(REF, "initial2.mat"))["ydata"].astype(np.float32)
    lab = sio.loadmat(os.path.join(REF, "labls.mat"))["train_labels"].reshape(-1).astype(np.int32)
    np.savez_compressed(os.path.join(HERE, "mnist500.npz"), X=X, Y0=Y0, labels=lab)


def oracle_vectors():
    from oracle import oracle as O

    f = np.load(os.path.join(HERE, "mnist500.npz"))
    X, Y0 = f["X"], f["Y0"]
    k, u = 90, 30.0
    I, D = O.knn_exact(X, k)
    P = O.perplexity_equalize(I, D, u)
    Ps = O.symmetrize(P)
    out = dict(I=I, D=D, P_rowptr=P.rowptr, P_col=P.col, P_val=P.val,
               S_rowptr=Ps.rowptr, S_col=Ps.col, S_val=Ps.val)
    for tag, Y in (("y0", Y0.astype(np.float64)), ("ys", Y0.astype(np.float64) * 2e4)):
        Fa = O.attractive(Ps, Y)
        Fr, Z, ng = O.repulsive_interp(Y.copy(), 0.7, O.ZETA_ALGEBRAIC)
        Frd, Zd, _ = O.repulsive_interp(Y.copy(), 0.7, O.ZETA_DIRECT)
        out[f"{tag}_Frep_direct"] = Frd
        out[f"{tag}_Z_direct"] = np.float64(Zd)
        out[f"{tag}_Y"] = Y.astype(np.float32)
        out[f"{tag}_Fattr"] = Fa
        out[f"{tag}_Frep"] = Fr
        out[f"{tag}_Z"] = np.float64(Z)
        out[f"{tag}_ng"] = np.int32(ng)
        out[f"{tag}_kl"] = np.float64(O.kl(Ps, Y, Z, 1.0))
    Y = Y0.astype(np.float64).copy()
    uY = np.zeros_like(Y)
    g = np.ones_like(Y)
    Zs, ngs = O.run(Ps, Y, uY, g, 0, 100, zmode=O.ZETA_DIRECT)
    out["run100_Y"] = Y
    out["run100_Z"] = Zs
    out["run100_ng"] = ngs
    np.savez_compressed(os.path.join(HERE, "mnist500_oracle.npz"), **out)


if __name__ == "__main__":
    if os.path.isdir(REF):
        convert_mat()
    oracle_vectors()
    print("fixtures written")
