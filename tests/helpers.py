"""Comparison helpers shared by the parity tests."""
import numpy as np


def rel_norm_err(a, ref):
    """max_i |a_i - ref_i| / max_i |ref_i| (normwise relative error)."""
    a = np.asarray(a, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    scale = np.max(np.abs(ref))
    if scale == 0.0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - ref)) / scale)


def disp_err(gpu_disp, ref_disp):
    """Nodal displacement error: max_i |u_i - u_i^ref| / max_i |u_i^ref| over
    node vectors."""
    d = np.linalg.norm(np.asarray(gpu_disp) - np.asarray(ref_disp), axis=0)
    s = np.linalg.norm(np.asarray(ref_disp), axis=0)
    return float(d.max() / s.max())


def ann_network(seed=0, h1=7, h2=5, constant=None):
    """Arrays of an ANN2SigLaw .npz file (ANN2SigLaw.C:137-158 reads them):
    a network that hardens with plastic strain and rate and softens with
    temperature over a copper-like range, or, with constant = c, one whose
    output is the constant c (w3 = 0, b3 = c)."""
    rng = np.random.default_rng(seed)
    w1 = rng.normal(size=(3, h1))
    w1[0] = 3.0 * np.abs(w1[0])
    w1[1] = 0.3 * np.abs(w1[1])
    w1[2] = -np.abs(w1[2])
    net = {
        "w1": w1,
        "b1": 0.5 * rng.normal(size=(h1, 1)),
        "w2": rng.uniform(-0.3, 1.0, size=(h1, h2)),
        "b2": 0.5 * rng.normal(size=(h2, 1)) - 1.0,
        "w3": rng.uniform(0.0, 0.25, size=(h2, 1)),
        "b3": np.array([[0.02]]),
        "logBase": np.array([1.0]),
        "minEntries": np.array([0.0, 0.0, 20.0, 90e6]),
        "maxEntries": np.array([2.0, 14.0, 1083.0, 600e6]),
    }
    if constant is not None:
        net["w3"] = np.zeros((h2, 1))
        net["b3"] = np.array([[float(constant)]])
    return net


def ann_forward(net, epsp, depsp, T):
    """The network's yield stress evaluated the way the file stores it:
    row-vector inputs times the stored (inputs x neurons) matrices."""
    mn, mx = net["minEntries"], net["maxEntries"]
    rg = mx - mn
    base = net["logBase"][0]
    x = np.array([(epsp - mn[0]) / rg[0],
                  (np.log(depsp / base) - mn[1]) / rg[1] if depsp > base else 0.0,
                  (T - mn[2]) / rg[2]])
    s1 = 1.0 / (1.0 + np.exp(-(x @ net["w1"] + net["b1"][:, 0])))
    s2 = 1.0 / (1.0 + np.exp(-(s1 @ net["w2"] + net["b2"][:, 0])))
    y = s2 @ net["w3"][:, 0] + net["b3"][0, 0]
    return rg[3] * y + mn[3]


def ann_law(path, net):
    """An ANN2SigLaw read from `net` written as an .npz file at `path`."""
    import dnlPython as dnl
    np.savez(str(path), **net)
    law = dnl.ANN2SigLaw("ann")
    law.setParameters(str(path))
    return law
