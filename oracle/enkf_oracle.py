"""numpy restatement of the EnKF analysis (EnkfAnalyze::pCalGainMatrix and
pAnalyze, enkfanalyze.cpp:113-183,343-495,578-597; clipPosition,
Matrix.cpp:170-185) in FP64 with LAPACK's SVD.

TEST INFRASTRUCTURE ONLY (what the GPU analysis is checked against).  The
data-row order is [nt][ng] here; the reference stores [ng][nt], a row
permutation that leaves every product unchanged."""
import numpy as np


def clip_position(s):
    rel = float(np.float32(0.001))
    n = len(s)
    acc = 0.0
    pos = 0
    while pos < n - 1:
        acc += abs(s[pos])
        if rel / (1 - rel) * acc > (n - pos + 1) * s[pos + 1]:
            break
        pos += 1
    return pos + 1


def vel_recover(v, dx, dt):
    dx, dt = np.float32(dx), np.float32(dt)
    t = (dx * dx) / (dt * dt * v)
    return np.sqrt(t, dtype=np.float32)


def vel_trans(v, dx, dt):
    dx, dt = np.float32(dx), np.float32(dt)
    return (dx * dx) / (dt * dt * v * v)


def member_mean(members):
    """pCreateAMean: FP32 sum in member order, then / N."""
    acc = np.zeros(members.shape[1:], np.float32)
    for m in members:
        acc = (acc + m).astype(np.float32)
    return (acc / np.float32(len(members))).astype(np.float32)


def sigma_of(hap, sigfac):
    """initPerturbSigma(max |HA'| by float-abs comparison, first occurrence)."""
    flat = hap.ravel()
    a = np.abs(flat.astype(np.float32))
    k = int(np.argmax(a))
    return abs(flat[k]) * float(np.float32(sigfac))


def analyze(members, ha, d, E, dx, dt, sigma=None, sigfac=None):
    """members [N][nx][nz] f32, ha [N][nt][ng] f32 (H on each member, no
    mute), d [nt][ng] f32 encoded observations, E [N][nt][ng] f64
    perturbation.  Returns (updated members, dict of intermediates)."""
    N = members.shape[0]
    HA = ha.reshape(N, -1).astype(np.float64).T          # nd x N
    Ep = E.reshape(N, -1).astype(np.float64).T           # nd x N
    D = d.reshape(-1).astype(np.float64)[:, None] + Ep
    HAp = HA - HA.sum(axis=1, keepdims=True) / N
    if sigma is None and sigfac is not None:
        sigma = sigma_of(HAp.T, sigfac)
    gamma = Ep - Ep.sum(axis=1, keepdims=True) / N
    band = HAp + gamma
    U, S, _ = np.linalg.svd(band, full_matrices=False)
    clip = clip_position(S)
    Uk = U[:, :clip]
    ssq = (1.0 / S[:clip]) * (1.0 / S[:clip])
    t4 = (HAp.T @ Uk) @ np.diag(ssq) @ (Uk.T @ (D - HA))
    A = members.reshape(N, -1)
    mean = member_mean(A)
    Ap = A.astype(np.float64) - mean.astype(np.float64)[None, :]
    t5 = Ap.T @ t4                                       # n x N
    out = np.empty_like(A)
    for i in range(N):
        v = vel_recover(A[i], dx, dt)
        v = (v.astype(np.float64) + t5[:, i]).astype(np.float32)
        out[i] = vel_trans(v, dx, dt)
    return out.reshape(members.shape), {"svals": S, "clip": clip, "t4": t4, "sigma": sigma}
